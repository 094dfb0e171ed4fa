"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ct_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("ct_mol_create", "ct_mol_create_from_fchk", "ct_mol_destroy", "ct_eval", "ct_esp",
                 "ct_backward", "ct_last_error", "ct_norm_rows"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    import chemtoolscuda_amd

    lib = ctypes.CDLL(chemtoolscuda_amd.LIBRARY_PATH)
    assert hasattr(lib, name), name


def test_error_path_without_gpu_work():
    import chemtoolscuda_amd

    lib = ctypes.CDLL(chemtoolscuda_amd.LIBRARY_PATH)
    lib.ct_last_error.restype = ctypes.c_char_p
    lib.ct_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.ct_version()
    out = ctypes.c_void_p()
    # null basis -> CT_E_INVALID with a message, no device work attempted
    st = lib.ct_mol_create(None, None, None, 0, None, 0, ctypes.byref(out))
    assert st == -1
    assert len(lib.ct_last_error()) > 0
    st = lib.ct_mol_create_from_fchk(b"/nonexistent.fchk", ctypes.byref(out))
    assert st == -4
    assert b"cannot open" in lib.ct_last_error()
