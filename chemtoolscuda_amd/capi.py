"""ctypes binding of the C ABI (include/chemtools_amd.h).

Used where device pointers cross the boundary (torch tensors' data_ptr() in
the autograd wrapper and the benchmark); the pybind11 ``Molecule`` covers the
host-array API.  Raises on any non-zero status with the library's message.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchemtools_amd.so")

Q = {
    "rho": 1 << 0, "grad": 1 << 1, "hess": 1 << 2, "lap": 1 << 3, "tau_pd": 1 << 4, "rdg": 1 << 5,
    "shannon": 1 << 6, "weizsacker": 1 << 7, "thomas_fermi": 1 << 8, "gge": 1 << 9, "general_ked": 1 << 10,
    "ked_alpha": 1 << 11,
}
WIDTH = {"grad": 3, "hess": 9}
PREC = {"fp64": 0, "mixed": 1}
LAYOUT = {"row": 0, "col": 1}


class BasisDesc(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int32), ("shell_atom", ctypes.c_void_p), ("shell_l", ctypes.c_void_p),
                ("shell_nprim", ctypes.c_void_p), ("shell_prim_off", ctypes.c_void_p), ("exps", ctypes.c_void_p),
                ("coefs", ctypes.c_void_p)]


class EvalParams(ctypes.Structure):
    _fields_ = [("gge_a", ctypes.c_double), ("gge_b", ctypes.c_double), ("ked_a", ctypes.c_double),
                ("ked_alpha", ctypes.c_double), ("mixed_tol", ctypes.c_double), ("rho_floor", ctypes.c_double),
                ("route", ctypes.c_int32), ("out_layout", ctypes.c_int32), ("rho_target", ctypes.c_void_p)]


class Outputs(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in Q]


class EvalReport(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("n_reevaluated", ctypes.c_int64), ("n_chunks", ctypes.c_int64),
                ("max_error_estimate", ctypes.c_double), ("factor_rank", ctypes.c_int64),
                ("active_fraction", ctypes.c_double)]


class KernelStats(ctypes.Structure):
    _fields_ = [(f"{k}_ms", ctypes.c_double) for k in ("ao", "contract", "finalize", "esp", "backward", "reeval")] + \
               [(f"{k}_calls", ctypes.c_int64) for k in ("ao", "contract", "finalize", "esp", "backward", "reeval")] + \
               [("screen_ms", ctypes.c_double), ("screen_calls", ctypes.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        try:  # share PyTorch's HIP runtime when it is installed (see __init__.py)
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: build with `python chemtoolscuda_amd/build_ext.py`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.ct_mol_create.argtypes = [ctypes.POINTER(BasisDesc), vp, vp, i32, vp, i32, ctypes.POINTER(vp)]
        L.ct_mol_create_from_fchk.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.ct_mol_create_from_file.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.ct_wavefunction_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.ct_wavefunction_sizes.argtypes = [vp, vp, vp, vp, vp]
        L.ct_wavefunction_copy.argtypes = [vp] * 11
        L.ct_wavefunction_destroy.argtypes = [vp]
        L.ct_mol_destroy.argtypes = [vp]
        L.ct_eval.argtypes = [vp, vp, i64, i32, ctypes.c_uint32, i32, ctypes.POINTER(EvalParams),
                              ctypes.POINTER(Outputs), ctypes.POINTER(EvalReport)]
        L.ct_esp.argtypes = [vp, vp, i64, i32, i32, vp]
        L.ct_eval_noalloc.argtypes = [vp, vp, i64, i32, ctypes.c_uint32, i32, ctypes.POINTER(EvalParams),
                                      ctypes.POINTER(Outputs), ctypes.POINTER(EvalReport)]
        L.ct_boys.argtypes = [vp, i64, i32, i32, vp]
        L.ct_mol_esp_pairs.argtypes = [vp, vp]
        L.ct_backward.argtypes = [vp, vp, i64, i32, vp, i32, vp, vp]
        L.ct_mol_set_coordinates.argtypes = [vp, vp]
        L.ct_mol_set_density_matrix.argtypes = [vp, vp]
        L.ct_mol_set_workspace_limit.argtypes = [vp, i64]
        L.ct_mol_set_screening.argtypes = [vp, ctypes.c_double]
        L.ct_set_device_allocation_limit.argtypes = [i64]
        L.ct_device_allocation_held.argtypes = [ctypes.POINTER(i64)]
        L.ct_measure_peaks.argtypes = [ctypes.POINTER(ctypes.c_double)] * 4
        L.ct_shell_screening_radius.argtypes = [i32, i32, vp, vp, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.ct_mol_synchronize.argtypes = [vp]
        L.ct_mol_enable_profiling.argtypes = [vp, ctypes.c_int]
        L.ct_mol_kernel_stats.argtypes = [vp, ctypes.POINTER(KernelStats)]
        L.ct_mol_nbasis.argtypes = [vp]
        L.ct_mol_natom.argtypes = [vp]
        L.ct_last_error.restype = ctypes.c_char_p
        L.ct_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise RuntimeError(f"chemtools_amd status {status}: {lib().ct_last_error().decode(errors='replace')}")


class Handle:
    """Owns a ct_mol*; built from a flat basis (``BasisSet.flat()``) and host arrays."""

    def __init__(self, flat, P, charges):
        import numpy as np

        self._keep = {k: np.ascontiguousarray(flat[k], dtype=np.int32) for k in
                      ("shell_atom", "shell_l", "shell_nprim", "shell_prim_off")}
        self._keep["exps"] = np.ascontiguousarray(flat["exps"], dtype=np.float64)
        self._keep["coefs"] = np.ascontiguousarray(flat["coefs"], dtype=np.float64)
        xyz = np.ascontiguousarray(flat["atom_xyz"], dtype=np.float64)
        Z = np.ascontiguousarray(charges, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        k = self._keep
        desc = BasisDesc(len(k["shell_l"]), k["shell_atom"].ctypes.data, k["shell_l"].ctypes.data,
                         k["shell_nprim"].ctypes.data, k["shell_prim_off"].ctypes.data, k["exps"].ctypes.data,
                         k["coefs"].ctypes.data)
        h = ctypes.c_void_p()
        check(lib().ct_mol_create(ctypes.byref(desc), xyz.ctypes.data, Z.ctypes.data, xyz.shape[0], P.ctypes.data,
                                  P.shape[0], ctypes.byref(h)))
        self.ptr = h
        self.nbasis = P.shape[0]
        self.natom = xyz.shape[0]

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                lib().ct_mol_destroy(self.ptr)
            except TypeError:  # interpreter shutdown: module globals already cleared
                pass
            self.ptr = None

    def eval(self, points_ptr: int, n: int, outputs: dict, precision: str = "fp64", params: dict | None = None,
             layout: str = "row"):
        """outputs: {quantity: pointer}; pointers may be host or device.  layout:
        points (n, 3) "row" or (3, n) "col"; params["out_layout"] = 1 gives grad
        (3, n) and hess (9, n)."""
        q = 0
        o = Outputs()
        for name, ptr in outputs.items():
            q |= Q[name]
            setattr(o, name, ptr)
        par = EvalParams(**(params or {}))
        rep = EvalReport()
        check(lib().ct_eval(self.ptr, points_ptr, n, LAYOUT[layout], q, PREC[precision], ctypes.byref(par),
                            ctypes.byref(o), ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in EvalReport._fields_}

    def eval_noalloc(self, points_ptr: int, n: int, outputs: dict, precision: str = "fp64",
                     params: dict | None = None, layout: str = "row"):
        """ct_eval_noalloc: like eval(), device pointers only, no allocation."""
        q = 0
        o = Outputs()
        for name, ptr in outputs.items():
            q |= Q[name]
            setattr(o, name, ptr)
        par = EvalParams(**(params or {}))
        rep = EvalReport()
        check(lib().ct_eval_noalloc(self.ptr, points_ptr, n, LAYOUT[layout], q, PREC[precision], ctypes.byref(par),
                                    ctypes.byref(o), ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in EvalReport._fields_}

    def esp(self, points_ptr: int, n: int, out_ptr: int, precision: str = "fp64", layout: str = "row"):
        check(lib().ct_esp(self.ptr, points_ptr, n, LAYOUT[layout], PREC[precision], out_ptr))

    def esp_pairs_by_L(self):
        """Screened primitive pairs of the ESP pre-contraction by total L (0..8)."""
        import numpy as np

        out = np.zeros(9, dtype=np.int64)
        check(lib().ct_mol_esp_pairs(self.ptr, out.ctypes.data))
        return out

    def backward(self, points_ptr: int, n: int, w_ptr: int, dP_ptr, dR_ptr, precision: str = "fp64",
                 layout: str = "row"):
        check(lib().ct_backward(self.ptr, points_ptr, n, LAYOUT[layout], w_ptr, PREC[precision], dP_ptr, dR_ptr))

    def set_coordinates(self, ptr: int):
        check(lib().ct_mol_set_coordinates(self.ptr, ptr))

    def set_density_matrix(self, ptr: int):
        check(lib().ct_mol_set_density_matrix(self.ptr, ptr))

    def set_workspace_limit(self, nbytes: int):
        check(lib().ct_mol_set_workspace_limit(self.ptr, int(nbytes)))

    def set_screening(self, eps: float):
        check(lib().ct_mol_set_screening(self.ptr, float(eps)))

    def synchronize(self):
        check(lib().ct_mol_synchronize(self.ptr))

    def enable_profiling(self, on: bool = True):
        check(lib().ct_mol_enable_profiling(self.ptr, 1 if on else 0))

    def kernel_stats(self) -> dict:
        st = KernelStats()
        check(lib().ct_mol_kernel_stats(self.ptr, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in KernelStats._fields_}


def read_wavefunction(path: str) -> dict:
    """Reads a .fchk / .wfn / .wfx / Molden file with the library's native
    readers, on the host only (no device work): the shell table (the
    ``BasisSet.flat()`` layout), P, coordinates, nuclear charges and atomic
    numbers -- exactly what ``Molecule(path)`` uploads."""
    import numpy as np

    L = lib()
    h = ctypes.c_void_p()
    check(L.ct_wavefunction_read(os.fsencode(path), ctypes.byref(h)))
    try:
        n = [ctypes.c_int32() for _ in range(4)]
        check(L.ct_wavefunction_sizes(h, *[ctypes.byref(v) for v in n]))
        nshell, nprim, nbasis, natom = (v.value for v in n)
        out = {
            "shell_atom": np.zeros(nshell, np.int32), "shell_l": np.zeros(nshell, np.int32),
            "shell_nprim": np.zeros(nshell, np.int32), "shell_prim_off": np.zeros(nshell, np.int32),
            "exps": np.zeros(nprim), "coefs": np.zeros(nprim), "P": np.zeros((nbasis, nbasis)),
            "atom_xyz": np.zeros((natom, 3)), "charges": np.zeros(natom), "numbers": np.zeros(natom, np.int32),
        }
        keys = ["shell_atom", "shell_l", "shell_nprim", "shell_prim_off", "exps", "coefs", "P", "atom_xyz", "charges",
                "numbers"]
        check(L.ct_wavefunction_copy(h, *[out[k].ctypes.data for k in keys]))
    finally:
        L.ct_wavefunction_destroy(h)
    return out


def set_device_allocation_limit(nbytes: int) -> None:
    """Process-wide cap on the library's device buffers (-1: none); requests
    beyond it fail like a failed hipMalloc (the chunker halves and retries)."""
    check(lib().ct_set_device_allocation_limit(int(nbytes)))


def device_allocation_held() -> int:
    """Device memory the library's buffers hold now (all handles; bytes)."""
    v = ctypes.c_int64()
    check(lib().ct_device_allocation_held(ctypes.byref(v)))
    return v.value


def measure_peaks() -> dict:
    """ct_measure_peaks: bf16 and FP64 matrix-core TFLOP/s, FP64 vector FMA
    TFLOP/s and HBM copy GB/s (read + write) sustained by the current device."""
    v = [ctypes.c_double() for _ in range(4)]
    check(lib().ct_measure_peaks(*[ctypes.byref(x) for x in v]))
    return {"bf16_mfma_tflops": v[0].value, "f64_mfma_tflops": v[1].value, "f64_valu_tflops": v[2].value,
            "hbm_copy_gbs": v[3].value}


def boys(T, m: int, lmax_pairs: int):
    """F_0..F_m(T) of the device Boys function as the ESP kernels evaluate it
    for primitive pairs up to total angular momentum lmax_pairs: (len(T), m+1)."""
    import numpy as np

    t = np.ascontiguousarray(T, dtype=np.float64)
    F = np.zeros((t.size, m + 1))
    check(lib().ct_boys(t.ctypes.data, t.size, int(m), int(lmax_pairs), F.ctypes.data))
    return F


def shell_screening_radius(l: int, exps, coefs, eps: float) -> float:
    """Cutoff radius of one shell (host only): beyond it every function of the
    shell and its first and second derivatives are at most eps in magnitude."""
    import numpy as np

    a = np.ascontiguousarray(exps, dtype=np.float64)
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    r = ctypes.c_double()
    check(lib().ct_shell_screening_radius(int(l), a.size, a.ctypes.data, c.ctypes.data, float(eps), ctypes.byref(r)))
    return r.value
