// extern "C" boundary (include/chemtools_amd.h).  Every entry point converts
// exceptions into a status code plus a thread-local message, so errors reach
// Python as readable RuntimeErrors instead of the reference's raw
// cudaError_t throw (include/cuda_utils.cuh:55-60, seen in Python as
// "Caught an unknown exception!").

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "esp.h"
#include "mol.h"

struct ct_mol {
  ct::Mol* impl;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return CT_OK;
  } catch (const ct::Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return CT_E_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CT_E_INVALID;
  }
}

ct::HostBasis to_host_basis(const ct_basis_desc* b) {
  CT_REQUIRE(b != nullptr && b->nshell > 0 && b->shell_atom && b->shell_l && b->shell_nprim && b->shell_prim_off &&
                 b->exps && b->coefs,
             "incomplete basis description");
  ct::HostBasis h;
  h.shell_atom.assign(b->shell_atom, b->shell_atom + b->nshell);
  h.shell_l.assign(b->shell_l, b->shell_l + b->nshell);
  h.shell_nprim.assign(b->shell_nprim, b->shell_nprim + b->nshell);
  h.shell_prim_off.assign(b->shell_prim_off, b->shell_prim_off + b->nshell);
  size_t nprim = 0;
  for (int s = 0; s < b->nshell; ++s) {
    CT_REQUIRE(b->shell_nprim[s] > 0 && b->shell_prim_off[s] >= 0, "bad primitive counts");
    nprim = std::max(nprim, (size_t)b->shell_prim_off[s] + b->shell_nprim[s]);
  }
  h.exps.assign(b->exps, b->exps + nprim);
  h.coefs.assign(b->coefs, b->coefs + nprim);
  int M = 0;
  for (int l : h.shell_l) M += ct::shell_nfunc(l);
  h.nbasis = M;
  return h;
}

ct::Mol& M(ct_mol* m) {
  CT_REQUIRE(m != nullptr && m->impl != nullptr, "null molecule handle");
  return *m->impl;
}
const ct::Mol& M(const ct_mol* m) {
  CT_REQUIRE(m != nullptr && m->impl != nullptr, "null molecule handle");
  return *m->impl;
}

}  // namespace

extern "C" {

int ct_mol_create(const ct_basis_desc* basis, const double* atom_xyz, const double* charges, int32_t natom,
                  const double* P, int32_t nbasis, ct_mol** out) {
  return guard([&] {
    CT_REQUIRE(out != nullptr && atom_xyz != nullptr && P != nullptr, "null argument");
    *out = nullptr;
    ct::HostBasis hb = to_host_basis(basis);
    CT_REQUIRE(hb.nbasis == nbasis, "nbasis does not match the shell table (" + std::to_string(hb.nbasis) + ")");
    auto* h = new ct_mol{nullptr};
    try {
      h->impl = new ct::Mol(hb, atom_xyz, charges, nullptr, natom, P);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

namespace {

int create_from(ct::FchkData (*load)(const std::string&), const char* path, ct_mol** out) {
  return guard([&] {
    CT_REQUIRE(out != nullptr && path != nullptr, "null argument");
    *out = nullptr;
    ct::FchkData d = load(path);
    auto* h = new ct_mol{nullptr};
    try {
      h->impl = new ct::Mol(d.basis, d.atom_xyz.data(), d.charges.data(), d.numbers.data(),
                            (int)d.numbers.size(), d.P.data());
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

}  // namespace

int ct_mol_create_from_fchk(const char* path, ct_mol** out) { return create_from(ct::load_fchk, path, out); }

int ct_mol_create_from_file(const char* path, ct_mol** out) { return create_from(ct::load_wavefunction, path, out); }

struct ct_wavefunction {
  ct::FchkData d;
};

int ct_wavefunction_read(const char* path, ct_wavefunction** out) {
  return guard([&] {
    CT_REQUIRE(out != nullptr && path != nullptr, "null argument");
    *out = nullptr;
    auto* w = new ct_wavefunction{ct::load_wavefunction(path)};
    *out = w;
  });
}

int ct_wavefunction_destroy(ct_wavefunction* wfn) {
  return guard([&] { delete wfn; });
}

int ct_wavefunction_sizes(const ct_wavefunction* wfn, int32_t* nshell, int32_t* nprim, int32_t* nbasis,
                          int32_t* natom) {
  return guard([&] {
    CT_REQUIRE(wfn != nullptr, "null wavefunction");
    const auto& d = wfn->d;
    if (nshell) *nshell = (int32_t)d.basis.shell_l.size();
    if (nprim) *nprim = (int32_t)d.basis.exps.size();
    if (nbasis) *nbasis = (int32_t)d.basis.nbasis;
    if (natom) *natom = (int32_t)d.numbers.size();
  });
}

int ct_wavefunction_copy(const ct_wavefunction* wfn, int32_t* shell_atom, int32_t* shell_l, int32_t* shell_nprim,
                         int32_t* shell_prim_off, double* exps, double* coefs, double* P, double* atom_xyz,
                         double* charges, int32_t* numbers) {
  return guard([&] {
    CT_REQUIRE(wfn != nullptr, "null wavefunction");
    const auto& d = wfn->d;
    const auto& b = d.basis;
    auto put = [](auto* dst, const auto& src) {
      if (dst) std::copy(src.begin(), src.end(), dst);
    };
    put(shell_atom, b.shell_atom);
    put(shell_l, b.shell_l);
    put(shell_nprim, b.shell_nprim);
    put(shell_prim_off, b.shell_prim_off);
    put(exps, b.exps);
    put(coefs, b.coefs);
    put(P, d.P);
    put(atom_xyz, d.atom_xyz);
    put(charges, d.charges);
    put(numbers, d.numbers);
  });
}

int ct_mol_destroy(ct_mol* mol) {
  return guard([&] {
    if (!mol) return;
    delete mol->impl;
    delete mol;
  });
}

int ct_mol_nbasis(const ct_mol* mol) { return mol && mol->impl ? mol->impl->nbasis() : CT_E_INVALID; }
int ct_mol_natom(const ct_mol* mol) { return mol && mol->impl ? mol->impl->natom() : CT_E_INVALID; }

int ct_mol_get_coordinates(const ct_mol* mol, double* xyz) {
  return guard([&] {
    CT_REQUIRE(xyz, "null output");
    const auto& v = M(mol).coordinates();
    std::memcpy(xyz, v.data(), sizeof(double) * v.size());
  });
}

int ct_mol_get_numbers(const ct_mol* mol, int32_t* numbers) {
  return guard([&] {
    CT_REQUIRE(numbers, "null output");
    const auto& v = M(mol).numbers();
    std::memcpy(numbers, v.data(), sizeof(int32_t) * v.size());
  });
}

int ct_mol_get_charges(const ct_mol* mol, double* charges) {
  return guard([&] {
    CT_REQUIRE(charges, "null output");
    const auto& v = M(mol).charges();
    std::memcpy(charges, v.data(), sizeof(double) * v.size());
  });
}

int ct_mol_get_density_matrix(const ct_mol* mol, double* P) {
  return guard([&] {
    CT_REQUIRE(P, "null output");
    const auto& v = M(mol).density_matrix();
    std::memcpy(P, v.data(), sizeof(double) * v.size());
  });
}

int ct_mol_set_coordinates(ct_mol* mol, const double* atom_xyz) {
  return guard([&] { M(mol).set_coordinates(atom_xyz); });
}

int ct_mol_set_density_matrix(ct_mol* mol, const double* P) {
  return guard([&] { M(mol).set_density_matrix(P); });
}

int ct_mol_set_workspace_limit(ct_mol* mol, int64_t bytes) {
  return guard([&] {
    CT_REQUIRE(bytes > 0, "workspace limit must be positive");
    M(mol).set_workspace_limit(bytes);
  });
}

int ct_set_device_allocation_limit(int64_t bytes) {
  return guard([&] {
    CT_REQUIRE(bytes >= -1, "allocation limit must be -1 (none) or a byte count");
    ct::set_device_allocation_limit(bytes);
  });
}
int ct_device_allocation_held(int64_t* bytes) {
  return guard([&] {
    CT_REQUIRE(bytes, "null output");
    *bytes = ct::device_allocation_held();
  });
}
int ct_measure_peaks(double* bf16_mfma_tflops, double* f64_mfma_tflops, double* f64_valu_tflops, double* hbm_gbs) {
  return guard([&] {
    CT_REQUIRE(bf16_mfma_tflops && f64_mfma_tflops && f64_valu_tflops && hbm_gbs, "null output");
    ct::measure_peaks(bf16_mfma_tflops, f64_mfma_tflops, f64_valu_tflops, hbm_gbs);
  });
}
int ct_mol_set_screening(ct_mol* mol, double eps) {
  return guard([&] {
    CT_REQUIRE(eps >= 0.0 && std::isfinite(eps), "screening threshold must be finite and non-negative");
    M(mol).set_screening(eps);
  });
}

int ct_shell_screening_radius(int32_t l, int32_t nprim, const double* exps, const double* coefs, double eps,
                              double* radius) {
  return guard([&] {
    CT_REQUIRE(exps && coefs && radius, "null pointer");
    CT_REQUIRE(nprim > 0, "shell without primitives");
    (void)ct::shell_nfunc(l);  // rejects unsupported l
    std::vector<double> cn(nprim);
    for (int p = 0; p < nprim; ++p) {
      CT_REQUIRE(exps[p] > 0.0 && std::isfinite(exps[p]), "exponents must be positive and finite");
      cn[p] = coefs[p] * ct::prim_radial_norm(exps[p], std::abs(l));
    }
    *radius = ct::shell_screen_radius(l, nprim, exps, cn.data(), eps);
  });
}

int ct_eval(ct_mol* mol, const double* points, int64_t n, int32_t layout, uint32_t quantities, int32_t precision,
            const ct_eval_params* params, const ct_outputs* out, ct_eval_report* report) {
  return guard([&] {
    CT_REQUIRE(out != nullptr, "null outputs");
    ct_eval_params p{};
    if (params) p = *params;
    if (p.mixed_tol <= 0) p.mixed_tol = 1e-6;
    if (p.rho_floor <= 0) p.rho_floor = 1e-10;
    M(mol).eval(points, n, layout, quantities, precision, p, *out, report);
  });
}

int ct_eval_noalloc(ct_mol* mol, const double* points, int64_t n, int32_t layout, uint32_t quantities,
                    int32_t precision, const ct_eval_params* params, const ct_outputs* out, ct_eval_report* report) {
  return guard([&] {
    CT_REQUIRE(out != nullptr, "null outputs");
    ct_eval_params p{};
    if (params) p = *params;
    if (p.mixed_tol <= 0) p.mixed_tol = 1e-6;
    if (p.rho_floor <= 0) p.rho_floor = 1e-10;
    ct::Mol& m = M(mol);
    const ct::Workspace w{static_cast<char*>(ct::noalloc_scratch(m.device())), (size_t)CT_NOALLOC_SCRATCH_BYTES};
    m.eval(points, n, layout, quantities, precision, p, *out, report, nullptr, &w);
  });
}
int ct_eval_cubic(ct_mol* mol, const double* origin, const double* axes, const int64_t* shape, uint32_t quantities,
                  int32_t precision, const ct_eval_params* params, const ct_outputs* out, ct_eval_report* report) {
  return guard([&] {
    CT_REQUIRE(out != nullptr && origin != nullptr && axes != nullptr && shape != nullptr, "null argument");
    ct::CubicGrid g;
    for (int k = 0; k < 3; ++k) {
      CT_REQUIRE(shape[k] > 0, "grid shape must be positive");
      g.origin[k] = origin[k];
      g.shape[k] = shape[k];
    }
    for (int k = 0; k < 9; ++k) g.axes[k] = axes[k];
    CT_REQUIRE(shape[0] <= (int64_t(1) << 62) / shape[1] / shape[2], "grid too large");
    ct_eval_params p{};
    if (params) p = *params;
    if (p.mixed_tol <= 0) p.mixed_tol = 1e-6;
    if (p.rho_floor <= 0) p.rho_floor = 1e-10;
    M(mol).eval(nullptr, shape[0] * shape[1] * shape[2], CT_LAYOUT_ROW, quantities, precision, p, *out, report, &g);
  });
}

int ct_norm_rows(const double* v, int64_t n, double* out) {
  return guard([&] {
    CT_REQUIRE(n >= 0, "negative size");
    if (n == 0) return;
    CT_REQUIRE(v && out, "null pointer");
    const bool vd = ct::is_device_pointer(v), od = ct::is_device_pointer(out);
    ct::DeviceBuffer tmp;
    if ((!vd || !od) && !tmp.reserve(sizeof(double) * 4 * (size_t)n))
      throw ct::Error(CT_E_NOMEM, "cannot allocate");
    const double* dv = v;
    double* dout = out;
    if (!vd) {
      CT_HIP_CHECK(hipMemcpy(tmp.get(), v, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
      dv = static_cast<const double*>(tmp.get());
    }
    if (!od) dout = static_cast<double*>(tmp.get()) + 3 * n;
    ct::launch_norm_rows(dv, n, dout, nullptr);
    if (!od) CT_HIP_CHECK(hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost));
    CT_HIP_CHECK(hipDeviceSynchronize());
  });
}

int ct_esp(ct_mol* mol, const double* points, int64_t n, int32_t layout, int32_t precision, double* out) {
  return guard([&] { M(mol).esp(points, n, layout, precision, out); });
}
int ct_mol_esp_pairs(ct_mol* mol, int64_t* pairs_by_L) {
  return guard([&] {
    CT_REQUIRE(pairs_by_L, "null output");
    M(mol).esp_pairs_by_L(pairs_by_L);
  });
}
int ct_boys(const double* T, int64_t n, int32_t m, int32_t lmax_pairs, double* F) {
  return guard([&] {
    CT_REQUIRE(n >= 0, "negative size");
    CT_REQUIRE(m >= 0 && m <= lmax_pairs && lmax_pairs <= ct::kEspLmax, "need 0 <= m <= lmax_pairs <= 8");
    if (n == 0) return;
    CT_REQUIRE(T && F, "null pointer");
    const bool td = ct::is_device_pointer(T), fd = ct::is_device_pointer(F);
    const std::vector<double> tab = ct::build_boys_tables();
    const size_t nf = (size_t)n * (m + 1);
    ct::DeviceBuffer buf;
    if (!buf.reserve(sizeof(double) * (tab.size() + (td ? 0 : n) + (fd ? 0 : nf))))
      throw ct::Error(CT_E_NOMEM, "cannot allocate");
    double* dtab = static_cast<double*>(buf.get());
    double* cur = dtab + tab.size();
    CT_HIP_CHECK(hipMemcpy(dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    const double* dT = T;
    if (!td) {
      CT_HIP_CHECK(hipMemcpy(cur, T, sizeof(double) * n, hipMemcpyHostToDevice));
      dT = cur;
      cur += n;
    }
    double* dF = fd ? F : cur;
    ct::launch_boys_probe(dT, n, m, lmax_pairs, dtab, dF, nullptr);
    if (!fd) CT_HIP_CHECK(hipMemcpy(F, dF, sizeof(double) * nf, hipMemcpyDeviceToHost));
    CT_HIP_CHECK(hipDeviceSynchronize());
  });
}

int ct_backward(ct_mol* mol, const double* points, int64_t n, int32_t layout, const double* dL_drho,
                int32_t precision, double* dL_dP, double* dL_dR) {
  return guard([&] { M(mol).backward(points, n, layout, dL_drho, precision, dL_dP, dL_dR); });
}

int ct_mol_synchronize(ct_mol* mol) {
  return guard([&] { M(mol).synchronize(); });
}

int ct_mol_enable_profiling(ct_mol* mol, int enable) {
  return guard([&] {
    M(mol).synchronize();
    M(mol).timer().collect();
    M(mol).timer().enable(enable != 0);
  });
}

int ct_mol_kernel_stats(ct_mol* mol, ct_kernel_stats* out) {
  return guard([&] {
    CT_REQUIRE(out, "null output");
    M(mol).synchronize();
    M(mol).timer().collect();
    *out = M(mol).timer().stats();
  });
}

const char* ct_last_error(void) { return g_last_error.c_str(); }

const char* ct_version(void) { return "chemtools-amd 0.1 (gfx950)"; }

}  // extern "C"
