// Fused AO evaluation: values and first/second derivatives of every basis
// function on a chunk of points, one lane per point.
//
// Replaces the reference AO device functions (src/evaluate_density.cu:25-411,
// src/evaluate_gradient.cu:18-1126, src/evaluate_laplacian.cu:15-421,
// src/evaluate_hessian.cu:17-1836).  Differences that matter on CDNA4:
//   * the shell table is a device array read at a wave-uniform index, so the
//     compiler issues scalar (SMEM) loads into SGPRs -- no constant-bank size
//     limit and no per-chunk re-upload;
//   * each primitive's exponential is evaluated once per (point, shell) and
//     reused by every component and derivative through the three radial sums
//     S0, S1, S2 (the reference re-evaluates normalisation sqrt()s and does a
//     global read-modify-write per primitive and component);
//   * normalisation is folded into the coefficients on the host;
//   * every output element is written exactly once (no += on uninitialised
//     memory), rows beyond M are written as zeros for the padded contraction;
//   * a shell is skipped for the whole wave when every lane has
//     a_min r^2 > screen (exp underflow region in parity mode).
//
// Output layout: ao[(d * Mp + i) * ld + p], d = derivative slot, i = basis
// function, p = point in the chunk (points contiguous -> coalesced stores and
// directly the B operand layout of the MFMA contraction).

#include "ct_internal.h"
#include "generated/ao_poly.h"

namespace ct {

namespace {

template <typename T, int MODE>
__global__ __launch_bounds__(256) void ao_eval_kernel(const ShellMeta* __restrict__ meta,
                                                      const ShellGeom* __restrict__ geom,
                                                      const PrimData* __restrict__ prim, int nshell, int nbasis,
                                                      int nbasis_pad, const double* __restrict__ pts, int64_t npts,
                                                      int64_t ld, T* __restrict__ ao, T screen) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // p < ld by construction
  const int64_t pc = p < npts ? p : (npts - 1);
  const double px = pts[3 * pc + 0], py = pts[3 * pc + 1], pz = pts[3 * pc + 2];
  const int64_t sd = (int64_t)nbasis_pad * ld;  // stride between derivative slots

  for (int s = 0; s < nshell; ++s) {
    const ShellMeta m = meta[s];
    const ShellGeom g = geom[s];
    const T x = (T)(px - g.cx), y = (T)(py - g.cy), z = (T)(pz - g.cz);
    const T r2 = x * x + y * y + z * z;
    T s0 = 0, s1 = 0, s2 = 0;
    const bool far = (T)g.amin * r2 > screen;
    if (!__all(far)) {
      for (int k = 0; k < m.nprim; ++k) {
        const PrimData pd = prim[m.prim_off + k];
        const T a = (T)pd.a;
        const T e = (T)pd.c * exp(-a * r2);
        s0 += e;
        if constexpr (MODE != MODE_RHO) {
          const T m2a = T(-2) * a;
          s1 += m2a * e;
          if constexpr (MODE == MODE_LAP || MODE == MODE_HESS) s2 += m2a * m2a * e;
        }
      }
    }
    T* o = ao + (int64_t)m.first * ld + p;
    if constexpr (MODE == MODE_RHO) ao_shell_dispatch_val<T>(m.code, x, y, z, s0, s1, s2, o, ld, sd);
    if constexpr (MODE == MODE_GRAD) ao_shell_dispatch_grad<T>(m.code, x, y, z, s0, s1, s2, o, ld, sd);
    if constexpr (MODE == MODE_LAP) ao_shell_dispatch_lap<T>(m.code, x, y, z, s0, s1, s2, o, ld, sd);
    if constexpr (MODE == MODE_HESS) ao_shell_dispatch_hess<T>(m.code, x, y, z, s0, s1, s2, o, ld, sd);
  }
  // zero rows of the padded basis so the contraction never reads garbage
  constexpr int nder = mode_nder(MODE);
  for (int d = 0; d < nder; ++d)
    for (int i = nbasis; i < nbasis_pad; ++i) ao[((int64_t)d * nbasis_pad + i) * ld + p] = T(0);
}

}  // namespace

void launch_ao_f64(int mode, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                   int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, double* ao,
                   double screen, hipStream_t s) {
  const dim3 block(256), grid((unsigned)(ld / 256));
  switch (mode) {
    case MODE_RHO:
      hipLaunchKernelGGL((ao_eval_kernel<double, MODE_RHO>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
    case MODE_GRAD:
      hipLaunchKernelGGL((ao_eval_kernel<double, MODE_GRAD>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
    case MODE_LAP:
      hipLaunchKernelGGL((ao_eval_kernel<double, MODE_LAP>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
    default:
      hipLaunchKernelGGL((ao_eval_kernel<double, MODE_HESS>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
  }
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
