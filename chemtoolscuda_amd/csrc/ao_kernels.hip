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

// exp(-a r^2) with the argument formed in FP64.  For the FP32 path only the
// reduced argument goes through the FP32 exponential: with t = a r^2 / ln 2,
// n = rint(t), exp(-a r^2) = 2^-(t - n) * 2^-n, so the FP32 rounding of the
// argument (|t - n| <= 1/2) costs at most ~2^-25 relative instead of
// |a r^2| * 2^-24 (which reaches 5e-6 near the exp underflow limit).
template <typename T>
__device__ __forceinline__ T gauss_exp(double a, double r2d);
template <>
__device__ __forceinline__ double gauss_exp<double>(double a, double r2d) {
  return exp(-a * r2d);
}
template <>
__device__ __forceinline__ float gauss_exp<float>(double a, double r2d) {
  const double t = a * r2d * 1.4426950408889634074;
  const double n = rint(t);
  const float f = (float)(t - n);
  return ldexpf(exp2f(-f), -(int)n);
}

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
    const double dx = px - g.cx, dy = py - g.cy, dz = pz - g.cz;
    const double r2d = dx * dx + dy * dy + dz * dz;
    const T x = (T)dx, y = (T)dy, z = (T)dz;
    T s0 = 0, s1 = 0, s2 = 0;
    const bool far = g.amin * r2d > (double)screen;
    if (!__all(far)) {
      for (int k = 0; k < m.nprim; ++k) {
        const PrimData pd = prim[m.prim_off + k];
        const T a = (T)pd.a;
        const T e = (T)pd.c * gauss_exp<T>(pd.a, r2d);
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

void launch_ao_f32(int mode, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                   int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, float* ao, float screen,
                   hipStream_t s) {
  const dim3 block(256), grid((unsigned)(ld / 256));
  switch (mode) {
    case MODE_RHO:
      hipLaunchKernelGGL((ao_eval_kernel<float, MODE_RHO>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
    case MODE_GRAD:
      hipLaunchKernelGGL((ao_eval_kernel<float, MODE_GRAD>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
    case MODE_LAP:
      hipLaunchKernelGGL((ao_eval_kernel<float, MODE_LAP>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
    default:
      hipLaunchKernelGGL((ao_eval_kernel<float, MODE_HESS>), grid, block, 0, s, meta, geom, prim, nshell, nbasis,
                         nbasis_pad, pts, npts, ld, ao, screen);
      break;
  }
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
