// ESP kernel: one lane per point, all lanes walk the same list of primitive
// pairs (wave-uniform record addresses -> scalar loads of the coefficients),
// grouped by total angular momentum L so each group runs fully unrolled code.
//
// Boys function F_0..F_L(T) in FP64, a method of its own: for T < 40 a
// tenth-order Taylor expansion about the nearest point of a 0.125-spaced table
// of F_m (Helgaker, Jorgensen & Olsen sec. 9.8.2; table built on the host by
// the convergent series), followed by the stable downward recursion
// F_n = (2T F_{n+1} + e^{-T}) / (2n+1); for T >= 40, F_0 = sqrt(pi/T)/2
// (erfc(sqrt(40)) < 1e-18) and the upward recursion
// F_{n+1} = ((2n+1) F_n - e^{-T}) / (2T), stable there.  (The reference's
// erf-based upward recursion, include/boys_functions.cuh:54-72, is NaN at
// T = 0 and unstable at small T.)  The table lives in LDS.
//
// ESP(C) = sum_A Z_A / |C - R_A| - sum_pairs S_pair(C)
// (reference src/evaluate_electrostatic.cu:10356-10371 for the nuclear term).
// Two instantiations: all-FP64, and mixed (pair geometry, T and the sum over
// pairs in FP64; Boys function, Hermite polynomial and the pair's sum over N
// in FP32).  The mixed Boys function uses an FP32 table that also holds
// e^{-T_g}, six Taylor terms (truncation < 1e-10 of F_L at |dT| <= 1/16),
// e^{-T} = e^{-T_g} e^{-dT} with the hardware exponential (no argument
// reduction error at large T), and the hardware reciprocal square root
// for T >= 40: F_N to a few FP32 ulps, no FP64 exp, divide or square root.

#include "esp.h"

namespace ct {

namespace {

template <int L>
__device__ __forceinline__ void boys(double T, const double* __restrict__ tab, double (&F)[L + 1]) {
  if (T < kBoysTmax) {
    const int g = (int)(T * (1.0 / kBoysStep) + 0.5);
    const double dT = T - g * kBoysStep;
    const double* row = tab + g * kBoysM + L;
    // the divisions by small integers are multiplications by their
    // reciprocals (constants after unrolling): no FP64 divide sequences
    double f = 0.0, term = 1.0;
#pragma unroll
    for (int k = 0; k <= kBoysK; ++k) {
      f += row[k] * term;
      term *= -dT * (1.0 / (k + 1));
    }
    F[L] = f;
    if constexpr (L > 0) {
      const double e = exp(-T);
#pragma unroll
      for (int n = L - 1; n >= 0; --n) F[n] = (2.0 * T * F[n + 1] + e) * (1.0 / (2 * n + 1));
    }
  } else {
    const double rT = 1.0 / T;
    F[0] = 0.5 * sqrt(M_PI * rT);
    if constexpr (L > 0) {
      const double e = exp(-T);
      const double inv2T = 0.5 * rT;
#pragma unroll
      for (int n = 0; n < L; ++n) F[n + 1] = ((2 * n + 1) * F[n] - e) * inv2T;
    }
  }
}

// FP32 Boys function of the mixed kernel; T arrives in FP64 so that the grid
// point and dT are exact before rounding.
template <int L>
__device__ __forceinline__ void boys(double Td, const float* __restrict__ tab, float (&F)[L + 1]) {
  const float T = (float)Td;
  if (Td < kBoysTmax) {
    const int g = (int)(Td * (1.0 / kBoysStep) + 0.5);
    const float x = (float)(g * kBoysStep - Td);  // -dT
    const float* row = tab + g * kBoysRow32;
    // Horner form of sum_k F_{L+k}(T_g) x^k / k!
    float f = row[L + kBoysK32];
#pragma unroll
    for (int k = kBoysK32 - 1; k >= 0; --k) f = row[L + k] + (x * (1.0f / (k + 1))) * f;
    F[L] = f;
    if constexpr (L > 0) {
      const float e = row[kBoysM32] * __expf(x);
      const float t2 = 2.0f * T;
#pragma unroll
      for (int n = L - 1; n >= 0; --n) F[n] = (t2 * F[n + 1] + e) * (1.0f / (2 * n + 1));
    }
  } else {
    const float rs = __builtin_amdgcn_rsqf(T);
    F[0] = 0.88622692545275801f * rs;  // sqrt(pi) / 2
    if constexpr (L > 0) {
      const float e = __expf(-T);
      const float inv2T = 0.5f * rs * rs;
#pragma unroll
      for (int n = 0; n < L; ++n) F[n + 1] = ((2 * n + 1) * F[n] - e) * inv2T;
    }
  }
}

// One primitive pair: F_0..F_L(T), then sum_N F_N poly_N.  TP = double: all
// in FP64.  TP = float (mixed kernel): Q - C and T in FP64, then X, Y, Z and
// the coefficients rounded to FP32 once, the Boys function, polynomial and
// sum over N in FP32; the sums over pairs stay FP64.
template <int L, typename TP>
__device__ __forceinline__ double pair_value(const double* __restrict__ rec, const TP* __restrict__ c, double cx,
                                             double cy, double cz, const TP* __restrict__ tab) {
  const double X = rec[0] - cx, Y = rec[1] - cy, Z = rec[2] - cz, p = rec[3];
  const double T = p * (X * X + Y * Y + Z * Z);
  TP F[L + 1];
  boys<L>(T, tab, F);
  TP xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = TP(1);
#pragma unroll
  for (int i = 1; i <= L; ++i) {
    xp[i] = xp[i - 1] * (TP)X;
    yp[i] = yp[i - 1] * (TP)Y;
    zp[i] = zp[i - 1] * (TP)Z;
  }
  TP s = TP(0);
  int idx = 0;
#pragma unroll
  for (int N = 0; N <= L; ++N) {
    TP poly = TP(0);
#pragma unroll
    for (int tot = (2 * N - L > 0 ? 2 * N - L : 0); tot <= N; ++tot)
#pragma unroll
      for (int i = tot; i >= 0; --i)
#pragma unroll
        for (int j = tot - i; j >= 0; --j) poly += c[idx++] * (xp[i] * yp[j] * zp[tot - i - j]);
    s += F[N] * poly;
  }
  return (double)s;
}

template <int L, typename TP>
__device__ __forceinline__ double segment_sum(const double* __restrict__ rec, const TP* __restrict__ coef, int count,
                                              double cx, double cy, double cz, const TP* __restrict__ tab) {
  constexpr int nc = esp_ncoef(L);
  constexpr int stride = 4 + nc;
  double s = 0.0;
  for (int k = 0; k < count; ++k) {
    const double* r = rec + (int64_t)k * stride;
    const TP* c;
    if constexpr (sizeof(TP) == 8)
      c = reinterpret_cast<const TP*>(r + 4);
    else
      c = coef + (int64_t)k * nc;
    s += pair_value<L, TP>(r, c, cx, cy, cz, tab);
  }
  return s;
}

template <typename TP>
__global__ __launch_bounds__(256) void esp_kernel(const double* __restrict__ pts, int64_t n,
                                                  const double* __restrict__ records,
                                                  const float* __restrict__ records32,
                                                  const EspSegment* __restrict__ segs, int nseg,
                                                  const double* __restrict__ atoms, int natom,
                                                  const double* __restrict__ boys_tab,
                                                  const float* __restrict__ boys_tab32, double* __restrict__ out) {
  constexpr int ntab = sizeof(TP) == 8 ? kBoysNG * kBoysM : kBoysNG * kBoysRow32;
  __shared__ TP tab[ntab];
  const TP* src;
  if constexpr (sizeof(TP) == 8)
    src = boys_tab;
  else
    src = boys_tab32;
  for (int k = threadIdx.x; k < ntab; k += blockDim.x) tab[k] = src[k];
  __syncthreads();
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t pc = p < n ? p : n - 1;
  const double cx = pts[3 * pc], cy = pts[3 * pc + 1], cz = pts[3 * pc + 2];
  double el = 0.0;
  for (int sg = 0; sg < nseg; ++sg) {
    const EspSegment seg = segs[sg];
    const double* rec = records + seg.off;
    const TP* c32 = reinterpret_cast<const TP*>(records32 + (sizeof(TP) == 4 ? seg.off32 : 0));
    switch (seg.L) {
      case 0: el += segment_sum<0, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      case 1: el += segment_sum<1, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      case 2: el += segment_sum<2, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      case 3: el += segment_sum<3, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      case 4: el += segment_sum<4, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      case 5: el += segment_sum<5, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      case 6: el += segment_sum<6, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      case 7: el += segment_sum<7, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
      default: el += segment_sum<8, TP>(rec, c32, seg.count, cx, cy, cz, tab); break;
    }
  }
  double nuc = 0.0;
  for (int a = 0; a < natom; ++a) {
    const double dx = cx - atoms[4 * a], dy = cy - atoms[4 * a + 1], dz = cz - atoms[4 * a + 2];
    nuc += atoms[4 * a + 3] / sqrt(dx * dx + dy * dy + dz * dz);
  }
  if (p < n) out[p] = nuc - el;
}

}  // namespace

void launch_esp(const double* pts, int64_t n, const double* records, const float* records32, const EspSegment* segs,
                int nseg, const double* atoms, int natom, const double* boys_tab, const float* boys_tab32, double* out,
                hipStream_t s) {
  if (n <= 0) return;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (records32 == nullptr)
    hipLaunchKernelGGL(esp_kernel<double>, grid, block, 0, s, pts, n, records, records32, segs, nseg, atoms, natom,
                       boys_tab, boys_tab32, out);
  else
    hipLaunchKernelGGL(esp_kernel<float>, grid, block, 0, s, pts, n, records, records32, segs, nseg, atoms, natom,
                       boys_tab, boys_tab32, out);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
