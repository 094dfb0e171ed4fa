// ESP kernel: one lane per point, all lanes walk the same list of primitive
// pairs (wave-uniform record addresses -> scalar loads of the coefficients),
// grouped by total angular momentum L so each group runs fully unrolled code.
//
// Boys function F_0..F_L(T), in FP64 in both kernels, a method of its own
// (the reference's erf-based upward recursion, include/boys_functions.cuh:54-72,
// is NaN at T = 0 and unstable at small T):
//   * T < tmax: F_L(T) by an eighth-order Taylor expansion about the nearest
//     point T_g of a 0.125-spaced table of F_m (Helgaker, Jorgensen & Olsen
//     sec. 9.8.2; the table is built on the host by the convergent series),
//     e^{-T} = e^{-T_g} e^{-dT} with e^{-T_g} from the same table row and
//     e^{-dT} by its Taylor polynomial (|dT| <= 1/16: no FP64 exp), then the
//     stable downward recursion F_n = (2T F_{n+1} + e^{-T}) / (2n+1);
//   * T >= tmax (50 for pairs up to L = 4, 60 up to L = 8): F_0 =
//     sqrt(pi / T) / 2 from the hardware reciprocal square root refined by two
//     Newton steps, and the upward recursion F_{n+1} = (2n+1) F_n / (2T),
//     whose dropped e^{-T} term is below 2e-17 of every F_n there.
// The table lives in LDS (45 KB for pairs up to L = 4, 69 KB up to L = 8).
//
// ESP(C) = sum_A Z_A / |C - R_A| - sum_pairs S_pair(C)
// (reference src/evaluate_electrostatic.cu:10356-10371 for the nuclear term).
// Two instantiations per grid: all-FP64, and mixed: the pair geometry, T, the
// Boys function and the sum over pairs in FP64; the Hermite polynomial and the
// pair's sum over N in FP32 (F_N rounded once).

#include "esp.h"

namespace ct {

namespace {

template <int L, int LMAXG>
__device__ __forceinline__ void boys_f64(double T, const double* __restrict__ tab, double (&F)[L + 1]) {
  using G = BoysGrid<LMAXG>;
  if (T < G::tmax) {
    const int g = (int)(T * (1.0 / G::step) + 0.5);
    const double x = g * G::step - T;  // -dT, exact
    const double* row = tab + g * G::row;
    // Horner form of sum_k F_{L+k}(T_g) x^k / k!
    double f = row[L + G::K];
#pragma unroll
    for (int k = G::K - 1; k >= 0; --k) f = fma(x * (1.0 / (k + 1)), f, row[L + k]);
    F[L] = f;
    if constexpr (L > 0) {
      // e^x, |x| <= 1/16: degree 9, remainder < 3e-19
      constexpr double c[10] = {1.0,
                                1.0,
                                0.5,
                                1.6666666666666666574e-01,
                                4.1666666666666664354e-02,
                                8.3333333333333332177e-03,
                                1.3888888888888889419e-03,
                                1.9841269841269841253e-04,
                                2.4801587301587301566e-05,
                                2.7557319223985892511e-06};
      double ex = c[9];
#pragma unroll
      for (int k = 8; k >= 0; --k) ex = fma(ex, x, c[k]);
      const double e = row[G::ncol] * ex;
      const double t2 = 2.0 * T;
#pragma unroll
      for (int n = L - 1; n >= 0; --n) F[n] = fma(t2, F[n + 1], e) * (1.0 / (2 * n + 1));
    }
  } else {
    // 1/sqrt(T): hardware estimate, two Newton steps y += y (1 - T y^2) / 2
    double y = __builtin_amdgcn_rsq(T);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const double r = fma(-(T * y), y, 1.0);
      y = fma(0.5 * y, r, y);
    }
    F[0] = 0.88622692545275801365 * y;  // sqrt(pi) / 2
    if constexpr (L > 0) {
      const double inv2T = 0.5 * y * y;
#pragma unroll
      for (int n = 0; n < L; ++n) F[n + 1] = ((2 * n + 1) * inv2T) * F[n];
    }
  }
}

// One primitive pair: F_0..F_L(T) in FP64, then sum_N F_N poly_N.  TP = double:
// all in FP64.  TP = float (mixed kernel): X, Y, Z, the coefficients and F_N
// rounded to FP32 once; the polynomial and the sum over N in FP32.
template <int L, int LMAXG, typename TP>
__device__ __forceinline__ double pair_value(const double* __restrict__ rec, const TP* __restrict__ c, double cx,
                                             double cy, double cz, const double* __restrict__ tab) {
  const double X = rec[0] - cx, Y = rec[1] - cy, Z = rec[2] - cz, p = rec[3];
  const double T = p * (X * X + Y * Y + Z * Z);
  double F[L + 1];
  boys_f64<L, LMAXG>(T, tab, F);
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
    s += (TP)F[N] * poly;
  }
  return (double)s;
}

// sum_A Z_A / |C - R_A| (reference src/evaluate_electrostatic.cu:10356-10371)
__device__ __forceinline__ double nuclear_term(double cx, double cy, double cz, const double* __restrict__ atoms,
                                               int natom) {
  double nuc = 0.0;
  for (int a = 0; a < natom; ++a) {
    const double dx = cx - atoms[4 * a], dy = cy - atoms[4 * a + 1], dz = cz - atoms[4 * a + 2];
    nuc += atoms[4 * a + 3] / sqrt(dx * dx + dy * dy + dz * dz);
  }
  return nuc;
}

template <int L, int LMAXG, typename TP>
__device__ __forceinline__ double segment_sum(const double* __restrict__ rec, const TP* __restrict__ coef, int count,
                                              double cx, double cy, double cz, const double* __restrict__ tab) {
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
    s += pair_value<L, LMAXG, TP>(r, c, cx, cy, cz, tab);
  }
  return s;
}

// Grid: x over points, y over slices of the pair list (blockIdx.y takes
// pairs [npairs y / S, npairs (y+1) / S) in list order).  S = 1: the kernel
// adds the nuclear term and writes the ESP; S > 1: it writes its slice's
// electronic sum to part[y * n + p] and esp_reduce_kernel finishes.
template <typename TP, int LMAXG, int BLOCK>
__global__ __launch_bounds__(BLOCK) void esp_kernel(const double* __restrict__ pts, int64_t n,
                                                  const double* __restrict__ records,
                                                  const float* __restrict__ records32,
                                                  const EspSegment* __restrict__ segs, int nseg, int64_t npairs,
                                                  const double* __restrict__ atoms, int natom,
                                                  const double* __restrict__ boys_tab, double* __restrict__ out,
                                                  double* __restrict__ part) {
  using G = BoysGrid<LMAXG>;
  __shared__ double tab[G::size];
  for (int k = threadIdx.x; k < G::size; k += blockDim.x) tab[k] = boys_tab[k];
  __syncthreads();
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t pc = p < n ? p : n - 1;
  const double cx = pts[3 * pc], cy = pts[3 * pc + 1], cz = pts[3 * pc + 2];
  const int64_t lo = npairs * blockIdx.y / gridDim.y, hi = npairs * (blockIdx.y + 1) / gridDim.y;
  double el = 0.0;
  int64_t first = 0;  // list index of the segment's first pair
  for (int sg = 0; sg < nseg && first < hi; ++sg) {
    const EspSegment seg = segs[sg];
    const int64_t b = lo > first ? lo - first : 0;
    const int64_t e = hi - first < seg.count ? hi - first : seg.count;
    first += seg.count;
    if (b >= e) continue;
    const int count = (int)(e - b);
    const double* rec = records + seg.off + b * (4 + esp_ncoef(seg.L));
    const TP* c32 = reinterpret_cast<const TP*>(records32 + (sizeof(TP) == 4 ? seg.off32 + b * esp_ncoef(seg.L) : 0));
    switch (seg.L) {
      case 0: el += segment_sum<0, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
      case 1: el += segment_sum<1, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
      case 2: el += segment_sum<2, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
      case 3: el += segment_sum<3, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
      default:
        if constexpr (LMAXG > 4) {
          switch (seg.L) {
            case 4: el += segment_sum<4, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
            case 5: el += segment_sum<5, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
            case 6: el += segment_sum<6, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
            case 7: el += segment_sum<7, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
            default: el += segment_sum<8, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab); break;
          }
        } else {
          el += segment_sum<4, LMAXG, TP>(rec, c32, count, cx, cy, cz, tab);
        }
        break;
    }
  }
  if (p >= n) return;
  if (gridDim.y > 1) {
    part[(int64_t)blockIdx.y * n + p] = el;
    return;
  }
  out[p] = nuclear_term(cx, cy, cz, atoms, natom) - el;
}

// out[p] = nuclear term - sum of the slices' electronic sums, in slice order
__global__ __launch_bounds__(256) void esp_reduce_kernel(const double* __restrict__ pts, int64_t n, int nslice,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ atoms, int natom,
                                                         double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double el = 0.0;
  for (int s = 0; s < nslice; ++s) el += part[(int64_t)s * n + p];
  out[p] = nuclear_term(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], atoms, natom) - el;
}

template <int M, int LMAXG>
__global__ __launch_bounds__(256) void boys_probe_kernel(const double* __restrict__ T, int64_t n,
                                                         const double* __restrict__ boys_tab, double* __restrict__ F) {
  using G = BoysGrid<LMAXG>;
  __shared__ double tab[G::size];
  for (int k = threadIdx.x; k < G::size; k += blockDim.x) tab[k] = boys_tab[k];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[M + 1];
  boys_f64<M, LMAXG>(T[i], tab, f);
#pragma unroll
  for (int m = 0; m <= M; ++m) F[i * (M + 1) + m] = f[m];
}

template <int M, int LMAXG>
void probe(const double* T, int64_t n, const double* tab, double* F, hipStream_t s) {
  hipLaunchKernelGGL((boys_probe_kernel<M, LMAXG>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, T, n, tab, F);
}

}  // namespace

void launch_esp(const double* pts, int64_t n, const double* records, const float* records32, const EspSegment* segs,
                int nseg, int64_t npairs, int lmax, const double* atoms, int natom, const double* boys_tabs,
                double* out, double* part, int nslice, int block, hipStream_t s) {
  if (n <= 0) return;
  CT_REQUIRE(lmax >= 0 && lmax <= kEspLmax, "ESP: pair angular momentum out of range");
  CT_REQUIRE(nslice >= 1 && (nslice == 1 || part != nullptr), "ESP: slices need a partial-sum buffer");
  CT_REQUIRE(block == 256 || block == 512 || (block == 1024 && lmax <= 4), "ESP: unsupported block size");
  const dim3 grid((unsigned)((n + block - 1) / block), (unsigned)nslice), blk(block);
  const double* tab = lmax <= 4 ? boys_tabs : boys_tabs + BoysGrid<4>::size;
#define CT_ESP_LAUNCH(TP, G, B)                                                                                  \
  hipLaunchKernelGGL((esp_kernel<TP, G, B>), grid, blk, 0, s, pts, n, records, records32, segs, nseg, npairs, atoms, \
                     natom, tab, out, part)
  const bool mixed = records32 != nullptr;
  if (lmax <= 4) {
    if (block == 1024) {
      if (mixed) CT_ESP_LAUNCH(float, 4, 1024); else CT_ESP_LAUNCH(double, 4, 1024);
    } else if (block == 512) {
      if (mixed) CT_ESP_LAUNCH(float, 4, 512); else CT_ESP_LAUNCH(double, 4, 512);
    } else {
      if (mixed) CT_ESP_LAUNCH(float, 4, 256); else CT_ESP_LAUNCH(double, 4, 256);
    }
  } else {
    if (block == 512) {
      if (mixed) CT_ESP_LAUNCH(float, 8, 512); else CT_ESP_LAUNCH(double, 8, 512);
    } else {
      if (mixed) CT_ESP_LAUNCH(float, 8, 256); else CT_ESP_LAUNCH(double, 8, 256);
    }
  }
#undef CT_ESP_LAUNCH
  CT_HIP_CHECK(hipGetLastError());
  if (nslice > 1) {
    hipLaunchKernelGGL(esp_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pts, n, nslice, part,
                       atoms, natom, out);
    CT_HIP_CHECK(hipGetLastError());
  }
}

void launch_boys_probe(const double* T, int64_t n, int m, int lmax_pairs, const double* boys_tabs, double* F,
                       hipStream_t s) {
  if (n <= 0) return;
  CT_REQUIRE(m >= 0 && m <= lmax_pairs && lmax_pairs <= kEspLmax, "Boys order out of range (0 <= m <= lmax <= 8)");
  if (lmax_pairs <= 4) {
    switch (m) {
      case 0: probe<0, 4>(T, n, boys_tabs, F, s); break;
      case 1: probe<1, 4>(T, n, boys_tabs, F, s); break;
      case 2: probe<2, 4>(T, n, boys_tabs, F, s); break;
      case 3: probe<3, 4>(T, n, boys_tabs, F, s); break;
      default: probe<4, 4>(T, n, boys_tabs, F, s); break;
    }
  } else {
    const double* tab8 = boys_tabs + BoysGrid<4>::size;
    switch (m) {
      case 0: probe<0, 8>(T, n, tab8, F, s); break;
      case 1: probe<1, 8>(T, n, tab8, F, s); break;
      case 2: probe<2, 8>(T, n, tab8, F, s); break;
      case 3: probe<3, 8>(T, n, tab8, F, s); break;
      case 4: probe<4, 8>(T, n, tab8, F, s); break;
      case 5: probe<5, 8>(T, n, tab8, F, s); break;
      case 6: probe<6, 8>(T, n, tab8, F, s); break;
      case 7: probe<7, 8>(T, n, tab8, F, s); break;
      default: probe<8, 8>(T, n, tab8, F, s); break;
    }
  }
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
