// Electrostatic potential: host-side Hermite pre-contraction and the device
// kernel's data layout.
//
// The reference builds, for every grid point, the full nuclear-attraction
// matrix V_ij(C) of every shell pair, primitive pair and component pair
// (src/evaluate_electrostatic.cu:10076-10158) and then contracts it with P in
// a host loop of four library calls per point (:10320-10344).  Here P is
// folded into the McMurchie-Davidson Hermite expansion once per molecule:
//
//   ESP_el(C) = sum_ij P_ij V_ij(C)
//             = sum_{primitive pairs ab} sum_tuv D^ab_tuv R_tuv(p, Q - C),
//   D^ab_tuv  = (2 pi / p) sum_{i in a's shell, j in b's shell} P_ij c_a c_b
//               sum_{monomials} E^x_t E^y_u E^z_v,
//
// and with R_tuv = d^t/dX^t d^u/dY^u d^v/dZ^v F_0(p r^2) written through
// Hermite polynomials, R_tuv = (-1)^{t+u+v} sum_ijk h_ti h_uj h_vk p^N X^i Y^j
// Z^k F_N(p r^2), N = (t+u+v+i+j+k)/2, the whole pair becomes
//
//   S_ab(C) = sum_{N=0}^{L} F_N(T) sum_{ijk} g_{N,ijk} X^i Y^j Z^k,
//
// with point-independent coefficients g (max(0, 2N-L) <= i+j+k <= N).  Per
// point and pair the device evaluates one Boys vector and one polynomial.
#pragma once

#include "mol.h"

namespace ct {

// Boys-function tables.  F_m(T_g) for T_g = g * step, g = 0 .. ng - 1, and
// e^{-T_g}; a row holds F_0 .. F_{lmax + K} and then e^{-T_g}.  Below tmax,
// F_L(T) = sum_{k <= K} F_{L+k}(T_g) (-dT)^k / k! with |dT| <= step / 2: the
// remainder is at most e^{1/16} (1/16)^9 / 9! < 5e-17 of F_L (|F_m'| < F_m).
// From tmax on, F_0 = sqrt(pi / T) / 2 and the upward recursion without the
// e^{-T} term: the dropped terms stay below 2e-17 of every F_n, n <= lmax, at
// tmax = 50 for lmax = 4 and tmax = 60 for lmax = 8.  Two grids: pairs up to
// L = 4 (shells up to d) and up to L = 8 (up to g).
template <int LMAX>
struct BoysGrid {
  static constexpr double step = 0.125;
  static constexpr int K = 8;
  static constexpr int ncol = LMAX + K + 1;  // F_0 .. F_{LMAX+K}
  static constexpr int row = ncol + 1;       // + e^{-T_g}
  static constexpr double tmax = LMAX <= 4 ? 50.0 : 60.0;
  static constexpr int ng = (int)(tmax / step) + 1;
  static constexpr int size = ng * row;  // doubles
};
constexpr int kEspLmax = 8;
// the two tables, one after the other: BoysGrid<4> then BoysGrid<8>
constexpr int64_t kBoysTableDoubles = BoysGrid<4>::size + BoysGrid<8>::size;

// Number of g coefficients of a pair of total angular momentum L; the device
// and the host enumerate them in the same nested order:
//   for N = 0..L, for tot = max(0, 2N-L)..N, for i = tot..0, for j = tot-i..0, k = tot-i-j
__host__ __device__ constexpr int esp_ncoef(int L) {
  int n = 0;
  for (int N = 0; N <= L; ++N)
    for (int tot = (2 * N - L > 0 ? 2 * N - L : 0); tot <= N; ++tot) n += (tot + 1) * (tot + 2) / 2;
  return n;
}

struct EspSegment {
  int32_t L;
  int32_t count;  // primitive pairs in the segment
  int64_t off;    // offset (doubles) of the first record; a record is [Qx, Qy, Qz, p, g...]
  int64_t off32;  // mixed precision: offset (floats) of the segment's FP32 g coefficients
};

struct EspData {
  std::vector<double> records;
  std::vector<EspSegment> segs;
  int64_t npairs = 0;
  int64_t screened = 0;
};

EspData build_esp_pairs(const HostBasis& basis, const std::vector<double>& xyz, const std::vector<double>& P,
                        double screen);
// Both Boys tables (kBoysTableDoubles doubles).
std::vector<double> build_boys_tables();

// FP32 copy of the g coefficients (esp_ncoef(L) floats per pair, segment by
// segment) for the mixed-precision kernel; fills EspSegment::off32.
std::vector<float> esp_coefficients_f32(const std::vector<double>& records, std::vector<EspSegment>& segs);

// records32 == nullptr: all-FP64 kernel.  Otherwise the mixed kernel: the
// pair geometry Q - C, T = p |Q - C|^2, the Boys function F_0..F_L(T) and the
// sum over pairs in FP64; the Hermite polynomial sum_ijk g X^i Y^j Z^k and the
// pair's sum over N in FP32.  lmax: largest pair L of the segments (selects
// the Boys grid); boys_tabs: build_boys_tables() in device memory; npairs:
// the pairs of all segments.  nslice > 1 splits the pair list over that many
// workgroups per point block (part: nslice * n doubles of partial sums, added
// in slice order); block: 256 or 512 threads.
void launch_esp(const double* pts, int64_t n, const double* records, const float* records32, const EspSegment* segs,
                int nseg, int64_t npairs, int lmax, const double* atoms, int natom, const double* boys_tabs,
                double* out, double* part, int nslice, int block, hipStream_t s);
// F_0..F_m(T) at n values of T (device pointers; F is (n, m + 1) row-major)
// as the ESP kernels evaluate them for pairs up to lmax_pairs (which selects
// the grid); 0 <= m <= lmax_pairs <= 8.
void launch_boys_probe(const double* T, int64_t n, int m, int lmax_pairs, const double* boys_tabs, double* F,
                       hipStream_t s);

}  // namespace ct
