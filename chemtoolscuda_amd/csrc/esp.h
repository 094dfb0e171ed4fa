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

// Boys-function Taylor table: F_m(T_g) for T_g = g * kBoysStep, g = 0..kBoysNG-1
// (T < kBoysTmax), m = 0..kBoysM-1.  F_L(T) = sum_{k<=kBoysK} F_{L+k}(T_g) (-dT)^k / k!.
constexpr int kEspLmax = 8;
constexpr int kBoysK = 10;
constexpr int kBoysM = kEspLmax + kBoysK + 1;
constexpr double kBoysStep = 0.125;
constexpr double kBoysTmax = 40.0;
constexpr int kBoysNG = 321;

// FP32 table of the mixed kernel, same grid: row g holds F_0..F_{kBoysM32-1}(T_g)
// and e^{-T_g} at column kBoysM32; rows kBoysRow32 floats apart (odd, so that
// lanes at different grid points read different LDS banks).  Six Taylor terms
// (k = 0..5, |dT| <= 1/16) leave a truncation of (1/16)^6 / 6! < 1e-10 of F_L.
constexpr int kBoysK32 = 5;
constexpr int kBoysM32 = kEspLmax + kBoysK32 + 1;
constexpr int kBoysRow32 = kBoysM32 + 1;

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
std::vector<double> build_boys_table();
std::vector<float> build_boys_table_f32();

// FP32 copy of the g coefficients (esp_ncoef(L) floats per pair, segment by
// segment) for the mixed-precision kernel; fills EspSegment::off32.
std::vector<float> esp_coefficients_f32(const std::vector<double>& records, std::vector<EspSegment>& segs);

// records32 == nullptr: all-FP64 kernel (boys_tab).  Otherwise the mixed
// kernel (boys_tab32): the pair geometry Q - C, T = p |Q - C|^2 and the sum
// over pairs in FP64; the Boys function, the Hermite polynomial
// sum_ijk g X^i Y^j Z^k and the pair's sum over N in FP32.
void launch_esp(const double* pts, int64_t n, const double* records, const float* records32, const EspSegment* segs,
                int nseg, const double* atoms, int natom, const double* boys_tab, const float* boys_tab32, double* out,
                hipStream_t s);

}  // namespace ct
