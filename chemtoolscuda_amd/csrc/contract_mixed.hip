// Mixed-precision contraction: FP32 AO values, Y = P [phi | d phi] with
// FP32 quality at bf16 matrix-core rates (the split-bf16 tile engine in
// mixed_tile.h), and FP64 per-point dot products in the epilogue.
//
// The seventh, magnitude product |P_hi| * |phi_hi| gives the per-point
// B = sum_i |phi_i| (|P| |phi|)_i that scales every rounding error of rho;
// the finaliser flags points whose estimate gamma * B exceeds the tolerance
// for FP64 re-evaluation.
//
// Decomposition: one workgroup owns a 128-row block of P and TC GEMM
// columns: 32 points x {phi, d_x phi, d_y phi, d_z phi} (lap / Hessian
// modes, TC = 128) or 64 points of phi (rho / grad modes).  The epilogue dots
// each wave's rows with the matching AO row quads (one 16-byte load per
// quad) in FP64, reduces over lane groups and the four waves in a fixed
// order, and writes one partial per (row block, dot, point); the finaliser
// sums the row blocks.

#include <type_traits>

#include "mixed_tile.h"

namespace ct {

namespace {

using namespace mixed;

// GEMM columns per workgroup: 128 = 32 points x 4 column types in the lap /
// Hessian modes; 64 points of one type in the rho / grad modes, where every
// column also carries the bound accumulator.
template <int MODE>
struct Cols {
  static constexpr int NT = (MODE == MODE_RHO || MODE == MODE_GRAD) ? 1 : 4;
  static constexpr int TC = NT == 1 ? 64 : 128;
  static constexpr int NCB = TC / 16;  // 16-column MFMA blocks
  static constexpr int PTS = TC / NT;  // points per workgroup
  __host__ __device__ static constexpr int type(int c) { return NT == 1 ? 0 : c / PTS; }
  __host__ __device__ static constexpr int point(int c) { return NT == 1 ? c : c % PTS; }
};

// Dots produced from a GEMM column of type t (phi, d_x phi, d_y phi, d_z phi):
// how many, and for the k-th one its slot in the dot layout and the AO array
// it pairs the column with.
template <int MODE>
__host__ __device__ constexpr int ndots_of(int t) {
  if (MODE == MODE_RHO) return 1;
  if (MODE == MODE_GRAD) return 4;
  if (MODE == MODE_LAP) return t == 0 ? 5 : 1;
  return t == 0 ? 10 : t;  // HESS: t_xy.. lists grow with t
}
template <int MODE>
__host__ __device__ constexpr int dot_slot(int t, int k) {
  if (MODE == MODE_RHO || MODE == MODE_GRAD || t == 0) return k;
  if (MODE == MODE_LAP) return 4 + t;
  return (t == 1 ? 10 : t == 2 ? 11 : 13) + k;  // t_xx | t_xy t_yy | t_xz t_yz t_zz
}
template <int MODE>
__host__ __device__ constexpr int dot_array(int t, int k) {
  if (MODE == MODE_RHO || MODE == MODE_GRAD || t == 0) return k;
  if (MODE == MODE_LAP) return t;
  return 1 + k;
}

template <int N>
using IC = std::integral_constant<int, N>;

constexpr int kMaxDots = 11;  // 10 Hessian-mode dots of the phi column + the bound

template <int MODE>
__global__ __launch_bounds__(256, 2) void contract_mixed_kernel(const u4* __restrict__ Pf, int mp,
                                                                const float* __restrict__ ao, int64_t ld,
                                                                double* __restrict__ part, int nrt,
                                                                int64_t nptile) {
  using C = Cols<MODE>;
  constexpr int TC = C::TC, NCB = C::NCB;
  constexpr int ND = mode_ndot(MODE);
  constexpr int ND1 = ND + 1;
  constexpr int NBC = C::NT == 1 ? NCB : 2;  // column blocks of type phi (they carry the bound)
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][NPL][TC][LROW];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
  const int r = lane & 15, q = lane >> 4;
  // XCD-aware tile order: workgroup g runs on XCD g % 8 (round-robin
  // dispatch); the nrt row blocks of one point tile get consecutive slots on
  // the same XCD so its AO tile is fetched from HBM once and re-read from L2.
  const int g = blockIdx.x;
  const int xcd = g & 7, slot = g >> 3;
  const int rt = slot % nrt;
  const int64_t ptile = (int64_t)(slot / nrt) * 8 + xcd;
  if (ptile >= nptile) return;
  const int64_t pbase = ptile * C::PTS;
  const int i0 = rt * TR;
  const int64_t mq = mp / 4;  // row quads per AO array

  // Lap mode: the derivative columns enter only tau_kk = d_k phi^T P d_k phi,
  // a quadratic form, so they use the symmetric-P shortcut of y_tile (0.6 of
  // their K range for five row blocks).  The Hessian mode's t_kl (k != l) are
  // not quadratic forms and use the full K.
  constexpr bool SYM = MODE == MODE_LAP;
  constexpr bool F64 = MODE != MODE_HESS;  // the Hessian mode has no registers for FP64 sums
  YTile<TC, NBC, F64> yt;
  const float* bsrc = ao + ((int64_t)C::type(tid & (TC - 1)) * mq * ld + pbase + C::point(tid & (TC - 1))) * 4;
  y_tile<TC, NBC, SYM, F64>(Pf, mp, ld, i0, tid, bsrc, Bs, yt, C::type(tid & (TC - 1)) != 0, i0 / KB,
                            (i0 + TR) / KB);

  // ---------------------------------------------------------------- epilogue
  double* red = reinterpret_cast<double*>(&Bs[0][0][0][0]);  // [NW][TC][kMaxDots]
  const int64_t quad0 = (i0 + RW * w) / 4 + q;
  auto per_cb = [&](auto cbtag) {
    constexpr int CB = decltype(cbtag)::value;
    constexpr int t = C::type(16 * CB);
    constexpr int nd = ndots_of<MODE>(t);
    constexpr bool bnd = CB < NBC;
    const int64_t p = pbase + C::point(16 * CB) + r;
    double d[kMaxDots];
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k) d[k] = 0.0;
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
      f4 v[nd];
#pragma unroll
      for (int k = 0; k < nd; ++k)
        v[k] = *reinterpret_cast<const f4*>(ao + (((int64_t)dot_array<MODE>(t, k) * mq + quad0 + 4 * rb) * ld + p) * 4);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double y = yt.y(rb, CB, j);
#pragma unroll
        for (int k = 0; k < nd; ++k) d[k] += (double)v[k][j] * y;
        if constexpr (bnd) d[kMaxDots - 1] += fabs((double)v[0][j]) * (double)yt.bound[rb][CB % NBC][j];
      }
    }
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k) {
      if (k < nd || (bnd && k == kMaxDots - 1)) {
        double x = d[k];
        x += __shfl_xor(x, 16);
        x += __shfl_xor(x, 32);
        d[k] = x;
      }
    }
    if (q == 0) {
      double* dst = red + ((int64_t)w * TC + 16 * CB + r) * kMaxDots;
#pragma unroll
      for (int k = 0; k < kMaxDots; ++k)
        if (k < nd || (bnd && k == kMaxDots - 1)) dst[k] = d[k];
    }
  };
  per_cb(IC<0>{});
  per_cb(IC<1>{});
  per_cb(IC<2>{});
  per_cb(IC<3>{});
  if constexpr (NCB == 8) {
    per_cb(IC<4>{});
    per_cb(IC<5>{});
    per_cb(IC<6>{});
    per_cb(IC<7>{});
  }
  __syncthreads();
  if (tid < TC) {
    const int c = tid;
    const int t = C::type(c);
    const int nd = ndots_of<MODE>(t);
    const int64_t p = pbase + C::point(c);
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k) {
      const bool bound = (k == kMaxDots - 1) && t == 0;
      if (k < nd || bound) {
        double sum = 0.0;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) sum += red[((int64_t)ww * TC + c) * kMaxDots + k];
        const int slot_k = bound ? ND : dot_slot<MODE>(t, k);
        part[((int64_t)rt * ND1 + slot_k) * ld + p] = sum;
      }
    }
  }
}

}  // namespace

int mixed_points_per_block(int mode) {
  return mode == MODE_RHO ? Cols<MODE_RHO>::PTS : mode == MODE_GRAD ? Cols<MODE_GRAD>::PTS : Cols<MODE_LAP>::PTS;
}

void mixed_fragment_order_P(const uint16_t* planes, int mp, uint16_t* out) {
  // out[((pl * mp/16 + b16) * nk + kb) * 64 + lane][e] = plane pl of
  // P[16 b16 + lane % 16][32 kb + 8 (lane / 16) + e]
  const int nk = mp / KB, nb = mp / 16;
  const size_t plane = (size_t)mp * mp;
  size_t o = 0;
  for (int pl = 0; pl < NPL; ++pl)
    for (int b16 = 0; b16 < nb; ++b16)
      for (int kb = 0; kb < nk; ++kb)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e)
            out[o++] = planes[pl * plane + (size_t)(16 * b16 + (l & 15)) * mp + 32 * kb + 8 * (l >> 4) + e];
}

void launch_contract_mixed(int mode, const uint16_t* Pf, int nbasis_pad, const float* ao, int64_t ld, double* part,
                           hipStream_t s) {
  CT_REQUIRE(nbasis_pad % TR == 0, "padded basis size must be a multiple of 128");
  const int ppb = mixed_points_per_block(mode);
  const int nrt = nbasis_pad / TR;
  const int64_t nptile = ld / ppb;
  const dim3 block(256), grid((unsigned)(((nptile + 7) / 8) * 8 * nrt));
  const u4* P = reinterpret_cast<const u4*>(Pf);
  switch (mode) {
    case MODE_RHO:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_RHO>, grid, block, 0, s, P, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
    case MODE_GRAD:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_GRAD>, grid, block, 0, s, P, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
    case MODE_LAP:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_LAP>, grid, block, 0, s, P, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
    default:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_HESS>, grid, block, 0, s, P, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
  }
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
