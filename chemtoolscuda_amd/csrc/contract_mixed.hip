// Mixed-precision contraction: FP32 AO values, P and the AOs split into three
// bf16 pieces each (x = hi + mid + lo exactly for FP32 x), six bf16 MFMA
// products per tile (hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid) accumulated
// in FP32 -- FP32-quality Y = P [phi | d phi] at bf16 matrix-core rates -- and
// FP64 per-point dot products in the epilogue.
//
// A seventh product |P_hi| * |phi_hi| (bf16, FP32 accumulation; bf16(|x|) =
// |bf16(x)|, so both are sign-bit masks of the hi pieces) gives the
// per-point magnitude B = sum_i |phi_i| (|P| |phi|)_i that scales every
// rounding error of rho; the finaliser flags points whose estimate
// gamma * B exceeds the tolerance for FP64 re-evaluation.
//
// Decomposition: one workgroup (4 waves) owns a 128-row block of P and 128
// GEMM columns: 128 points of one column type (rho / grad
// modes) or 32 points x {phi, d_x phi, d_y phi, d_z phi} (lap / Hessian
// modes).  K advances 32 at a time through double-buffered LDS tiles:
//   A: P pieces (precomputed on the host, row-major, k contiguous),
//   B: FP32 AOs read coalesced from the k-major AO buffer and split into
//      bf16 pieces on the way into LDS, stored column-major (k contiguous)
//      so each lane's MFMA operand is one 16-byte LDS read.
// Each wave computes a 64 x 64 sub-tile (4 x 4 MFMA 16x16x32 blocks).  The
// epilogue dots its rows with the matching FP32 AO rows in FP64, reduces
// across lane groups and the two row halves, and writes one partial per
// (row block, dot, point); the finaliser sums the row blocks.
//
// MFMA 16x16x32 bf16 lane layout (lane l, r = l % 16, q = l / 16), measured:
//   A: A[r][8q + e], B: B[8q + e][r] (e = 0..7), D: D[4q + j][r] (j = 0..3).

#include <type_traits>

#include "ct_internal.h"

namespace ct {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int TR = 128;         // rows of P per workgroup
constexpr int TC = 128;         // GEMM columns per workgroup
constexpr int KB = 32;          // k per stage (= MFMA depth)
constexpr int LROW = KB + 8;    // bf16 per LDS row (80 bytes: staggers banks)
constexpr int NPA = 3;          // A pieces: hi, mid, lo (|P| hi = |hi| is a bit mask)
constexpr int NPB = 3;          // B pieces: hi, mid, lo

__device__ __forceinline__ bf16x8 as_bf(u4 v) { return __builtin_bit_cast(bf16x8, v); }

template <int MODE>
struct Cols {
  static constexpr int NT = (MODE == MODE_RHO || MODE == MODE_GRAD) ? 1 : 4;
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
__global__ __launch_bounds__(256, 1) void contract_mixed_kernel(const uint16_t* __restrict__ Pp, int mp,
                                                                const float* __restrict__ ao, int64_t ld,
                                                                double* __restrict__ part, int nrt,
                                                                int64_t nptile) {
  using C = Cols<MODE>;
  constexpr int ND = mode_ndot(MODE);
  constexpr int ND1 = ND + 1;
  __shared__ __attribute__((aligned(16))) uint16_t As[2][NPA][TR][LROW];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][NPB][TC][LROW];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
  const int wr = w >> 1, wc = w & 1;
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
  const int64_t sd = (int64_t)mp * ld;
  const int nk = mp / KB;

  // staging roles
  const int arow = tid >> 1, akh = tid & 1;  // A: row, 16-wide k half
  const int bcol = tid & (TC - 1), bkh = tid >> 7;
  const float* bsrc = ao + C::type(bcol) * sd + (int64_t)(16 * bkh) * ld + pbase + C::point(bcol);
  const uint16_t* asrc = Pp + (int64_t)(i0 + arow) * mp + 16 * akh;
  const int64_t pplane = (int64_t)mp * mp;

  // Global -> register staging of k-block kb+1 is issued before computing
  // k-block kb; the registers go to the other LDS buffer afterwards.
  u4 ra[1][NPA][2];
  float rb[1][16];
  auto load_stage = [&](auto buf, int kb) {
    constexpr int b = decltype(buf)::value;
    const int k0 = kb * KB;
#pragma unroll
    for (int pl = 0; pl < NPA; ++pl) {
      const u4* src = reinterpret_cast<const u4*>(asrc + pl * pplane + k0);
      ra[b][pl][0] = src[0];
      ra[b][pl][1] = src[1];
    }
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) rb[b][kk] = bsrc[(int64_t)(k0 + kk) * ld];
  };
  auto store_stage = [&](int s, auto buf) {
    constexpr int b = decltype(buf)::value;
#pragma unroll
    for (int pl = 0; pl < NPA; ++pl) {
      u4* d = reinterpret_cast<u4*>(&As[s][pl][arow][16 * akh]);
      d[0] = ra[b][pl][0];
      d[1] = ra[b][pl][1];
    }
    // exact three-piece split with the hardware RNE pack (v_cvt_pk_bf16_f32)
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      u4 vh, vm, vl;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int k = 8 * half + 2 * e;
        const float x0 = rb[b][k], x1 = rb[b][k + 1];
        bf16x2 h, m, l;
        h[0] = (__bf16)x0;
        h[1] = (__bf16)x1;
        const float r0 = x0 - (float)h[0], r1 = x1 - (float)h[1];
        m[0] = (__bf16)r0;
        m[1] = (__bf16)r1;
        l[0] = (__bf16)(r0 - (float)m[0]);
        l[1] = (__bf16)(r1 - (float)m[1]);
        vh[e] = __builtin_bit_cast(uint32_t, h);
        vm[e] = __builtin_bit_cast(uint32_t, m);
        vl[e] = __builtin_bit_cast(uint32_t, l);
      }
      *reinterpret_cast<u4*>(&Bs[s][0][bcol][16 * bkh + 8 * half]) = vh;
      *reinterpret_cast<u4*>(&Bs[s][1][bcol][16 * bkh + 8 * half]) = vm;
      *reinterpret_cast<u4*>(&Bs[s][2][bcol][16 * bkh + 8 * half]) = vl;
    }
  };
  using B0 = IC<0>;

  // Column blocks carrying phi (type 0) also accumulate the bound: all four in
  // the one-type modes, blocks 0 and 1 of the wc == 0 waves in the four-type
  // modes (columns 0..31 are phi of the 32 points).
  constexpr int NB = C::NT == 1 ? 4 : 2;
  const bool has_bound = C::NT == 1 || wc == 0;

  f4 acc[4][4], accb[4][NB];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < NB; ++b) accb[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  }

  const u4 absmask = u4{0x7FFF7FFFu, 0x7FFF7FFFu, 0x7FFF7FFFu, 0x7FFF7FFFu};
  auto compute = [&](int s) {
    bf16x8 bh[4], bm[4], bl[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int col = 64 * wc + 16 * cb + r;
      bh[cb] = as_bf(*reinterpret_cast<const u4*>(&Bs[s][0][col][8 * q]));
      bm[cb] = as_bf(*reinterpret_cast<const u4*>(&Bs[s][1][col][8 * q]));
      bl[cb] = as_bf(*reinterpret_cast<const u4*>(&Bs[s][2][col][8 * q]));
    }
    auto load_a = [&](int plane, bf16x8 (&a)[4]) {
#pragma unroll
      for (int rbk = 0; rbk < 4; ++rbk)
        a[rbk] = as_bf(*reinterpret_cast<const u4*>(&As[s][plane][64 * wr + 16 * rbk + r][8 * q]));
    };
    // One A piece live at a time.  Products with the smallest pieces come
    // first so the large hi*hi term is added last; within a sweep the 16
    // accumulators are independent, so dependent MFMAs are 16 issues apart.
#define CT_MFMA_SWEEP(A, B)                                                                        \
  _Pragma("unroll") for (int rbk = 0; rbk < 4; ++rbk) _Pragma("unroll") for (int cb = 0; cb < 4; ++cb) \
      acc[rbk][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[rbk], B[cb], acc[rbk][cb], 0, 0, 0);
    bf16x8 a[4];
    load_a(2, a);          // lo
    CT_MFMA_SWEEP(a, bh)
    load_a(1, a);          // mid
    CT_MFMA_SWEEP(a, bm)
    CT_MFMA_SWEEP(a, bh)
    load_a(0, a);          // hi
    CT_MFMA_SWEEP(a, bl)
    CT_MFMA_SWEEP(a, bm)
    CT_MFMA_SWEEP(a, bh)
#undef CT_MFMA_SWEEP
    if (has_bound) {
#pragma unroll
      for (int rbk = 0; rbk < 4; ++rbk)
#pragma unroll
        for (int cb = 0; cb < NB; ++cb)
          accb[rbk][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              as_bf(__builtin_bit_cast(u4, a[rbk]) & absmask), as_bf(__builtin_bit_cast(u4, bh[cb]) & absmask),
              accb[rbk][cb], 0, 0, 0);
    }
  };

  load_stage(B0{}, 0);
  store_stage(0, B0{});
  __syncthreads();
  for (int kb = 0; kb < nk; ++kb) {
    const int s = kb & 1;
    if (kb + 1 < nk) load_stage(B0{}, kb + 1);
    compute(s);
    if (kb + 1 < nk) store_stage(s ^ 1, B0{});
    __syncthreads();
  }

  // ---------------------------------------------------------------- epilogue
  // Per column block (all compile-time for a given wave column half): issue
  // every AO load of a row block at once, then accumulate in FP64.
  __syncthreads();
  double* red = reinterpret_cast<double*>(&As[0][0][0][0]);  // 2 x 16 lanes x 4 blocks x kMaxDots doubles
  double dacc[4][kMaxDots];
  auto epilogue_dots = [&](auto wctag) {
    constexpr int WC = decltype(wctag)::value;
    auto per_cb = [&](auto cbtag) {
      constexpr int CB = decltype(cbtag)::value;
      constexpr int col0 = 64 * WC + 16 * CB;
      constexpr int t = C::type(col0);
      constexpr int nd = ndots_of<MODE>(t);
      constexpr bool bnd = (t == 0) && (CB < NB);
      const int64_t p = pbase + C::point(col0) + r;
#pragma unroll
      for (int k = 0; k < kMaxDots; ++k) dacc[CB][k] = 0.0;
#pragma unroll
      for (int rbk = 0; rbk < 4; ++rbk) {
        float v[4][nd];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t i = i0 + 64 * wr + 16 * rbk + 4 * q + j;
#pragma unroll
          for (int k = 0; k < nd; ++k) v[j][k] = ao[dot_array<MODE>(t, k) * sd + i * ld + p];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double y = (double)acc[rbk][CB][j];
#pragma unroll
          for (int k = 0; k < nd; ++k) dacc[CB][k] += (double)v[j][k] * y;
          if constexpr (bnd) dacc[CB][kMaxDots - 1] += fabs((double)v[j][0]) * (double)accb[rbk][CB < NB ? CB : 0][j];
        }
      }
#pragma unroll
      for (int k = 0; k < kMaxDots; ++k) {
        if (k < nd || (bnd && k == kMaxDots - 1)) {
          double x = dacc[CB][k];
          x += __shfl_xor(x, 16);
          x += __shfl_xor(x, 32);
          dacc[CB][k] = x;
        }
      }
      // combine the two row halves (wr = 0, 1) through LDS
      double* slot = red + ((WC * 16 + r) * 4 + CB) * kMaxDots;
      if (wr == 1 && q == 0) {
#pragma unroll
        for (int k = 0; k < kMaxDots; ++k)
          if (k < nd || (bnd && k == kMaxDots - 1)) slot[k] = dacc[CB][k];
      }
    };
    per_cb(IC<0>{});
    per_cb(IC<1>{});
    per_cb(IC<2>{});
    per_cb(IC<3>{});
  };
  auto epilogue_write = [&](auto wctag) {
    constexpr int WC = decltype(wctag)::value;
    auto write_cb = [&](auto cbtag) {
      constexpr int CB = decltype(cbtag)::value;
      constexpr int col0 = 64 * WC + 16 * CB;
      constexpr int t = C::type(col0);
      constexpr int nd = ndots_of<MODE>(t);
      constexpr bool bnd = (t == 0) && (CB < NB);
      const int64_t p = pbase + C::point(col0) + r;
      const double* slot = red + ((WC * 16 + r) * 4 + CB) * kMaxDots;
      if (wr == 0 && q == 0) {
#pragma unroll
        for (int k = 0; k < nd; ++k)
          part[((int64_t)rt * ND1 + dot_slot<MODE>(t, k)) * ld + p] = dacc[CB][k] + slot[k];
        if constexpr (bnd)
          part[((int64_t)rt * ND1 + ND) * ld + p] = dacc[CB][kMaxDots - 1] + slot[kMaxDots - 1];
      }
    };
    write_cb(IC<0>{});
    write_cb(IC<1>{});
    write_cb(IC<2>{});
    write_cb(IC<3>{});
  };
  if (wc == 0)
    epilogue_dots(IC<0>{});
  else
    epilogue_dots(IC<1>{});
  __syncthreads();
  if (wc == 0)
    epilogue_write(IC<0>{});
  else
    epilogue_write(IC<1>{});
}

}  // namespace

int mixed_points_per_block(int mode) { return (mode == MODE_RHO || mode == MODE_GRAD) ? TC : TC / 4; }

void launch_contract_mixed(int mode, const uint16_t* Pp, int nbasis_pad, const float* ao, int64_t ld, double* part,
                           hipStream_t s) {
  const int ppb = mixed_points_per_block(mode);
  const int nrt = nbasis_pad / TR;
  const int64_t nptile = ld / ppb;
  const dim3 block(256), grid((unsigned)(((nptile + 7) / 8) * 8 * nrt));
  switch (mode) {
    case MODE_RHO:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_RHO>, grid, block, 0, s, Pp, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
    case MODE_GRAD:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_GRAD>, grid, block, 0, s, Pp, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
    case MODE_LAP:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_LAP>, grid, block, 0, s, Pp, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
    default:
      hipLaunchKernelGGL(contract_mixed_kernel<MODE_HESS>, grid, block, 0, s, Pp, nbasis_pad, ao, ld, part, nrt, nptile);
      break;
  }
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
