// Mixed-precision contraction: FP32 AO values, P and the AOs split into three
// bf16 pieces each (x = hi + mid + lo exactly for FP32 x), six bf16 MFMA
// products per tile (lo*hi, mid*mid, hi*lo, mid*hi, hi*mid, hi*hi) summed per
// k-step and accumulated in FP32 -- FP32-quality Y = P [phi | d phi] at bf16
// matrix-core rates -- and FP64 per-point dot products in the epilogue.
//
// A seventh product |P_hi| * |phi_hi| (bf16, FP32 accumulation; bf16(|x|) =
// |bf16(x)|, so both are sign-bit masks of the hi pieces) gives the
// per-point magnitude B = sum_i |phi_i| (|P| |phi|)_i that scales every
// rounding error of rho; the finaliser flags points whose estimate
// gamma * B exceeds the tolerance for FP64 re-evaluation.
//
// Decomposition: one workgroup (4 waves, two workgroups per CU) owns a
// 128-row block of P and 128 GEMM columns: 128 points of one column type
// (rho / grad modes) or 32 points x {phi, d_x phi, d_y phi, d_z phi} (lap /
// Hessian modes).  Wave w owns rows 32w..32w+31 of the block and all 128
// columns (2 x 8 MFMA 16x16x32 blocks).  K advances 32 at a time:
//   A (P pieces): stored on the host in MFMA fragment order, so each wave
//     loads its 16 x 32 fragments straight into registers with one fully
//     coalesced 1 KiB load per (piece, 16-row block); prefetched two
//     k-steps ahead, never staged in LDS;
//   B (AOs): FP32 row quads read coalesced from HBM/L2 one k-step ahead,
//     split into bf16 pieces on the way into double-buffered LDS, stored
//     column-major (k contiguous, 80-byte rows: conflict-free 16-byte
//     fragment reads).
// 61 KiB of LDS and at most 256 VGPRs per wave let two workgroups share a
// CU, so one workgroup's epilogue and memory waits overlap the other's MFMAs.
// The epilogue dots each wave's rows with the matching AO row quads (one
// 16-byte load per quad) in FP64, reduces over lane groups and the four
// waves in a fixed order, and writes one partial per (row block, dot,
// point); the finaliser sums the row blocks.
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

constexpr int TR = 128;       // rows of P per workgroup
constexpr int KB = 32;        // k per stage (= MFMA depth)
constexpr int LROW = KB + 8;  // bf16 per LDS row (80 bytes: staggers banks)
constexpr int NPL = 3;        // pieces: hi, mid, lo
constexpr int NW = 4;         // waves per workgroup
constexpr int RW = TR / NW;   // rows per wave
constexpr int NRB = RW / 16;  // 16-row MFMA blocks per wave

__device__ __forceinline__ bf16x8 as_bf(u4 v) { return __builtin_bit_cast(bf16x8, v); }

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
  const int nk = mp / KB;   // even: mp is a multiple of 128
  const int64_t mq = mp / 4;  // row quads per AO array

  // A: fragment (piece pl, 16-row block b16, k-step kb) is the 64 x 16 bytes
  // at Pf[((pl * mp/16 + b16) * nk + kb) * 64 + lane].
  const u4* abase = Pf + (int64_t)((i0 + RW * w) / 16) * nk * 64 + lane;
  const int64_t aplane = (int64_t)(mp / 16) * nk * 64;
  u4 areg[NPL][NRB];
  auto load_a = [&](int kb) {
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) areg[pl][rb] = abase[pl * aplane + ((int64_t)rb * nk + kb) * 64];
  };

  // B: TC/64 (column, 8-row octet) units per thread, unit u = tid + 256 u
  // -> column u % TC, octet u / TC; an octet is two row quads.
  constexpr int NU = TC / 64;
  const int bc = tid & (TC - 1), bo = tid / TC;
  const float* bsrc = ao + ((int64_t)C::type(bc) * mq * ld + pbase + C::point(bc)) * 4;
  f4 rbuf[NU][2];
  auto load_b = [&](int kb) {
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        rbuf[u][h] =
            *reinterpret_cast<const f4*>(bsrc + (int64_t)(8 * kb + 2 * (bo + (256 / TC) * u) + h) * ld * 4);
  };
  // exact three-piece split with the hardware RNE pack (v_cvt_pk_bf16_f32)
  auto store_b = [&](int s) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      u4 vh, vm, vl;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float x0 = rbuf[u][e >> 1][2 * (e & 1)], x1 = rbuf[u][e >> 1][2 * (e & 1) + 1];
        bf16x2 hh, mm, ll;
        hh[0] = (__bf16)x0;
        hh[1] = (__bf16)x1;
        const float r0 = x0 - (float)hh[0], r1 = x1 - (float)hh[1];
        mm[0] = (__bf16)r0;
        mm[1] = (__bf16)r1;
        ll[0] = (__bf16)(r0 - (float)mm[0]);
        ll[1] = (__bf16)(r1 - (float)mm[1]);
        vh[e] = __builtin_bit_cast(uint32_t, hh);
        vm[e] = __builtin_bit_cast(uint32_t, mm);
        vl[e] = __builtin_bit_cast(uint32_t, ll);
      }
      const int k = 8 * (bo + (256 / TC) * u);
      *reinterpret_cast<u4*>(&Bs[s][0][bc][k]) = vh;
      *reinterpret_cast<u4*>(&Bs[s][1][bc][k]) = vm;
      *reinterpret_cast<u4*>(&Bs[s][2][bc][k]) = vl;
    }
  };

  // acc: hi*hi products; accs: the five smaller products (2^-8 of the size
  // and less).  Kept apart, the FP32 rounding of the running sums happens once
  // per k-step at the full magnitude instead of six times (the largest error
  // source of the path); Y = acc + accs is formed exactly in FP64.
  f4 acc[NRB][NCB], accs[NRB][NCB], accb[NRB][NBC];
#pragma unroll
  for (int a = 0; a < NRB; ++a) {
#pragma unroll
    for (int b = 0; b < NCB; ++b) acc[a][b] = accs[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < NBC; ++b) accb[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  }

  const u4 absmask = u4{0x7FFF7FFFu, 0x7FFF7FFFu, 0x7FFF7FFFu, 0x7FFF7FFFu};
  // Column blocks in pairs: 2 x NRB independent accumulators per product, so
  // dependent MFMAs are four issues apart.
  auto compute = [&](int s) {
#pragma unroll
    for (int cp = 0; cp < NCB / 2; ++cp) {
      bf16x8 bf[2][NPL];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl)
          bf[c][pl] = as_bf(*reinterpret_cast<const u4*>(&Bs[s][pl][16 * (2 * cp + c) + r][8 * q]));
      constexpr int pa[5] = {2, 1, 0, 1, 0};
      constexpr int pb[5] = {0, 1, 2, 0, 1};
#pragma unroll
      for (int pr = 0; pr < 5; ++pr)
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            accs[rb][2 * cp + c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(areg[pa[pr]][rb]), bf[c][pb[pr]],
                                                                            accs[rb][2 * cp + c], 0, 0, 0);
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int c = 0; c < 2; ++c)
          acc[rb][2 * cp + c] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(areg[0][rb]), bf[c][0], acc[rb][2 * cp + c], 0, 0, 0);
      if (2 * cp < NBC) {
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            accb[rb][(2 * cp + c) % NBC] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                as_bf(areg[0][rb] & absmask), as_bf(__builtin_bit_cast(u4, bf[c][0]) & absmask),
                accb[rb][(2 * cp + c) % NBC], 0, 0, 0);
      }
    }
  };

  // Pipeline: B global loads run one k-step ahead of their LDS store; the A
  // fragments of the next k-step are loaded after the last use of the current
  // ones (the other workgroup on the CU covers the wait).
  load_b(0);
  load_a(0);
  store_b(0);
  load_b(1);
  __syncthreads();
  for (int kb = 0; kb < nk; ++kb) {
    const int s = kb & 1;
    compute(s);
    if (kb + 1 < nk) {
      load_a(kb + 1);
      store_b(s ^ 1);
    }
    if (kb + 2 < nk) load_b(kb + 2);
    __syncthreads();
  }

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
        const double y = (double)acc[rb][CB][j] + (double)accs[rb][CB][j];
#pragma unroll
        for (int k = 0; k < nd; ++k) d[k] += (double)v[k][j] * y;
        if constexpr (bnd) d[kMaxDots - 1] += fabs((double)v[0][j]) * (double)accb[rb][CB % NBC][j];
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
  __syncthreads();  // all waves are done reading Bs
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
