// Split-bf16 tile engine shared by the mixed-precision contraction and the
// mixed adjoint: Y = A[i0:i0+128, :] * B for TC GEMM columns, FP32-quality,
// on v_mfma_f32_16x16x32_bf16 (A = P, or the transposed factor of P).
//
// P and the FP32 columns are split exactly into three bf16 pieces
// (x = hi + mid + lo), and Y is the sum of the six products with the
// largest terms (hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid; the dropped
// ones are <= 2^-23 of |P||B|), accumulated as described at YTile.  An
// optional seventh product
// |P_hi| * |B_hi| (sign-bit masks of the hi pieces) accumulates the
// magnitude sum that scales the rounding errors.
//
// One workgroup = 4 waves, two workgroups per CU; wave w owns rows
// 32w..32w+31 of the 128-row block and all TC columns.  K advances 32 per
// step:
//   A (P pieces): stored on the host in MFMA fragment order
//     (mixed_fragment_order), so each wave loads its 16 x 32 fragments
//     straight into registers with one coalesced 1 KiB load per (piece,
//     16-row block); the next step's fragments are loaded after the last use
//     of the current ones (the other workgroup on the CU covers the wait);
//   B: FP32 row quads (ao_quad_index layout, or ao_packed_index for the
//     screened factor route: quad and k-step strides given) read one k-step
//     ahead, split on the way into double-buffered LDS, column-major with
//     80-byte rows (conflict-free 16-byte fragment reads).
//
// MFMA 16x16x32 bf16 lane layout (lane l, r = l % 16, q = l / 16), measured:
//   A: A[r][8q + e], B: B[8q + e][r] (e = 0..7), D: D[4q + j][r] (j = 0..3).
#pragma once

#include <cstddef>

#include "ct_internal.h"

namespace ct {
namespace mixed {

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

// Exact three-piece split of two FP32 values (hardware RNE pack,
// v_cvt_pk_bf16_f32): the packed hi, mid and lo pairs.  The pieces are
// unpacked from the packed pair by a shift (low half) and a mask (high
// half): a bf16 is the top half of the FP32 of the same value.  Three packs,
// four bit operations and four subtractions per pair.
struct Split2 {
  uint32_t h, m, l;
};
__device__ __forceinline__ uint32_t pack_bf16(float x0, float x1) {
  bf16x2 v;
  v[0] = (__bf16)x0;
  v[1] = (__bf16)x1;
  uint32_t u = __builtin_bit_cast(uint32_t, v);
  // opaque to the optimiser, which would otherwise rebuild the low half as a
  // second conversion instead of one shift
  asm("" : "+v"(u));
  return u;
}
__device__ __forceinline__ Split2 split2(float x0, float x1) {
  const uint32_t h = pack_bf16(x0, x1);
  const float r0 = x0 - __uint_as_float(h << 16), r1 = x1 - __uint_as_float(h & 0xFFFF0000u);
  const uint32_t m = pack_bf16(r0, r1);
  const float s0 = r0 - __uint_as_float(m << 16), s1 = r1 - __uint_as_float(m & 0xFFFF0000u);
  return Split2{h, m, pack_bf16(s0, s1)};
}

// Accumulation of the six products, per 16-column block, two schemes:
//   FP64 sums (the first NF column blocks): per k-step and column block the
//     six products are summed from zero in FP32 (their 32-term MFMA sums and
//     five adds at the size of one k-step's contribution) and the chunk is
//     added to FP64 running sums: the only FP32 rounding at full magnitude is
//     gone (error ~2.4u |P||B|, close to the 2.1u floor of rounding the
//     inputs to FP32); 80 VALU operations per k-step and 16-row x 16-column
//     pair of blocks (conversion and FP64 add).
//   FP32 pair (the other blocks): hi*hi in one FP32 accumulator, the five
//     smaller products in another (~3.7u |P||B|), no VALU work per k-step.
// NF = TC / 16: all blocks FP64; NF = 0: none (the Hessian mode's larger
// epilogue).  Both cost the same registers.
template <int TC, int NBC, int NF = TC / 16>
struct YTile {
  static constexpr int NCB = TC / 16;
  static constexpr int N32 = NCB - NF;
  static_assert(NF >= 0 && NF <= NCB, "FP64 blocks");
  double yd[NRB][NF > 0 ? NF : 1][4];   // FP64 running sums of blocks 0..NF-1
  f4 hh[NRB][N32 > 0 ? N32 : 1];        // FP32 pair of blocks NF..NCB-1: hi*hi
  f4 sm[NRB][N32 > 0 ? N32 : 1];        //   and the five smaller products
  f4 bound[NRB][NBC > 0 ? NBC : 1];     // |P_hi| |B_hi| for column blocks 0..NBC-1
  __device__ __forceinline__ double y(int rb, int cb, int j) const {
    if (cb < NF) return yd[rb][cb < NF ? cb : 0][j];
    const int c = cb >= NF ? cb - NF : 0;
    return (double)hh[rb][c][j] + (double)sm[rb][c][j];
  }
};

// One k-step of the tile engine: the six split products of A (areg) with
// the LDS buffer s, column blocks in pairs (2 x NRB independent accumulators
// per product, so dependent MFMAs are four issues apart), added to the
// accumulators of t, and the bound product on the first NBC column blocks.
// dbl (TRI only): add the products twice.
template <int TC, int NBC, int NF, bool TRI>
__device__ __forceinline__ void tile_compute(const uint16_t (&Bs)[2][NPL][TC][LROW], int s, bool dbl,
                                             const u4 (&areg)[NPL][NRB], YTile<TC, NBC, NF>& t, int r, int q) {
  constexpr int NCB = TC / 16;
  const u4 absmask = u4{0x7FFF7FFFu, 0x7FFF7FFFu, 0x7FFF7FFFu, 0x7FFF7FFFu};
  const u4 exp1 = u4{0x00800080u, 0x00800080u, 0x00800080u, 0x00800080u};  // +1 in both bf16 exponents
  // products smallest first, hi*hi last
  constexpr int pa[6] = {2, 1, 0, 1, 0, 0};
  constexpr int pb[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
  for (int cp = 0; cp < (NCB + 1) / 2; ++cp) {
    // an odd NCB leaves the second block of the last pair empty (c < nc)
    const int nc = 2 * cp + 1 < NCB ? 2 : 1;
    bf16x8 bf[2][NPL];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl)
        if (c < nc) bf[c][pl] = as_bf(*reinterpret_cast<const u4*>(&Bs[s][pl][16 * (2 * cp + c) + r][8 * q]));
    f4 tc[NRB][2];
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int c = 0; c < 2; ++c) tc[rb][c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int pr = 0; pr < 6; ++pr)
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int cb = 2 * cp + c;
          const int c32 = cb >= NF ? cb - NF : 0;
          if (c >= nc) continue;
          if (cb < NF)
            tc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(areg[pa[pr]][rb]), bf[c][pb[pr]], tc[rb][c], 0,
                                                                0, 0);
          else if (pr < 5)
            t.sm[rb][c32] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(areg[pa[pr]][rb]), bf[c][pb[pr]],
                                                                    t.sm[rb][c32], 0, 0, 0);
          else
            t.hh[rb][c32] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(areg[0][rb]), bf[c][0], t.hh[rb][c32], 0,
                                                                    0, 0);
        }
    const double sc = (TRI && dbl) ? 2.0 : 1.0;
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int cb = 2 * cp + c;
        if (c < nc && cb < NF)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            t.yd[rb][cb < NF ? cb : 0][j] = fma((double)tc[rb][c][j], sc, t.yd[rb][cb < NF ? cb : 0][j]);
      }
    if constexpr (NBC > 0) {
      if (2 * cp < NBC) {
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) {
          const u4 pabs = (TRI && dbl) ? (areg[0][rb] & absmask) + exp1 : areg[0][rb] & absmask;
#pragma unroll
          for (int c = 0; c < 2; ++c)
            if (2 * cp + c < NBC)
              t.bound[rb][(2 * cp + c) % NBC] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                  as_bf(pabs), as_bf(__builtin_bit_cast(u4, bf[c][0]) & absmask), t.bound[rb][(2 * cp + c) % NBC],
                  0, 0, 0);
        }
      }
    }
  }
}

// Zeroes the accumulators of a tile.
template <int TC, int NBC, int NF>
__device__ __forceinline__ void tile_zero(YTile<TC, NBC, NF>& t) {
  using T = YTile<TC, NBC, NF>;
#pragma unroll
  for (int a = 0; a < NRB; ++a) {
#pragma unroll
    for (int b = 0; b < (NF > 0 ? NF : 1); ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j) t.yd[a][b][j] = 0.0;
#pragma unroll
    for (int b = 0; b < (T::N32 > 0 ? T::N32 : 1); ++b) t.hh[a][b] = t.sm[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < (NBC > 0 ? NBC : 1); ++b) t.bound[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  }
}

// Splits this thread's B units (FP32 row-quad pairs in rbuf) into bf16
// pieces and stores them into LDS buffer s.
template <int TC, int NU>
__device__ __forceinline__ void tile_store_b(uint16_t (&Bs)[2][NPL][TC][LROW], int s, const f4 (&rbuf)[NU][2],
                                             const int (&ucol)[NU], const int (&uoct)[NU], const bool (&uok)[NU]) {
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    if (!uok[u]) continue;
    u4 vh, vm, vl;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const Split2 sp = split2(rbuf[u][e >> 1][2 * (e & 1)], rbuf[u][e >> 1][2 * (e & 1) + 1]);
      vh[e] = sp.h;
      vm[e] = sp.m;
      vl[e] = sp.l;
    }
    const int k = 8 * uoct[u];
    *reinterpret_cast<u4*>(&Bs[s][0][ucol[u]][k]) = vh;
    *reinterpret_cast<u4*>(&Bs[s][1][ucol[u]][k]) = vm;
    *reinterpret_cast<u4*>(&Bs[s][2][ucol[u]][k]) = vl;
  }
}

// Computes the tile.  bsrc: this thread's B column at row quad 0, i.e. the
// column tid % TC; its row quad k lies at bsrc + k * ld * 4 (row-quad
// layout), or, with qstride and kstride (tile-packed layout), quad q of
// k-step kb at bsrc + kb * kstride + q * qstride.  Bs: the workgroup's LDS.
// Ends with a barrier (Bs free for reuse).
//
// K runs over the k-steps k_begin .. mp/32 - 1 (or .. k_end - 1: the rest of
// A's columns are zero).  TRI (symmetric P, columns
// used only in quadratic forms x^T P x): k_begin is the row block's diagonal
// k-step and wave w, whose 32 rows are k-step kd = k_begin + w, uses the
// k-steps from kd on, adding the products of k-steps right of kd twice.  By
// x^T P x = sum_I x_I^T (P_II x_I + 2 sum_{J>I} P_IJ x_J) over 32-row blocks I,
// summing the tile's dots over all row blocks gives the full quadratic form
// for about half the work.  The bound product is treated the same way
// (|x|^T |P| |x| is a quadratic form too), its doubling applied exactly to
// the |P_hi| operand (bf16 exponent + 1).
//
// TC a multiple of 64: bsrc is this thread's column (tid % TC).  Other TC
// (a multiple of 16): bcol(c) gives column c's row quad 0, and the TC x 4
// (column, octet) units of a k-step go round-robin to the 256 threads.
template <int TC, int NBC, int NF, bool TRI, bool LIST = false, class BCol = std::nullptr_t>
__device__ __forceinline__ void y_tile(const u4* __restrict__ Pf, int rows, int mp, int64_t ld, int i0, int tid,
                                       const float* __restrict__ bsrc, uint16_t (&Bs)[2][NPL][TC][LROW],
                                       YTile<TC, NBC, NF>& t, int k_begin, const int32_t* klist = nullptr,
                                       int kcount = 0, BCol bcol = nullptr, int64_t qstride = 0,
                                       int64_t kstride = 0, int k_end = -1) {
  static_assert(!TRI || NF == TC / 16, "the triangle needs the FP64 running sums");
  static_assert(!(TRI && LIST), "the triangle walks all k-steps from the diagonal");
  static_assert(TC % 16 == 0, "column blocks of 16");
  constexpr int NCB = TC / 16;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, q = lane >> 4;
  const int nk = mp / KB;

  // A (rows x mp: P, or the transposed factor of P): fragment (piece pl,
  // 16-row block b16, k-step kb) is the 64 x 16 bytes at
  // Pf[((pl * rows/16 + b16) * nk + kb) * 64 + lane]: a wave-uniform
  // fragment address (scalar arithmetic per k-step) plus the lane's 16-byte
  // slot.
  const u4* awave = Pf + (int64_t)((i0 + RW * w) / 16) * nk * 64;
  const int64_t aplane = (int64_t)(rows / 16) * nk * 64;
  u4 areg[NPL][NRB];
  auto load_a = [&](int kb) {
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
        const u4* frag = awave + (pl * aplane + ((int64_t)rb * nk + kb) * 64);
        areg[pl][rb] = frag[lane];
      }
  };

  // B: (column, 8-row octet) units, unit u -> column u % TC, octet u / TC
  // (an octet is two row quads); thread tid takes units tid + 256 i.
  constexpr bool POW = TC % 64 == 0;
  constexpr int NU = POW ? TC / 64 : (TC * 4 + 255) / 256;
  const float* ubase[NU];
  int ucol[NU], uoct[NU];
  bool uok[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    if constexpr (POW) {
      ucol[u] = tid & (TC - 1);
      uoct[u] = tid / TC + (256 / TC) * u;
      ubase[u] = bsrc;
      uok[u] = true;
    } else {
      const int uu = tid + 256 * u;
      uok[u] = uu < TC * 4;  // per wave: whole waves drop out of the last round
      const int uc = uok[u] ? uu : tid;
      ucol[u] = uc % TC;
      uoct[u] = uc / TC;
      ubase[u] = bcol(ucol[u]);
    }
  }
  // Every thread loads all NU units, a unit past the last one (!uok) from its
  // first unit's address (a cache hit; only the store skips it): with the
  // loads unconditional the compiler counts them statically and waits for
  // the next step's B only where it is split into LDS, a whole k-step after
  // the load.  Masked loads make the count dynamic, and the wait for A at
  // the top of each k-step then also drains the B prefetch.
  // The lane's row quads of k-step 0 (per unit and half); k-step kb lies a
  // wave-uniform 32 ld floats per step further.
  // (qstride, kstride: floats between consecutive row quads and k-steps of
  // a column; 0: the row-quad layout's 4 ld and 32 ld)
  const int64_t qs = qstride ? qstride : ld * 4, ks = kstride ? kstride : 32 * ld;
  const float* bq[NU][2];
#pragma unroll
  for (int u = 0; u < NU; ++u)
#pragma unroll
    for (int h = 0; h < 2; ++h) bq[u][h] = ubase[u] + (int64_t)(2 * uoct[u] + h) * qs;
  f4 rbuf[NU][2];
  auto load_b = [&](int kb) {
    const int64_t koff = (int64_t)kb * ks;
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) rbuf[u][h] = *reinterpret_cast<const f4*>(bq[u][h] + koff);
  };
  auto store_b = [&](int s) { tile_store_b<TC, NU>(Bs, s, rbuf, ucol, uoct, uok); };

  tile_zero(t);

  auto compute = [&](int s, bool dbl) { tile_compute<TC, NBC, NF, TRI>(Bs, s, dbl, areg, t, r, q); };

  // The loads run unconditionally, clamped to the last k-step on the final
  // iterations (those values are unused; the LDS buffer written then is the
  // idle one): conditional loads would keep two copies of the B registers
  // alive across the k-step.
  // The k-steps: LIST, klist[0 .. kcount) (the k-steps a point tile needs,
  // increasing; list and length are the same for the whole workgroup, so the
  // entries are scalar loads, fetched three steps ahead); else
  // k_begin .. nk - 1.
  const int kn = LIST ? kcount : (k_end >= 0 ? k_end : nk) - k_begin;
  auto kat = [&](int i) {
    if constexpr (LIST)
      return klist[i];
    else
      return k_begin + i;
  };
  if (kn <= 0) return;
  const int kd = k_begin + w;  // TRI: this wave's diagonal k-step (wave-uniform)
  int kcur = kat(0), knext = kat(kn > 1 ? 1 : 0), knext2 = kat(kn > 2 ? 2 : kn - 1);
  load_b(kcur);
  load_a(TRI ? kd : kcur);
  store_b(0);
  load_b(knext);
  __syncthreads();
  for (int i = 0; i < kn; ++i) {
    const int kb = kcur;
    const int k3 = kat(i + 3 < kn ? i + 3 : kn - 1);
    const int s = i & 1;
    if (!TRI || kb >= kd) compute(s, kb > kd);
    if (!TRI || knext > kd) load_a(knext);  // A(kd) is loaded before the loop
    store_b(s ^ 1);
    load_b(knext2);
    __syncthreads();
    kcur = knext;
    knext = knext2;
    knext2 = k3;
  }
}


}  // namespace mixed
}  // namespace ct
