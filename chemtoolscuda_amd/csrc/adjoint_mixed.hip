// Mixed-precision adjoint of rho (no reference counterpart; the backward of
// the autograd wrapper when the forward runs in mixed precision):
//
//   dL/dP_ij = sum_g w_g phi_i(g) phi_j(g)                 (w_g = dL/drho_g)
//   dL/dR_A  = -2 sum_{i on A} sum_g w_g grad phi_i(g) (P phi(g))_i
//
// with FP32 AOs (row quads, computed in FP64 and rounded once) and the
// split-bf16 matrix cores, FP64 everywhere the sums leave a tile:
//   * gram_mixed: one workgroup per (upper-triangle 128 x 128 tile of G,
//     slice of points); both operands (phi rows of the two blocks, the second
//     scaled by w) are split into bf16 pieces on the way into LDS, the six
//     products go to separate hi*hi / small FP32 accumulators, and each slice
//     writes its FP64 partial tile (and its transpose); gram_reduce adds the
//     slices in a fixed order.
//   * contract_adjoint_mixed: Y = P phi with the shared tile engine
//     (mixed_tile.h), its products accumulated in FP32 pairs (hi*hi apart
//     from the five smaller ones: no FP64 work per k-step, which would
//     outweigh the matrix-core work of this 64-column tile); the epilogue
//     sums w_g d_k phi_i(g) Y_i(g) over the workgroup's points per row i in
//     FP64, over several point tiles per workgroup, then adds the row totals
//     into T[i][k] with FP64 atomics.

#include "mixed_tile.h"

namespace ct {

namespace {

using namespace mixed;

// ------------------------------------------------------------------ dL/dR
constexpr int kAdjTC = 64;  // points per tile

__global__ __launch_bounds__(256, 2) void contract_adjoint_mixed_kernel(const u4* __restrict__ Pf, int mp, int m,
                                                                        const float* __restrict__ ao, int64_t ld,
                                                                        const double* __restrict__ w, int64_t npts,
                                                                        int nrt, int64_t nptile, int tiles_per_wg,
                                                                        double* __restrict__ T) {
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][NPL][kAdjTC][LROW];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, q = lane >> 4;
  // XCD-aware order as in the forward contraction: the nrt row blocks of one
  // group of point tiles run on the same XCD.
  const int g = blockIdx.x;
  const int xcd = g & 7, slot = g >> 3;
  const int rt = slot % nrt;
  const int64_t grp = (int64_t)(slot / nrt) * 8 + xcd;
  const int64_t t0 = grp * tiles_per_wg;
  if (t0 >= nptile) return;
  const int64_t t1 = t0 + tiles_per_wg < nptile ? t0 + tiles_per_wg : nptile;
  const int i0 = rt * TR;
  const int64_t mq = mp / 4;
  const int64_t quad0 = (i0 + RW * wv) / 4 + q;
  // P's rows and columns beyond M are zero: K stops at the last k-step with a
  // real function, and a wave whose rows are all padding adds nothing to T
  const bool live = i0 + RW * wv < m;
  const int kend = (m + KB - 1) / KB;

  double rs[NRB][4][3];  // row sums: rows i0 + 32 wv + 16 rb + 4 q + j, axis k
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) rs[rb][j][k] = 0.0;

  for (int64_t pt = t0; pt < t1; ++pt) {
    const int64_t pbase = pt * kAdjTC;
    YTile<kAdjTC, 0, 0> yt;
    y_tile<kAdjTC, 0, 0, false>(Pf, mp, mp, ld, i0, tid, ao + (pbase + (tid & (kAdjTC - 1))) * 4, Bs, yt, 0,
                                 nullptr, 0, nullptr, 0, 0, kend);
    if (!live) continue;
#pragma unroll
    for (int cb = 0; cb < kAdjTC / 16; ++cb) {
      const int64_t p = pbase + 16 * cb + r;
      const double wp = p < npts ? w[p] : 0.0;
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
        f4 dphi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          dphi[k] = *reinterpret_cast<const f4*>(ao + (((int64_t)(1 + k) * mq + quad0 + 4 * rb) * ld + p) * 4);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double wy = wp * yt.y(rb, cb, j);
#pragma unroll
          for (int k = 0; k < 3; ++k) rs[rb][j][k] += wy * (double)dphi[k][j];
        }
      }
    }
  }
  // sum over the 16 lanes (points) of each row group, then one atomic per row and axis
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double v = rs[rb][j][k];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        rs[rb][j][k] = v;
      }
  if (r == 0 && live) {
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t i = i0 + RW * wv + 16 * rb + 4 * q + j;
#pragma unroll
        for (int k = 0; k < 3; ++k) atomicAdd(&T[i * 3 + k], rs[rb][j][k]);
      }
  }
}

// ------------------------------------------------------------------ dL/dP
// LDS: one stage of both operands, [operand][piece][row][k], 80-byte rows.
__global__ __launch_bounds__(256, 2) void gram_mixed_kernel(const float* __restrict__ ao, int mp, int m, int64_t ld,
                                                            const double* __restrict__ w, int64_t npts,
                                                            const int2* __restrict__ tiles, int ntri,
                                                            int64_t slice_len, double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) uint16_t S[2][NPL][TR][LROW];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, q = lane >> 4;
  const int tile = blockIdx.x % ntri;
  const int slice = blockIdx.x / ntri;
  const int I = tiles[tile].x * TR, J = tiles[tile].y * TR;
  const int64_t pbeg = (int64_t)slice * slice_len;
  const int64_t pend = pbeg + slice_len < npts ? pbeg + slice_len : npts;

  // staging units of this thread: row quad rq, point quad pq (4 rows x 4
  // points) of both operands, phi rows of block I (op 0) and w phi rows of
  // block J (op 1): every thread scales the same number of values by w
  const int rq = tid >> 3, pq = tid & 7;
  const float* src[2] = {ao + ((int64_t)(I / 4 + rq) * ld) * 4, ao + ((int64_t)(J / 4 + rq) * ld) * 4};
  // rows at or past the last real k-step are not in the AO buffer: zeros
  const int kend_rows = (m + KB - 1) / KB * KB;
  const bool real[2] = {I + 4 * rq < kend_rows, J + 4 * rq < kend_rows};
  f4 st[2][4];
  double ws[4];
  auto load = [&](int64_t p0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t p = p0 + 4 * pq + e;
      const bool ok = p < pend;
#pragma unroll
      for (int op = 0; op < 2; ++op)
        st[op][e] = ok && real[op] ? *reinterpret_cast<const f4*>(src[op] + p * 4) : f4{0.f, 0.f, 0.f, 0.f};
      ws[e] = ok ? w[p] : 0.0;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int op = 0; op < 2; ++op)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        uint32_t vh[2], vm[2], vl[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          float x0 = st[op][2 * e][j], x1 = st[op][2 * e + 1][j];
          if (op == 1) {
            x0 = (float)(ws[2 * e] * (double)x0);
            x1 = (float)(ws[2 * e + 1] * (double)x1);
          }
          const Split2 sp = split2(x0, x1);
          vh[e] = sp.h;
          vm[e] = sp.m;
          vl[e] = sp.l;
        }
        typedef uint32_t u2 __attribute__((ext_vector_type(2)));
        *reinterpret_cast<u2*>(&S[op][0][4 * rq + j][4 * pq]) = u2{vh[0], vh[1]};
        *reinterpret_cast<u2*>(&S[op][1][4 * rq + j][4 * pq]) = u2{vm[0], vm[1]};
        *reinterpret_cast<u2*>(&S[op][2][4 * rq + j][4 * pq]) = u2{vl[0], vl[1]};
      }
  };

  f4 hh[NRB][8], sm[NRB][8];
#pragma unroll
  for (int a = 0; a < NRB; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) hh[a][b] = sm[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  load(pbeg);
  for (int64_t p0 = pbeg; p0 < pend; p0 += KB) {
    store();
    __syncthreads();
    if (p0 + KB < pend) load(p0 + KB);
    bf16x8 a[NPL][NRB];
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
        a[pl][rb] = as_bf(*reinterpret_cast<const u4*>(&S[0][pl][RW * wv + 16 * rb + r][8 * q]));
#pragma unroll
    for (int cp = 0; cp < 4; ++cp) {
      bf16x8 bf[2][NPL];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl)
          bf[c][pl] = as_bf(*reinterpret_cast<const u4*>(&S[1][pl][16 * (2 * cp + c) + r][8 * q]));
      constexpr int pa[5] = {2, 1, 0, 1, 0};
      constexpr int pb[5] = {0, 1, 2, 0, 1};
#pragma unroll
      for (int pr = 0; pr < 5; ++pr)
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            sm[rb][2 * cp + c] =
                __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[pa[pr]][rb], bf[c][pb[pr]], sm[rb][2 * cp + c], 0, 0, 0);
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int c = 0; c < 2; ++c)
          hh[rb][2 * cp + c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][rb], bf[c][0], hh[rb][2 * cp + c], 0, 0, 0);
    }
    __syncthreads();
  }
  // partial tile of this slice (and its transpose for an off-diagonal tile)
  double* out = part + (int64_t)slice * mp * mp;
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int cb = 0; cb < 8; ++cb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = I + RW * wv + 16 * rb + 4 * q + j;
        const int jj = J + 16 * cb + r;
        const double v = (double)hh[rb][cb][j] + (double)sm[rb][cb][j];
        out[(int64_t)i * mp + jj] = v;
        if (I != J) out[(int64_t)jj * mp + i] = v;
      }
}

}  // namespace

int adjoint_mixed_tiles_per_wg(int64_t npts, int mp) {
  const int64_t nptile = (npts + kAdjTC - 1) / kAdjTC;
  const int nrt = mp / TR;
  // about 4096 workgroups, at least one point tile each
  const int64_t t = (nptile * nrt + 4095) / 4096;
  return (int)(t < 1 ? 1 : t);
}

void launch_contract_adjoint_mixed(const uint16_t* Pf, int mp, int m, const float* ao, int64_t ld, const double* w,
                                   int64_t npts, double* T, hipStream_t s) {
  CT_REQUIRE(mp % TR == 0 && m > 0 && m <= mp, "padded basis size must be a multiple of 128");
  const int nrt = mp / TR;
  const int64_t nptile = (npts + kAdjTC - 1) / kAdjTC;  // ld >= nptile * 64: ld is a multiple of 256
  const int tpw = adjoint_mixed_tiles_per_wg(npts, mp);
  const int64_t ngrp = (nptile + tpw - 1) / tpw;
  const dim3 grid((unsigned)(((ngrp + 7) / 8) * 8 * nrt)), block(256);
  hipLaunchKernelGGL(contract_adjoint_mixed_kernel, grid, block, 0, s, reinterpret_cast<const u4*>(Pf), mp, m, ao, ld,
                     w, npts, nrt, nptile, tpw, T);
  CT_HIP_CHECK(hipGetLastError());
}

void gram_mixed_plan(int64_t npts, int mp, int* nslice, int64_t* slice_len, int max_slices) {
  const int nb = mp / TR;
  const int ntri = nb * (nb + 1) / 2;
  // The (tile, slice) workgroups run two per CU in rounds: the call takes
  // (rounds) x (points per slice).  Of the slice counts up to 64 (slices of
  // at least 1024 points) the one with the least of that product, the
  // smallest on a tie (fewer partial tiles to add): 10 tiles (M = 416) on 256
  // CUs take 51 slices, one round, where 64 slices ran a second round for
  // 128 of 640 workgroups.
  int dev = 0, cu = 0;
  CT_HIP_CHECK(hipGetDevice(&dev));
  CT_HIP_CHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t slots = 2 * (int64_t)(cu > 0 ? cu : 1);
  int64_t cap = (npts + 1023) / 1024;
  if (cap > 64) cap = 64;
  if (cap > max_slices) cap = max_slices;
  if (cap < 1) cap = 1;
  int64_t best = 1;
  double best_cost = 0.0;
  for (int64_t ns = 1; ns <= cap; ++ns) {
    const int64_t len = ((npts + ns - 1) / ns + KB - 1) / KB * KB;
    const int64_t used = (npts + len - 1) / len;
    const int64_t rounds = (ntri * used + slots - 1) / slots;
    const double cost = (double)rounds * (double)len;
    if (ns == 1 || cost < best_cost) {
      best = ns;
      best_cost = cost;
    }
  }
  int64_t len = ((npts + best - 1) / best + KB - 1) / KB * KB;
  *slice_len = len;
  *nslice = (int)((npts + len - 1) / len);
}

void launch_gram_mixed(const float* ao, int mp, int m, int64_t ld, const double* w, int64_t npts, const int32_t* tiles,
                       int nslice, int64_t slice_len, double* part, hipStream_t s) {
  const int nb = mp / TR;
  const int ntri = nb * (nb + 1) / 2;
  hipLaunchKernelGGL(gram_mixed_kernel, dim3((unsigned)(ntri * nslice)), dim3(256), 0, s, ao, mp, m, ld, w, npts,
                     reinterpret_cast<const int2*>(tiles), ntri, slice_len, part);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
