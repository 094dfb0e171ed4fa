// FP64 contraction on the matrix cores with fused per-point epilogues.
//
// The reference forms rho = sum_i phi_i (P phi)_i with a cuBLAS DGEMM into an
// (N x M) scratch matrix, a Hadamard kernel and a DGEMV against a vector of
// ones, once per derivative column -- 8 DGEMMs and 6 AO passes for
// rho + grad + lap (src/evaluate_density.cu:780-823,
// src/evaluate_gradient.cu:1318-1375, src/evaluate_laplacian.cu:479-821,
// src/evaluate_hessian.cu:1906-2230).  Here one kernel computes
// Y = P [phi | d_x phi | d_y phi | d_z phi] tile by tile with
// v_mfma_f64_16x16x4f64 and immediately dots each Y tile with the matching
// AO rows, so Y never leaves the registers and every quantity of a mode comes
// out of a single pass.
//
// GEMM columns: rho and gradient modes take one column type, phi, for 64
// points (four column blocks of 16); lap and Hessian modes take four, phi,
// d_x phi, d_y phi, d_z phi of 16 points.  Two work decompositions:
//   * contract_f64_pointset_kernel (gradient, lap and Hessian modes): the four
//     waves of a workgroup share a point set and split the rows of P, the
//     point set's AO rows staged through LDS once per round of 256 rows.
//   * contract_f64_rho_kernel (rho mode): a pair of wavefronts owns 64 points
//     and all rows of P, each wave taking half of the 64-row blocks of P and
//     streaming all of K per block into a 64 x 64 FP64 tile (16 MFMA
//     accumulators), its operands read straight from global memory in
//     coalesced 128-byte row segments (A[i][k] = P[k][i], P being symmetric;
//     B[k][j] = AO[k][p0 + j]); the pair adds its partial sums through LDS.
//     Its split form (one wave per 64 points and 64-row block) serves small
//     point counts.
//
// MFMA 16x16x4 f64 lane layout on gfx950 (lane l, r = l % 16, q = l / 16),
// measured with an in-bounds probe:
//   A operand: A[r][q];  B operand: B[q][r];  D: D[q + 4 j][r], j = 0..3
// (rows interleaved -- unlike the f32/bf16 16x16 variants, whose D rows are
// 4 q + j).

#include "ct_internal.h"

namespace ct {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int kIBlock = 64;  // rows of P per wave pass of the rho kernel; the basis padding is a multiple

// Epilogue slots and products of the point-set kernel.  Slot l of a lane's
// row: derivative ep_slot_d(l) of the point in column block ep_slot_cb(l)
// (gradient mode: four blocks of 16 points; lap and Hessian: one block).
// Product (cb, l): the Y column block cb times slot l, accumulated into dot
// ep_dot(cb, l), or -1 if the mode forms no such product.
//   grad: cb x {phi, d_x, d_y, d_z phi}     -> dot 4 cb + d
//   lap:  Y(phi) x {phi, d_k phi, lap phi}  -> dots 0..4;  Y(d_k phi) x d_k phi -> dot 4 + k
//   hess: Y(phi) x {phi, d_k, d_kl phi}     -> dots 0..9;  Y(d_k phi) x d_l phi (l <= k) -> 10..15
// Each dot receives one product per (row block, row), in row order.
__host__ __device__ constexpr int ep_slots(int mode) { return mode == MODE_GRAD ? 16 : mode == MODE_LAP ? 5 : 10; }
__host__ __device__ constexpr int ep_slot_cb(int mode, int l) { return mode == MODE_GRAD ? l / 4 : 0; }
__host__ __device__ constexpr int ep_slot_d(int mode, int l) { return mode == MODE_GRAD ? l % 4 : l; }
__host__ __device__ constexpr int ep_dot(int mode, int cb, int l) {
  if (mode == MODE_GRAD) return l / 4 == cb ? l : -1;
  if (cb == 0) return l;
  if (mode == MODE_LAP) return l == cb ? 4 + cb : -1;
  // hess: t_kl for l <= k, in the order xx, xy, yy, xz, yz, zz
  const int base = cb == 1 ? 10 : cb == 2 ? 11 : 13;
  return (l >= 1 && l <= cb) ? base + l - 1 : -1;
}

// Rho mode, wave pairs: 64 points (four column blocks of 16) per pair of
// waves.  ISPLIT: one 64-row block of P per wave (blockIdx.y), each writing
// its own partial rho [block][point] that reduce_row_blocks adds in order --
// enough waves for small point counts.  Unsplit: waves 2h and 2h + 1 of a
// workgroup share a point set; wave 2h takes the row blocks b with b % 4 in
// {0, 3}, wave 2h + 1 those with b % 4 in {1, 2} (equal work under the
// triangular k ranges below, whose length falls linearly with b).  For
// nbasis_pad = 128 both forms add the same terms in the same order.
// Two waves per SIMD: left to itself the compiler keeps the 64 x 64 FP64
// accumulator tile in 128 AGPRs beside the other registers (one wave per
// SIMD); the bound fits both in 256.
template <bool ISPLIT>
__global__ __launch_bounds__(256, 2) void contract_f64_rho_kernel(const double* __restrict__ P, int mp,
                                                               const double* __restrict__ ao, int64_t ld,
                                                               double* __restrict__ dots) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  const int half = ISPLIT ? 0 : (wave & 1);
  const int64_t p0 = ISPLIT ? ((int64_t)blockIdx.x * 4 + wave) * 64 : ((int64_t)blockIdx.x * 2 + (wave >> 1)) * 64;
  const double* bptr = ao + (int64_t)q * ld + p0 + r;  // column block cb at + 16 cb
  double dacc[4] = {0.0, 0.0, 0.0, 0.0};

  const int i_begin = ISPLIT ? (int)blockIdx.y * kIBlock : 0;
  const int i_end = ISPLIT ? i_begin + kIBlock : mp;
  if (ISPLIT) dots += (int64_t)blockIdx.y * ld;
  for (int i0 = i_begin; i0 < i_end; i0 += kIBlock) {
    if (!ISPLIT) {
      const int m4 = (i0 / kIBlock) & 3;
      if ((m4 == 0 || m4 == 3) != (half == 0)) continue;
    }
    d4 acc[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = d4{0.0, 0.0, 0.0, 0.0};

    // rho = phi^T P phi needs only the diagonal-and-right blocks: by the
    // symmetry of P, sum_IJ x_I^T P_IJ x_J = sum_I x_I^T (P_II x_I +
    // 2 sum_{J>I} P_IJ x_J), so k >= i0 only, with the B values doubled
    // (exactly) right of the diagonal block.  k order: right of the diagonal
    // block, then the diagonal block -- the epilogue then reads AO rows
    // i0..i0+63 while they are still in L2 from the last k-steps.
    const double* aptr = P + (int64_t)q * mp + i0 + r;
    const int ka = i0 + kIBlock;
    const int kfirst = ka < mp ? ka : i0;
    double a[4], b[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) a[rb] = aptr[(int64_t)kfirst * mp + 16 * rb];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) b[cb] = bptr[(int64_t)kfirst * ld + 16 * cb];

    // one k-step on the operands in a/b, prefetching k-step kn
    auto kstep = [&](int kn, bool dbl) {
      double an[4], bn[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) an[rb] = aptr[(int64_t)kn * mp + 16 * rb];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) bn[cb] = bptr[(int64_t)kn * ld + 16 * cb];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
          acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], dbl ? 2.0 * b[cb] : b[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) a[rb] = an[rb];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) b[cb] = bn[cb];
    };
    // k-steps kb, kb + 4, .. < ke; the last one prefetches knext (peeled, so
    // the loop's loads do not depend on the k index)
    auto phase = [&](int kb, int ke, int knext, bool dbl) {
      int k0 = kb;
      for (; k0 + 4 < ke; k0 += 4) kstep(k0 + 4, dbl);
      if (k0 < ke) kstep(knext, dbl);
    };
    phase(ka, mp, i0, true);
    phase(i0, i0 + kIBlock, i0, false);

    // epilogue: rho += phi_i (P phi)_i over the rows i0..i0+63 of this wave's
    // points, the 16 AO values of a 16-row block loaded as one batch
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      double v[4][4];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) v[jj][cb] = ao[(i0 + 16 * rb + q + 4 * jj) * ld + p0 + 16 * cb + r];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) dacc[cb] += v[jj][cb] * acc[rb][cb][jj];
    }
  }

  // reduce the four lane groups that hold the same point
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    double v = dacc[cb];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    dacc[cb] = v;
  }
  if (!ISPLIT) {
    __shared__ double pair_part[2][4][16];
    if (half == 1 && q == 0) {
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) pair_part[wave >> 1][cb][r] = dacc[cb];
    }
    __syncthreads();
    if (half == 1) return;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) dacc[cb] += pair_part[wave >> 1][cb][r];
  }
  if (q == 0) {
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) dots[p0 + 16 * cb + r] = dacc[cb];
  }
}

// Point-set workgroups (gradient, lap and Hessian modes; rho mode stays with
// the wave pairs above, whose independent waves run its short triangular k
// ranges faster): the four waves of a workgroup share one point set (the 64
// GEMM columns above) and split the rows of P between them, so the point
// set's AO columns are read from HBM once per round of rows instead of once
// per 64-row block (organic60 lap mode: 69 KB per point against 187 KB).
// A round covers up to 256 rows of P as 16 row blocks of 16; wave w holds
// blocks base + w, base + 7 - w, base + 8 + w, base + 15 - w (equal totals of
// the triangular k ranges), accumulating each block's 16 x 64 FP64 tile in
// the MFMA registers.  The k loop streams the point set's AO rows through LDS
// once per round, 16 rows (four MFMA k-steps) per chunk, double-buffered with
// one barrier per chunk; B is read from LDS one k-step ahead, A (P, a few MB)
// from global memory through a ring one chunk ahead.  The AO stream is read
// non-temporally: it is used once, and in L2 it would push out P.
// Lap mode's derivative columns enter only tau_kk = d_k phi^T P d_k phi, a
// symmetric form: the row block at rows i0..i0+15 needs k >= i0 only, doubled
// from i0 + 16 on (see contract_f64_rho_kernel), so the chunks below a round's
// first row carry only the phi columns.  AO bytes per point set and round:
// the phi columns over all k, the derivative columns over k >= the round's
// first row (all k in the gradient and Hessian modes), and the epilogue's AO
// rows of the round.
constexpr int kKC = 16;            // AO rows per LDS chunk
constexpr int kLdsPitch = 64 + 2;  // doubles per LDS row (the four k rows of a B read start 4 banks apart)

template <int MODE>
__global__ __launch_bounds__(256, 2) void contract_f64_pointset_kernel(const double* __restrict__ P, int mp,
                                                                    const double* __restrict__ ao, int64_t ld,
                                                                    double* __restrict__ dots) {
  static_assert(MODE != MODE_RHO, "rho mode: contract_f64_rho_kernel");
  constexpr int NT = MODE == MODE_GRAD ? 1 : 4;
  constexpr int PPB = NT == 1 ? 64 : 16;  // points per workgroup
  constexpr int ND = mode_ndot(MODE);
  constexpr int NACC = NT == 1 ? 4 * ND : ND;
  constexpr bool TRI = MODE == MODE_LAP;  // derivative columns triangular
  __shared__ double sB[2][kKC][kLdsPitch];
  static_assert(4 * NACC * 16 <= 2 * kKC * kLdsPitch, "the wave partials reuse the B buffers");

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: the k-loop's conditions are scalar
  const int r = lane & 15, q = lane >> 4;
  const int64_t p0 = (int64_t)blockIdx.x * PPB;
  const int64_t sd = (int64_t)mp * ld;
  const int S = mp >> 4;  // 16-row blocks of P (mp is a multiple of 128)
  const int nch = mp / kKC;

  // staging: thread tid moves two 16-byte pieces of a chunk, rows tid >> 5 and
  // 8 + (tid >> 5), LDS columns c, c + 1 with c = 2 (tid & 31); column c is
  // point c of the set (one column type) or point c % 16 of type c / 16.  In
  // chunks without derivative columns the derivative pieces re-read the phi
  // values of the same point (lines the workgroup fetches anyway): every
  // load stays unconditional.
  const int srow = tid >> 5, scol = (tid & 31) * 2;
  const bool sphi = NT == 1 || scol < 16;
  const double* gsrc = NT == 1 ? ao + (int64_t)srow * ld + p0 + scol
                               : ao + (int64_t)(scol >> 4) * sd + (int64_t)srow * ld + p0 + (scol & 15);
  const double* gphi = NT == 1 ? gsrc : ao + (int64_t)srow * ld + p0 + (scol & 15);
  const int64_t shalf = (int64_t)8 * ld;  // the second piece: 8 rows on
  auto stage_load = [&](int k, int kd_lo, d2& v0, d2& v1) {
    const double* src = (sphi || k >= kd_lo) ? gsrc : gphi;
    v0 = __builtin_nontemporal_load(reinterpret_cast<const d2*>(src + (int64_t)k * ld));
    v1 = __builtin_nontemporal_load(reinterpret_cast<const d2*>(src + (int64_t)k * ld + shalf));
  };
  auto stage_store = [&](int buf, const d2& v0, const d2& v1) {
    sB[buf][srow][scol] = v0.x;
    sB[buf][srow][scol + 1] = v0.y;
    sB[buf][srow + 8][scol] = v1.x;
    sB[buf][srow + 8][scol + 1] = v1.y;
  };

  double dacc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; ++k) dacc[k] = 0.0;

  for (int base = 0; base < S; base += 16) {
    const int nrb = (S - base) >= 16 ? 4 : (S - base) >> 2;  // row blocks per wave this round (4 or 2)
    auto blk = [&](int j) { return base + 8 * (j >> 1) + ((j & 1) ? 7 - wave : wave); };
    const int kd_lo = TRI ? base * 16 : 0;  // first k the derivative columns need
    d4 acc[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[j][cb] = d4{0.0, 0.0, 0.0, 0.0};

    {
      d2 v0, v1;
      stage_load(0, kd_lo, v0, v1);
      stage_store(0, v0, v1);
    }
    __syncthreads();

    // A of the chunk's four k-steps, each refilled with the next chunk's as it
    // is consumed.  Loads are unconditional (a block j >= nrb reads block
    // nrb - 1, the last chunk itself again): a load under a branch makes the
    // compiler wait for it where the branches join.
    const double* aptr[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) aptr[j] = P + (int64_t)q * mp + blk(j < nrb ? j : nrb - 1) * 16 + r;
    double a[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) a[s][j] = aptr[j][(int64_t)(4 * s) * mp];
    double b[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) b[cb] = sB[0][q][16 * cb + r];

    for (int ch = 0; ch < nch; ++ch) {
      const int buf = ch & 1;
      const int kc = ch * kKC;
      const int kn = ch + 1 < nch ? kc + kKC : kc;
      d2 v0, v1;
      stage_load(kn, kd_lo, v0, v1);  // stored once this chunk's MFMAs are issued
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        // B of the next k-step (after the last k-step of the chunk that is in
        // the other buffer, read after the barrier; this read is then dead)
        double bn[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) bn[cb] = sB[buf][(4 * (s + 1)) % kKC + q][16 * cb + r];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (j >= nrb) continue;
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
            if (!(TRI && cb > 0))
              acc[j][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s][j], b[cb], acc[j][cb], 0, 0, 0);
          // triangular columns: the chunk reaches block j when kc >= i0 and
          // lies right of its diagonal block when kc >= i0 + 16 (chunks and
          // blocks are both 16 rows).  The doubling is a factor on A, so the
          // block's MFMAs are one sequence under one branch (MFMAs on both
          // sides of an if/else made the compiler copy the accumulators after
          // every k-step).
          if (!TRI || kc < blk(j) * 16) continue;
          const double af = kc >= blk(j) * 16 + 16 ? 2.0 * a[s][j] : a[s][j];  // exact
#pragma unroll
          for (int cb = 1; cb < 4; ++cb)
            acc[j][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, b[cb], acc[j][cb], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) a[s][j] = aptr[j][(int64_t)(kn + 4 * s) * mp];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) b[cb] = bn[cb];
      }
      stage_store(buf ^ 1, v0, v1);
      __syncthreads();
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) b[cb] = sB[buf ^ 1][q][16 * cb + r];
    }

    // epilogue: dot each row block's Y tile with the AO rows of its 16 rows
    // (lane row q + 4 j of the block).  Per lane, slot l of a row is the AO
    // value of derivative ep_slot_d(l) at column block ep_slot_cb(l); the
    // values of JC rows x LC slots are loaded as one batch before their
    // products.
    constexpr int NL = ep_slots(MODE);
    constexpr int JC = MODE == MODE_LAP ? 2 : 1;
    constexpr int LC = MODE == MODE_GRAD ? 8 : 5;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= nrb) continue;
      const int64_t i0 = blk(j) * 16;
#pragma unroll
      for (int j0 = 0; j0 < 4; j0 += JC) {
#pragma unroll
        for (int l0 = 0; l0 < NL; l0 += LC) {
          double v[JC][LC];
#pragma unroll
          for (int jj = 0; jj < JC; ++jj) {
            const int64_t i = i0 + q + 4 * (j0 + jj);
#pragma unroll
            for (int ll = 0; ll < LC; ++ll)
              v[jj][ll] = __builtin_nontemporal_load(ao + ep_slot_d(MODE, l0 + ll) * sd + i * ld + p0 +
                                                     16 * ep_slot_cb(MODE, l0 + ll) + r);
          }
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int jj = 0; jj < JC; ++jj)
#pragma unroll
              for (int ll = 0; ll < LC; ++ll) {
                constexpr int kNone = -1;
                const int k = ep_dot(MODE, cb, l0 + ll);
                if (k != kNone) dacc[k] += v[jj][ll] * acc[j][cb][j0 + jj];
              }
        }
      }
    }
  }

  // the four lane groups of a point, then the four waves in order; the wave
  // partials go where the B chunks were, once every wave is done with them
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    double v = dacc[k];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    dacc[k] = v;
  }
  double(*red)[NACC][16] = reinterpret_cast<double(*)[NACC][16]>(&sB[0][0][0]);
  __syncthreads();
  if (q == 0) {
#pragma unroll
    for (int k = 0; k < NACC; ++k) red[wave][k][r] = dacc[k];
  }
  __syncthreads();
  for (int e = tid; e < NACC * 16; e += 256) {
    const int k = e >> 4, c = e & 15;
    const double v = ((red[0][k][c] + red[1][k][c]) + red[2][k][c]) + red[3][k][c];
    if constexpr (NT == 1) {
      dots[(int64_t)(k % ND) * ld + p0 + 16 * (k / ND) + c] = v;
    } else {
      dots[(int64_t)k * ld + p0 + c] = v;
    }
  }
}

__global__ __launch_bounds__(256) void reduce_row_blocks_kernel(const double* __restrict__ part, int nblk,
                                                                int64_t count, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * count + k];
  out[k] = s;
}


// Factor route in FP64 (P = L L^T, L stored k-major: L[k * rows + j]): one
// wave owns 64 GEMM columns -- NT column types x 64/NT points -- and all
// `rows` factor columns, 64 at a time; the tile psi = L^T [columns] stays in
// the MFMA accumulators and every dot is a lane-local sum over the factor
// columns (phi^T P x = psi^T (L^T x)), reduced over the four lane groups.
//   F1:  phi x 64 points (rho mode: rho);
//   F4:  {phi, d_x phi, d_y phi, d_z phi} x 16 points: rho, g_k and t_kk (lap
//        mode) or t_kl (Hessian mode); in the lap and Hessian modes it also
//        stores psi (psi[row * ld + point]) for the launches below, which
//        dot their columns with that psi instead of forming it again:
//   L1:  lap phi x 64 points: psi . lap psi;
//   H4:  {d_xx, d_xy, d_xz, d_yy phi} x 16 points: h_xx, h_xy, h_xz, h_yy;
//   H2:  {d_yz, d_zz phi} x 32 points: h_yz, h_zz.
// Split (npass > 0: few points, or three or more passes of L^T) each
// workgroup takes the 64 factor columns of one pass y only and writes its
// partial dots at dots + y * ndot * ld (summed in order afterwards): more
// waves for the same work, and the passes of a point group share its AO
// columns through one XCD's L2.
enum VF64 : int { F64_F1, F64_F4, F64_L1, F64_H4, F64_H2 };

// Three waves per SIMD (168 VGPRs) where that fits without spills; the
// lap and Hessian F4 tiles keep two (their 3-wave builds spill).  Few-point
// launches (the re-evaluation of a mixed call's flagged points) are latency
// bound: 1.14 -> 0.79 ms per step for 25 000 points.
// kmask (optional, the screened layout): per tile of tp points, the 32-function
// k-steps holding a shell within its cutoff radius (nwords 32-bit words per
// tile); a wave contracts the union of its tiles' k-steps only (the AO
// kernel writes every point of its 64-point waves for that union).  kmask
// null: every k-step.

// SCR: the screened layout (kmask given), at two waves per SIMD (the masked
// k-loop needs a few more registers than three waves leave); unscreened
// launches keep the plain k-loop and three waves where the variant allows.
template <int VF, int MODE, bool SCR>
__global__ __launch_bounds__(256, SCR || ((VF == F64_F4 || VF == F64_H4) && MODE != MODE_GRAD) ? 2 : 3) void contract_f64_factor_kernel(const double* __restrict__ L, int rows, int mp,
                                                                  const double* __restrict__ ao, int64_t ld,
                                                                  double* __restrict__ dots,
                                                                  const uint32_t* __restrict__ kmask, int nwords,
                                                                  int tp, double* __restrict__ psi, int npass) {
  constexpr int NT = VF == F64_F1 || VF == F64_L1 ? 1 : VF == F64_H2 ? 2 : 4;
  constexpr bool PSI = VF == F64_L1 || VF == F64_H4 || VF == F64_H2;  // dots with the stored psi
  constexpr int PPW = 64 / NT;   // points per wave
  constexpr int NPB = PPW / 16;  // 16-point column blocks per type
  constexpr bool HES = MODE == MODE_HESS, LAP = MODE == MODE_LAP;
  // dots per point: F1 [rho]; L1 [L0]; H4, H2 [h x NT]; F4 [rho, g_x, g_y,
  // g_z, (lap: t_xx, t_yy, t_zz | Hessian: t_xx, t_xy, t_yy, t_xz, t_yz, t_zz)]
  constexpr int NDV = VF == F64_F1 || VF == F64_L1 ? 1 : VF == F64_F4 ? (LAP ? 7 : HES ? 10 : 4) : NT;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int r = lane & 15, q = lane >> 4;
  // split (npass > 0): workgroup b runs on XCD b % 8 and takes pass
  // j % npass of point group (j / npass) * 8 + b % 8, j = b / 8, so the
  // passes of a group follow one another on one XCD and read its AO columns
  // from that XCD's L2 instead of HBM
  const bool split = npass > 0;
  int64_t grp = blockIdx.x;
  int pass = 0;
  if (split) {
    const int j = (int)(blockIdx.x >> 3);
    pass = j % npass;
    grp = (int64_t)(j / npass) * 8 + (blockIdx.x & 7);
    if (grp >= ld / (4 * PPW)) return;  // (no barriers in this kernel)
  }
  const int64_t p0 = (grp * 4 + wave) * PPW;
  const int64_t sd = (int64_t)mp * ld;
  // the wave's active k-steps: the union of its (one or two) tiles' masks,
  // read word by word when the k-loop needs the next active k-step (once per
  // eight MFMA steps; wave-uniform, scalar loads)
  const int nk = mp / 32;
  const int64_t t0 = kmask ? p0 / tp : 0, t1 = kmask ? (p0 + PPW - 1) / tp : 0;
  // the first active k-step after kb (nk: none)
  auto next_active = [&](int kb) {
    int k = kb + 1;
    while (k < nk) {
      const int w = k >> 5;
      uint32_t m = 0xFFFFFFFFu;
      if (kmask) {
        m = 0;
        for (int64_t t = t0; t <= t1; ++t) m |= kmask[t * nwords + w];
      }
      m >>= (k & 31);
      if (m) return k + __builtin_ctz(m);
      k = (w + 1) << 5;
    }
    return nk;
  };
  const int i_begin = split ? pass * 64 : 0;
  const int i_end = split ? i_begin + 64 : rows;
  dots += split ? (int64_t)pass * mode_ndot(MODE) * ld : 0;

  // AO array of type t (0 = phi, 1..3 = d phi, 4 = lap phi | 4..9 = d_xx,
  // d_xy, d_xz, d_yy, d_yz, d_zz phi)
  auto arr = [](int t) { return VF == F64_L1 ? 4 : VF == F64_H4 ? 4 + t : VF == F64_H2 ? 8 + t : t; };
  const double* bptr[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) bptr[cb] = ao + arr(cb / NPB) * sd + (int64_t)q * ld + p0 + 16 * (cb % NPB) + r;

  double dacc[NPB][NDV];
#pragma unroll
  for (int pb = 0; pb < NPB; ++pb)
#pragma unroll
    for (int k = 0; k < NDV; ++k) dacc[pb][k] = 0.0;

  for (int i0 = i_begin; i0 < i_end; i0 += 64) {
    d4 acc[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = d4{0.0, 0.0, 0.0, 0.0};
    // A[r][q] = L^T[i0 + 16 rb + r][k0 + q] = L[k0 + q][i0 + 16 rb + r];
    // the k-loop walks the active 32-function k-steps, eight MFMA steps each,
    // the operands of the next step prefetched (within a k-step, or the
    // first of the next active one)
    const double* aptr = L + (int64_t)q * rows + i0 + r;
    if constexpr (SCR) {
      int kb = next_active(-1);
      if (kb < nk) {
        int kb_next = next_active(kb);
        int k0 = 32 * kb;
        double a[4], b[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) a[rb] = aptr[(int64_t)k0 * rows + 16 * rb];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) b[cb] = bptr[cb][(int64_t)k0 * ld];
        for (;;) {
          const bool step_end = ((k0 + 4) & 31) == 0;  // last MFMA step of the k-step
          // prefetch (clamped on the very last step; unused)
          const int kn = !step_end ? k0 + 4 : (kb_next < nk ? 32 * kb_next : k0);
          double an[4], bn[4];
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) an[rb] = aptr[(int64_t)kn * rows + 16 * rb];
#pragma unroll
          for (int cb = 0; cb < 4; ++cb) bn[cb] = bptr[cb][(int64_t)kn * ld];
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
              acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) a[rb] = an[rb];
#pragma unroll
          for (int cb = 0; cb < 4; ++cb) b[cb] = bn[cb];
          if (step_end) {
            if (kb_next >= nk) break;
            kb = kb_next;
            kb_next = next_active(kb);
          }
          k0 = kn;
        }
      }
    } else {
      double a[4], b[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) a[rb] = aptr[16 * rb];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) b[cb] = bptr[cb][0];
      for (int k0 = 0; k0 < mp; k0 += 4) {
        const int kn = (k0 + 4 < mp) ? k0 + 4 : k0;  // prefetch (clamped on the last step; unused)
        double an[4], bn[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) an[rb] = aptr[(int64_t)kn * rows + 16 * rb];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) bn[cb] = bptr[cb][(int64_t)kn * ld];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) a[rb] = an[rb];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) b[cb] = bn[cb];
      }
    }
    // lane (r, q) holds rows q + 4 j of every 16-row block rb, column r of
    // every column block: all column types of its point(s)
    auto psi_at = [&](int rb, int j, int pb) {
      return psi + (int64_t)(i0 + 16 * rb + q + 4 * j) * ld + p0 + 16 * pb + r;
    };
#pragma unroll
    for (int pb = 0; pb < NPB; ++pb) {
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        double ps[4];  // the stored psi of the block's rows (PSI variants)
        if constexpr (PSI) {
#pragma unroll
          for (int j = 0; j < 4; ++j) ps[j] = *psi_at(rb, j, pb);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double y0 = PSI ? ps[j] : acc[rb][pb][j];
          if constexpr (VF == F64_F1) {
            dacc[pb][0] += y0 * y0;
          } else if constexpr (PSI) {
#pragma unroll
            for (int k = 0; k < NT; ++k) dacc[pb][k] += y0 * acc[rb][k * NPB + pb][j];
          } else {
            if constexpr (LAP || HES) *psi_at(rb, j, pb) = y0;
            const double yx = acc[rb][1][j], yy = acc[rb][2][j], yz = acc[rb][3][j];
            dacc[pb][0] += y0 * y0;
            dacc[pb][1] += y0 * yx;
            dacc[pb][2] += y0 * yy;
            dacc[pb][3] += y0 * yz;
            if constexpr (LAP) {
              dacc[pb][4] += yx * yx;
              dacc[pb][5] += yy * yy;
              dacc[pb][6] += yz * yz;
            }
            if constexpr (HES) {
              dacc[pb][4] += yx * yx;
              dacc[pb][5] += yx * yy;
              dacc[pb][6] += yy * yy;
              dacc[pb][7] += yx * yz;
              dacc[pb][8] += yy * yz;
              dacc[pb][9] += yz * yz;
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int pb = 0; pb < NPB; ++pb)
#pragma unroll
    for (int k = 0; k < NDV; ++k) {
      double v = dacc[pb][k];
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      dacc[pb][k] = v;
    }
  if (q == 0) {
#pragma unroll
    for (int pb = 0; pb < NPB; ++pb)
#pragma unroll
      for (int k = 0; k < NDV; ++k) {
        // slots of the mode's dot layout (mode_ndot)
        const int slot = VF == F64_L1   ? 4
                         : VF == F64_H4 ? 4 + k
                         : VF == F64_H2 ? 8 + k
                         : k < 4        ? k
                         : HES          ? k + 6
                                        : k + 1;
        dots[(int64_t)slot * ld + p0 + 16 * pb + r] = dacc[pb][k];
      }
  }
}

template <int VF, int MODE>
void launch_factor_var(const double* L, int rows, int mp, const double* ao, int64_t ld, double* dots, hipStream_t s,
                       bool split, const uint32_t* kmask, int nwords, int tp, double* psi) {
  constexpr int PPW = VF == F64_F1 || VF == F64_L1 ? 64 : VF == F64_H2 ? 32 : 16;
  // ld: multiple of 256; split: rows / 64 passes per point group, the
  // groups rounded up to a multiple of the 8 XCDs
  const int64_t ngroups = ld / (4 * PPW);
  const int npass = split ? rows / 64 : 0;
  const dim3 block(256), grid(split ? (unsigned)((ngroups + 7) / 8 * 8 * npass) : (unsigned)ngroups);
  if (kmask)
    hipLaunchKernelGGL((contract_f64_factor_kernel<VF, MODE, true>), grid, block, 0, s, L, rows, mp, ao, ld, dots,
                       kmask, nwords, tp, psi, npass);
  else
    hipLaunchKernelGGL((contract_f64_factor_kernel<VF, MODE, false>), grid, block, 0, s, L, rows, mp, ao, ld, dots,
                       kmask, nwords, tp, psi, npass);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace

namespace {
void factor_modes(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao, int64_t ld, double* dots,
                  hipStream_t s, bool split, const uint32_t* kmask, int nwords, int tp, double* psi) {
  CT_REQUIRE(rows_pad % 64 == 0 && rows_pad > 0 && ld % 256 == 0, "factor padding / point stride");
  CT_REQUIRE(!kmask || (nwords == (nbasis_pad / 32 + 31) / 32 && tp >= 16 && ld % tp == 0),
             "activity masks do not match the FP64 factor contraction");
  CT_REQUIRE(psi != nullptr || mode == MODE_RHO || mode == MODE_GRAD, "the lap and Hessian modes need the psi buffer");
  switch (mode) {
    case MODE_RHO:
      launch_factor_var<F64_F1, MODE_RHO>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp, psi);
      break;
    case MODE_GRAD:
      launch_factor_var<F64_F4, MODE_GRAD>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp, psi);
      break;
    case MODE_LAP:
      launch_factor_var<F64_F4, MODE_LAP>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp, psi);
      launch_factor_var<F64_L1, MODE_LAP>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp, psi);
      break;
    default:
      launch_factor_var<F64_F4, MODE_HESS>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp, psi);
      launch_factor_var<F64_H4, MODE_HESS>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp, psi);
      launch_factor_var<F64_H2, MODE_HESS>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp, psi);
      break;
  }
}
}  // namespace

void launch_contract_f64_factor(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao,
                                int64_t ld, double* dots, hipStream_t s, const uint32_t* kmask, int nwords, int tp,
                                double* psi) {
  factor_modes(mode, L, rows_pad, nbasis_pad, ao, ld, dots, s, false, kmask, nwords, tp, psi);
}

int64_t contract_f64_factor_split_workspace(int mode, int rows_pad, int64_t ld) {
  return (int64_t)(rows_pad / 64) * mode_ndot(mode) * ld;
}

void launch_contract_f64_factor_split(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao,
                                      int64_t ld, double* part, double* dots, hipStream_t s, const uint32_t* kmask,
                                      int nwords, int tp, double* psi) {
  factor_modes(mode, L, rows_pad, nbasis_pad, ao, ld, part, s, true, kmask, nwords, tp, psi);
  const int64_t count = (int64_t)mode_ndot(mode) * ld;
  hipLaunchKernelGGL(reduce_row_blocks_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, part,
                     rows_pad / 64, count, dots);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_contract_f64(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld, double* dots,
                         hipStream_t s) {
  const int ppb = (mode == MODE_RHO || mode == MODE_GRAD) ? 64 : 16;  // points per workgroup
  if (nbasis_pad % kRowBlock != 0 || ld % (2 * ppb) != 0) throw Error(CT_E_INVALID, "contract_f64: padding");
  const dim3 block(256), grid((unsigned)(ld / ppb));
  switch (mode) {
    case MODE_RHO:  // wave pairs: two point sets of 64 per workgroup
      hipLaunchKernelGGL((contract_f64_rho_kernel<false>), dim3((unsigned)(ld / 128)), block, 0, s, P, nbasis_pad,
                         ao, ld, dots);
      break;
    case MODE_GRAD:
      hipLaunchKernelGGL((contract_f64_pointset_kernel<MODE_GRAD>), grid, block, 0, s, P, nbasis_pad, ao, ld, dots);
      break;
    case MODE_LAP:
      hipLaunchKernelGGL((contract_f64_pointset_kernel<MODE_LAP>), grid, block, 0, s, P, nbasis_pad, ao, ld, dots);
      break;
    default:
      hipLaunchKernelGGL((contract_f64_pointset_kernel<MODE_HESS>), grid, block, 0, s, P, nbasis_pad, ao, ld, dots);
      break;
  }
  CT_HIP_CHECK(hipGetLastError());
}

int64_t contract_f64_split_workspace(int mode, int nbasis_pad, int64_t ld) {
  return (int64_t)(nbasis_pad / kIBlock) * mode_ndot(mode) * ld;
}

void launch_contract_f64_split(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld,
                               double* part, double* dots, hipStream_t s) {
  if (mode != MODE_RHO || nbasis_pad % kIBlock != 0 || ld % 256 != 0)
    throw Error(CT_E_INVALID, "contract_f64_split: rho mode, padded sizes");
  hipLaunchKernelGGL((contract_f64_rho_kernel<true>), dim3((unsigned)(ld / 256), (unsigned)(nbasis_pad / kIBlock)),
                     dim3(256), 0, s, P, nbasis_pad, ao, ld, part);
  CT_HIP_CHECK(hipGetLastError());
  const int64_t count = (int64_t)mode_ndot(mode) * ld;
  hipLaunchKernelGGL(reduce_row_blocks_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, part,
                     nbasis_pad / kIBlock, count, dots);
  CT_HIP_CHECK(hipGetLastError());
}

namespace {

// Largest violation of |R_ij| <= sqrt(R_ii R_jj) over i, j < m, R = A - L L^T
// (A: mp x mp row-major; L: row i of the factor at L[i * rows ..], rank
// columns; diag[i] = R_ii).  A positive semidefinite remainder satisfies it
// to rounding; an indefinite one does not, even with a zero diagonal.  NaN
// counts as an infinite violation.  16 x 16 threads per block, one (i, j)
// each; the block's largest value goes to *out with one atomic (non-negative
// doubles order like their bit patterns).
__global__ __launch_bounds__(256) void factor_residual_kernel(const double* __restrict__ A, int mp,
                                                              const double* __restrict__ L, int rows, int rank,
                                                              int m, const double* __restrict__ diag,
                                                              unsigned long long* __restrict__ out) {
  const int i = blockIdx.y * 16 + threadIdx.y;
  const int j = blockIdx.x * 16 + threadIdx.x;
  double v = 0.0;
  if (i < m && j < m) {
    double r = A[(int64_t)i * mp + j];
    const double* li = L + (int64_t)i * rows;
    const double* lj = L + (int64_t)j * rows;
    for (int k = 0; k < rank; ++k) r = fma(-li[k], lj[k], r);
    v = fabs(r) - sqrt(fmax(diag[i], 0.0) * fmax(diag[j], 0.0));
    v = v <= 1e300 ? fmax(v, 0.0) : 1e300;
  }
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  __shared__ double wmax[4];
  const int t = threadIdx.y * 16 + threadIdx.x;
  if ((t & 63) == 0) wmax[t >> 6] = v;
  __syncthreads();
  if (t == 0) {
    const double b = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    if (b > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(b));
  }
}

}  // namespace

void launch_factor_residual(const double* A, int mp, const double* L, int rows, int rank, int m, const double* diag,
                            unsigned long long* out, hipStream_t s) {
  if (m <= 0) return;
  const dim3 grid((unsigned)((m + 15) / 16), (unsigned)((m + 15) / 16)), block(16, 16);
  hipLaunchKernelGGL(factor_residual_kernel, grid, block, 0, s, A, mp, L, rows, rank, m, diag, out);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
