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
// Work decomposition: a pair of wavefronts owns a set of points and all rows
// of P, each wave taking half of the 64-row blocks of P.
//   * 1 GEMM column type (MODE_RHO, MODE_GRAD): 4 column blocks x 16 points
//     = 64 points per wave pair.
//   * 4 column types (MODE_LAP, MODE_HESS): the 4 column blocks are phi,
//     d_x phi, d_y phi, d_z phi of the same 16 points.
// For each of its 64-row blocks of P (4 MFMA row blocks) a wave streams all
// of K, accumulating a 64 x 64 FP64 tile (16 MFMA accumulators), then runs
// the epilogue; the pair adds its partial dots through LDS at the end.
// Operands are read straight from global memory with coalesced 128-byte row
// segments: A[i][k] = P[k][i] (P is symmetric), B[k][j] = AO[k][p0 + j]; P
// (<= a few MB) stays L2-resident, and the two waves of a pair stream the
// same AO columns at about the same time.
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

constexpr int kIBlock = 64;  // rows of P per wave pass (4 MFMA row blocks); the basis padding is a multiple

// Epilogue slots and products.  Slot l of a lane's row: derivative
// ep_slot_d(l) of the point in column block ep_slot_cb(l) (one column type:
// four blocks of 16 points; four column types: one block).  Product (cb, l):
// the Y column block cb times slot l, accumulated into dot ep_dot(cb, l), or
// -1 if the mode forms no such product.
//   rho:  cb x {phi}                        -> dot cb
//   grad: cb x {phi, d_x, d_y, d_z phi}     -> dot 4 cb + d
//   lap:  Y(phi) x {phi, d_k phi, lap phi}  -> dots 0..4;  Y(d_k phi) x d_k phi -> dot 4 + k
//   hess: Y(phi) x {phi, d_k, d_kl phi}     -> dots 0..9;  Y(d_k phi) x d_l phi (l <= k) -> 10..15
// Each dot receives one product per (row block, row), in row order.
__host__ __device__ constexpr int ep_slots(int mode) {
  return mode == MODE_RHO ? 4 : mode == MODE_GRAD ? 16 : mode == MODE_LAP ? 5 : 10;
}
__host__ __device__ constexpr int ep_slot_cb(int mode, int l) {
  return mode == MODE_RHO ? l : mode == MODE_GRAD ? l / 4 : 0;
}
__host__ __device__ constexpr int ep_slot_d(int mode, int l) {
  return mode == MODE_RHO ? 0 : mode == MODE_GRAD ? l % 4 : l;
}
__host__ __device__ constexpr int ep_dot(int mode, int cb, int l) {
  if (mode == MODE_RHO) return l == cb ? cb : -1;
  if (mode == MODE_GRAD) return l / 4 == cb ? l : -1;
  if (cb == 0) return l;
  if (mode == MODE_LAP) return l == cb ? 4 + cb : -1;
  // hess: t_kl for l <= k, in the order xx, xy, yy, xz, yz, zz
  const int base = cb == 1 ? 10 : cb == 2 ? 11 : 13;
  return (l >= 1 && l <= cb) ? base + l - 1 : -1;
}

// ISPLIT: one 64-row block of P per wave (blockIdx.y), each writing its own
// partial dots [block][dot][point] that reduce_row_blocks adds in order --
// enough waves for small point counts (the mixed path's re-evaluation).
// Unsplit: waves 2h and 2h + 1 of a workgroup share a point set; wave 2h
// takes the row blocks b with b % 4 in {0, 3}, wave 2h + 1 those with
// b % 4 in {1, 2} (equal work under the triangular k ranges below, whose
// length falls linearly with b).
// Two waves per SIMD: left to itself the compiler keeps the 64 x 64 FP64
// accumulator tile in 128 AGPRs beside ~190 VGPRs (one wave per SIMD); with
// the bound it fits both in 256 registers (organic60 FP64 contraction
// 183 -> 129 ms per step; the Hessian mode spills two registers).
template <int MODE, bool ISPLIT>
__global__ __launch_bounds__(256, 2) void contract_f64_kernel(const double* __restrict__ P, int mp,
                                                           const double* __restrict__ ao, int64_t ld,
                                                           double* __restrict__ dots) {
  constexpr int NT = (MODE == MODE_RHO || MODE == MODE_GRAD) ? 1 : 4;
  constexpr int PPW = NT == 1 ? 64 : 16;
  constexpr int ND = mode_ndot(MODE);
  constexpr int NACC = NT == 1 ? 4 * ND : ND;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  const int half = ISPLIT ? 0 : (wave & 1);
  const int64_t p0 = ISPLIT ? ((int64_t)blockIdx.x * 4 + wave) * PPW : ((int64_t)blockIdx.x * 2 + (wave >> 1)) * PPW;
  const int64_t sd = (int64_t)mp * ld;

  const double* bptr[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int t = NT == 1 ? 0 : cb;
    const int poff = NT == 1 ? 16 * cb : 0;
    bptr[cb] = ao + t * sd + (int64_t)q * ld + p0 + poff + r;
  }
  double dacc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; ++k) dacc[k] = 0.0;

  const int i_begin = ISPLIT ? (int)blockIdx.y * kIBlock : 0;
  const int i_end = ISPLIT ? i_begin + kIBlock : mp;
  if (ISPLIT) dots += (int64_t)blockIdx.y * ND * ld;
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

    // Quadratic forms x^T P x need only the diagonal-and-right blocks: by the
    // symmetry of P, sum_IJ x_I^T P_IJ x_J = sum_I x_I^T (P_II x_I +
    // 2 sum_{J>I} P_IJ x_J), so those columns use k >= i0 only, with their
    // values doubled (exactly) right of the diagonal block.  Rho mode: every
    // column (rho = phi^T P phi).  Lap mode: the derivative columns (cb 1..3),
    // which enter only tau_kk = d_k phi^T P d_k phi.
    constexpr bool TRI_ALL = MODE == MODE_RHO;
    constexpr bool TRI_DER = MODE == MODE_LAP;
    const double* aptr = P + (int64_t)q * mp + i0 + r;
    // k order: right of the diagonal block, left of it, the diagonal block
    // last -- the epilogue then reads AO rows i0..i0+63 while they are still
    // in L2 from the last k-steps
    const int ka = i0 + kIBlock;
    const int kfirst = ka < mp ? ka : (TRI_ALL || i0 == 0 ? i0 : 0);
    double a[4], b[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) a[rb] = aptr[(int64_t)kfirst * mp + 16 * rb];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) b[cb] = bptr[cb][(int64_t)kfirst * ld];

    // one k-step on the operands in a/b, prefetching k-step kn (all rows of P;
    // AO columns cb = 0 only unless load_all)
    auto kstep = [&](int kn, bool all_cols, bool dbl, bool load_all) {
      double an[4], bn[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) an[rb] = aptr[(int64_t)kn * mp + 16 * rb];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
        if (cb == 0 || load_all) bn[cb] = bptr[cb][(int64_t)kn * ld];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
          if (cb == 0 || all_cols)
            acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                a[rb], ((TRI_ALL || cb > 0) && dbl) ? 2.0 * b[cb] : b[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) a[rb] = an[rb];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) b[cb] = bn[cb];
    };
    // k-steps kb, kb + 4, .. < ke; the last one prefetches knext with every
    // column (peeled, so the loop's loads do not depend on the k index)
    auto phase = [&](int kb, int ke, int knext, bool all_cols, bool dbl, bool load_all) {
      int k0 = kb;
      for (; k0 + 4 < ke; k0 += 4) kstep(k0 + 4, all_cols, dbl, load_all);
      if (k0 < ke) kstep(knext, all_cols, dbl, true);
    };
    // phase B: left of the diagonal block (not in rho mode); lap mode needs
    // only the phi column there
    const int kb_end = TRI_ALL ? 0 : i0;
    phase(ka, mp, kb_end > 0 ? 0 : i0, true, TRI_ALL || TRI_DER, true);
    phase(0, kb_end, i0, !TRI_DER, false, !TRI_DER);
    phase(i0, i0 + kIBlock, i0, true, false, true);

    // epilogue: dot the Y tile with the AO rows i0..i0+63 of this wave's
    // points.  Per lane, slot l of row i is the AO value of derivative
    // ep_slot_d(l) at point column block ep_slot_cb(l); the values of JC rows
    // x LC slots are loaded as one batch before their products (lap mode: 20
    // loads in flight per 16-row block) rather than one at a time beside
    // each product.
    constexpr int NL = ep_slots(MODE);
    constexpr int JC = (MODE == MODE_GRAD || MODE == MODE_HESS) ? 1 : 4;
    constexpr int LC = MODE == MODE_GRAD ? 8 : MODE == MODE_HESS ? 5 : NL;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
      for (int j0 = 0; j0 < 4; j0 += JC) {
#pragma unroll
        for (int l0 = 0; l0 < NL; l0 += LC) {
          double v[JC][LC];
#pragma unroll
          for (int jj = 0; jj < JC; ++jj) {
            const int64_t i = i0 + 16 * rb + q + 4 * (j0 + jj);
#pragma unroll
            for (int ll = 0; ll < LC; ++ll)
              v[jj][ll] = ao[ep_slot_d(MODE, l0 + ll) * sd + i * ld + p0 + 16 * ep_slot_cb(MODE, l0 + ll) + r];
          }
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int jj = 0; jj < JC; ++jj)
#pragma unroll
              for (int ll = 0; ll < LC; ++ll) {
                constexpr int kNone = -1;
                const int k = ep_dot(MODE, cb, l0 + ll);
                if (k != kNone) dacc[k] += v[jj][ll] * acc[rb][cb][j0 + jj];
              }
        }
      }
    }
  }

  // reduce the four lane groups that hold the same point
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    double v = dacc[k];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    dacc[k] = v;
  }
  if (!ISPLIT) {
    __shared__ double pair_part[2][NACC][16];
    if (half == 1 && q == 0) {
#pragma unroll
      for (int k = 0; k < NACC; ++k) pair_part[wave >> 1][k][r] = dacc[k];
    }
    __syncthreads();
    if (half == 1) return;
#pragma unroll
    for (int k = 0; k < NACC; ++k) dacc[k] += pair_part[wave >> 1][k][r];
  }
  if (q == 0) {
    if constexpr (NT == 1) {
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int d = 0; d < ND; ++d) dots[(int64_t)d * ld + p0 + 16 * cb + r] = dacc[cb * ND + d];
    } else {
#pragma unroll
      for (int d = 0; d < ND; ++d) dots[(int64_t)d * ld + p0 + r] = dacc[d];
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

template <bool ISPLIT>
void launch_modes(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld, double* dots,
                  hipStream_t s) {
  const int ppw = (mode == MODE_RHO || mode == MODE_GRAD) ? 64 : 16;
  const dim3 block(256), grid((unsigned)(ld / ((ISPLIT ? 4 : 2) * ppw)), ISPLIT ? (unsigned)(nbasis_pad / kIBlock) : 1u);
  switch (mode) {
    case MODE_RHO:
      hipLaunchKernelGGL((contract_f64_kernel<MODE_RHO, ISPLIT>), grid, block, 0, s, P, nbasis_pad, ao, ld, dots);
      break;
    case MODE_GRAD:
      hipLaunchKernelGGL((contract_f64_kernel<MODE_GRAD, ISPLIT>), grid, block, 0, s, P, nbasis_pad, ao, ld, dots);
      break;
    case MODE_LAP:
      hipLaunchKernelGGL((contract_f64_kernel<MODE_LAP, ISPLIT>), grid, block, 0, s, P, nbasis_pad, ao, ld, dots);
      break;
    default:
      hipLaunchKernelGGL((contract_f64_kernel<MODE_HESS, ISPLIT>), grid, block, 0, s, P, nbasis_pad, ao, ld, dots);
      break;
  }
  CT_HIP_CHECK(hipGetLastError());
}

// Factor route in FP64 (P = L L^T, L stored k-major: L[k * rows + j]): one
// wave owns 64 GEMM columns -- NT column types x 64/NT points -- and all
// `rows` factor columns, 64 at a time; the tile psi = L^T [columns] stays in
// the MFMA accumulators and every dot is a lane-local sum over the factor
// columns (phi^T P x = psi^T (L^T x)), reduced over the four lane groups.
//   F1:  phi x 64 points (rho mode: rho);
//   F4:  {phi, d_x phi, d_y phi, d_z phi} x 16 points: rho, g_k and t_kk (lap
//        mode) or t_kl (Hessian mode);
//   F2:  {phi, lap phi} x 32 points: psi . lap psi;
//   FH0, FH1: {phi, d_kl phi for kl = xx, xy, xz | yy, yz, zz} x 16 points: h_kl.
// With gridDim.y > 1 (few points) each workgroup row y takes the 64 factor
// columns of pass y only and writes its partial dots at dots + y * ndot * ld
// (summed in order afterwards): more waves for the same work.
enum VF64 : int { F64_F1, F64_F4, F64_F2, F64_FH0, F64_FH1 };

// Three waves per SIMD (168 VGPRs) where that fits without spills; the
// lap and Hessian F4 tiles keep two (their 3-wave builds spill).  Few-point
// launches (the re-evaluation of a mixed call's flagged points) are latency
// bound: 1.14 -> 0.79 ms per step for 25 000 points.
// kmask (optional, the screened layout): per tile of tp points, the 32-function
// k-steps holding a shell within its cutoff radius (nwords 32-bit words per
// tile); a wave contracts the union of its tiles' k-steps only (a wave's
// points span at most two tiles).  kmask null: every k-step.

// SCR: the screened layout (kmask given), at two waves per SIMD (the masked
// k-loop needs a few more registers than three waves leave); unscreened
// launches keep the plain k-loop and three waves where the variant allows.
template <int VF, int MODE, bool SCR>
__global__ __launch_bounds__(256, SCR || (VF == F64_F4 && MODE != MODE_GRAD) ? 2 : 3) void contract_f64_factor_kernel(const double* __restrict__ L, int rows, int mp,
                                                                  const double* __restrict__ ao, int64_t ld,
                                                                  double* __restrict__ dots,
                                                                  const uint32_t* __restrict__ kmask, int nwords,
                                                                  int tp) {
  constexpr int NT = VF == F64_F1 ? 1 : VF == F64_F2 ? 2 : 4;
  constexpr int PPW = 64 / NT;   // points per wave
  constexpr int NPB = PPW / 16;  // 16-point column blocks per type
  constexpr bool HES = MODE == MODE_HESS, LAP = MODE == MODE_LAP;
  // dots per point: F1 [rho]; F2 [L0]; FH [h x 3]; F4 [rho, g_x, g_y, g_z,
  // (lap: t_xx, t_yy, t_zz | Hessian: t_xx, t_xy, t_yy, t_xz, t_yz, t_zz)]
  constexpr int NDV = VF == F64_F1 || VF == F64_F2 ? 1 : VF == F64_F4 ? (LAP ? 7 : HES ? 10 : 4) : 3;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int r = lane & 15, q = lane >> 4;
  const int64_t p0 = ((int64_t)blockIdx.x * 4 + wave) * PPW;
  const int64_t sd = (int64_t)mp * ld;
  const bool split = gridDim.y > 1;
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
        m = kmask[t0 * nwords + w];
        if (t1 != t0) m |= kmask[t1 * nwords + w];
      }
      m >>= (k & 31);
      if (m) return k + __builtin_ctz(m);
      k = (w + 1) << 5;
    }
    return nk;
  };
  const int i_begin = split ? (int)blockIdx.y * 64 : 0;
  const int i_end = split ? i_begin + 64 : rows;
  dots += split ? (int64_t)blockIdx.y * mode_ndot(MODE) * ld : 0;

  // AO array of type t (0 = phi, 1..3 = d phi, 4 = lap phi | 4..9 = d_kl phi)
  auto arr = [](int t) {
    return VF == F64_F2 ? (t == 0 ? 0 : 4) : VF == F64_FH0 ? (t == 0 ? 0 : 3 + t) : VF == F64_FH1 ? (t == 0 ? 0 : 6 + t) : t;
  };
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
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int pb = 0; pb < NPB; ++pb) {
          const double y0 = acc[rb][pb][j];
          if constexpr (VF == F64_F1) {
            dacc[pb][0] += y0 * y0;
          } else if constexpr (VF == F64_F2) {
            dacc[pb][0] += y0 * acc[rb][NPB + pb][j];
          } else if constexpr (VF == F64_FH0 || VF == F64_FH1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) dacc[pb][k] += y0 * acc[rb][1 + k][j];
          } else {
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
        const int slot = VF == F64_F2    ? 4
                         : VF == F64_FH0 ? 4 + k
                         : VF == F64_FH1 ? 7 + k
                         : k < 4         ? k
                         : HES           ? k + 6
                                         : k + 1;
        dots[(int64_t)slot * ld + p0 + 16 * pb + r] = dacc[pb][k];
      }
  }
}

template <int VF, int MODE>
void launch_factor_var(const double* L, int rows, int mp, const double* ao, int64_t ld, double* dots, hipStream_t s,
                       bool split = false, const uint32_t* kmask = nullptr, int nwords = 0, int tp = 64) {
  constexpr int PPW = VF == F64_F1 ? 64 : VF == F64_F2 ? 32 : 16;
  const dim3 block(256), grid((unsigned)(ld / (4 * PPW)), split ? (unsigned)(rows / 64) : 1u);  // ld: multiple of 256
  if (kmask)
    hipLaunchKernelGGL((contract_f64_factor_kernel<VF, MODE, true>), grid, block, 0, s, L, rows, mp, ao, ld, dots,
                       kmask, nwords, tp);
  else
    hipLaunchKernelGGL((contract_f64_factor_kernel<VF, MODE, false>), grid, block, 0, s, L, rows, mp, ao, ld, dots,
                       kmask, nwords, tp);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace

namespace {
void factor_modes(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao, int64_t ld, double* dots,
                  hipStream_t s, bool split, const uint32_t* kmask = nullptr, int nwords = 0, int tp = 64) {
  CT_REQUIRE(rows_pad % 64 == 0 && rows_pad > 0 && ld % 256 == 0, "factor padding / point stride");
  CT_REQUIRE(!kmask || (nwords == (nbasis_pad / 32 + 31) / 32 && tp >= 16 && ld % tp == 0),
             "activity masks do not match the FP64 factor contraction");
  switch (mode) {
    case MODE_RHO:
      launch_factor_var<F64_F1, MODE_RHO>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp);
      break;
    case MODE_GRAD:
      launch_factor_var<F64_F4, MODE_GRAD>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp);
      break;
    case MODE_LAP:
      launch_factor_var<F64_F4, MODE_LAP>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp);
      launch_factor_var<F64_F2, MODE_LAP>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp);
      break;
    default:
      launch_factor_var<F64_F4, MODE_HESS>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp);
      launch_factor_var<F64_FH0, MODE_HESS>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp);
      launch_factor_var<F64_FH1, MODE_HESS>(L, rows_pad, nbasis_pad, ao, ld, dots, s, split, kmask, nwords, tp);
      break;
  }
}
}  // namespace

void launch_contract_f64_factor(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao,
                                int64_t ld, double* dots, hipStream_t s, const uint32_t* kmask, int nwords, int tp) {
  factor_modes(mode, L, rows_pad, nbasis_pad, ao, ld, dots, s, false, kmask, nwords, tp);
}

int64_t contract_f64_factor_split_workspace(int mode, int rows_pad, int64_t ld) {
  return (int64_t)(rows_pad / 64) * mode_ndot(mode) * ld;
}

void launch_contract_f64_factor_split(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao,
                                      int64_t ld, double* part, double* dots, hipStream_t s, const uint32_t* kmask,
                                      int nwords, int tp) {
  factor_modes(mode, L, rows_pad, nbasis_pad, ao, ld, part, s, true, kmask, nwords, tp);
  const int64_t count = (int64_t)mode_ndot(mode) * ld;
  hipLaunchKernelGGL(reduce_row_blocks_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, part,
                     rows_pad / 64, count, dots);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_contract_f64(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld, double* dots,
                         hipStream_t s) {
  launch_modes<false>(mode, P, nbasis_pad, ao, ld, dots, s);
}

int64_t contract_f64_split_workspace(int mode, int nbasis_pad, int64_t ld) {
  return (int64_t)(nbasis_pad / kIBlock) * mode_ndot(mode) * ld;
}

void launch_contract_f64_split(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld,
                               double* part, double* dots, hipStream_t s) {
  launch_modes<true>(mode, P, nbasis_pad, ao, ld, part, s);
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
