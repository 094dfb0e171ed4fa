// Adjoint of rho (no reference counterpart; used by the autograd wrapper):
//
//   dL/dP_ij  = sum_g w_g phi_i(g) phi_j(g)                 (w_g = dL/drho_g)
//   dL/dR_A   = -2 sum_{i on A} sum_g w_g grad phi_i(g) (P phi(g))_i
//
// Two FP64 matrix-core kernels per point chunk:
//   * gram_f64: G = Phi diag(w) Phi^T, split over point slices.  One wave owns
//     a 64 x 64 tile of the upper triangle of G (over the row blocks holding
//     real functions) and one slice of points (the GEMM's K); each slice
//     writes its own partial tile and, off the diagonal, its transpose, and a
//     reduction adds the slices in a fixed order (deterministic).
//   * contract_adjoint_f64: the forward contraction's GEMM Y = P phi with an
//     epilogue that sums w_g d_k phi_i(g) Y_i(g) over the wave's points per
//     row i and adds the row totals into T[i][k] with FP64 atomics (the order
//     of those few additions per row is not fixed: results agree to the last
//     bits only).
// The per-atom sums of -2 T happen on the host.

#include "ct_internal.h"

namespace ct {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kGramSlice = 2048;  // points per partial (K of one wave)

// MFMA 16x16x4 f64 layout (lane l, r = l % 16, q = l / 16): A[r][q], B[q][r],
// D[q + 4j][r].
// One k-step covers 8 points: lane q reads points p0 + 2q and p0 + 2q + 1 of
// each of its rows as one 16-byte load and feeds them to two MFMAs (the
// point order inside a slice is fixed, so the sums are deterministic).  The
// next k-step's loads are issued before the current step's MFMAs.  Points
// at or beyond the slice end contribute zero (the AO buffer holds ld >= the
// rounded-up slice end columns; w is read only below npts).
__global__ __launch_bounds__(64, 2) void gram_f64_kernel(const double* __restrict__ ao, int mp, int64_t ld,
                                                      const double* __restrict__ w, int64_t npts,
                                                      double* __restrict__ part, int ntile) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x;
  const int r = lane & 15, q = lane >> 4;
  const int ntri = ntile * (ntile + 1) / 2;
  int t = blockIdx.x % ntri;
  const int slice = blockIdx.x / ntri;
  // (row block a, column block b >= a) of the t-th upper-triangle tile
  int ra = 0;
  while (t >= ntile - ra) {
    t -= ntile - ra;
    ++ra;
  }
  const int I = ra * 64, J = (ra + t) * 64;
  const int64_t pbeg = (int64_t)slice * kGramSlice;
  const int64_t pend = pbeg + kGramSlice < npts ? pbeg + kGramSlice : npts;
  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const double* arow = ao + (int64_t)(I + r) * ld + 2 * q;
  const double* brow = ao + (int64_t)(J + r) * ld + 2 * q;
  const int64_t r16 = 16 * ld;
  d2 a[4], b[4];
  auto load = [&](int64_t p, d2* av, d2* bv) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      av[k] = *reinterpret_cast<const d2*>(arow + k * r16 + p);
      bv[k] = *reinterpret_cast<const d2*>(brow + k * r16 + p);
    }
  };
  load(pbeg, a, b);
  for (int64_t p0 = pbeg; p0 < pend; p0 += 8) {
    const int64_t pn = p0 + 8 < pend ? p0 + 8 : p0;
    d2 an[4], bn[4];
    load(pn, an, bn);
    const int64_t pa = p0 + 2 * q;
    const bool ok0 = pa < pend, ok1 = pa + 1 < pend;
    const double w0 = ok0 ? w[pa] : 0.0, w1 = ok1 ? w[pa + 1] : 0.0;
    double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a0[k] = ok0 ? a[k][0] : 0.0;
      a1[k] = ok1 ? a[k][1] : 0.0;
      b0[k] = ok0 ? w0 * b[k][0] : 0.0;
      b1[k] = ok1 ? w1 * b[k][1] : 0.0;
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[rb], b0[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[rb], b1[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = an[k];
      b[k] = bn[k];
    }
  }
  double* out = part + (int64_t)slice * mp * mp;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t i = I + 16 * rb + q + 4 * j, jj = J + 16 * cb + r;
        out[i * mp + jj] = acc[rb][cb][j];
        if (I != J) out[jj * mp + i] = acc[rb][cb][j];
      }
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(const double* __restrict__ part, int nslice, int64_t count,
                                                          double* __restrict__ G) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  double s = G[k];
  for (int sl = 0; sl < nslice; ++sl) s += part[(int64_t)sl * count + k];
  G[k] = s;
}

// Y = P phi (64 points per wave pair, like the forward MODE_GRAD kernel) with
// the row-wise weighted epilogue.  Waves 2h and 2h + 1 of a workgroup share
// 64 points and take alternate 64-row blocks of P; their rows are disjoint,
// so each adds its own row totals into T.  The k-loop prefetches one k-step
// ahead; the epilogue loads the 12 AO values of two rows per lane as one
// batch before their products.
__global__ __launch_bounds__(256, 2) void contract_adjoint_f64_kernel(const double* __restrict__ P, int mp, int m,
                                                                      const double* __restrict__ ao, int64_t ld,
                                                                      const double* __restrict__ w, int64_t npts,
                                                                      double* __restrict__ T) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  const int half = wave & 1;
  const int64_t p0 = ((int64_t)blockIdx.x * 2 + (wave >> 1)) * 64;
  const int64_t sd = (int64_t)mp * ld;
  const double* bptr[4];
  double wl[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    bptr[cb] = ao + (int64_t)q * ld + p0 + 16 * cb + r;
    const int64_t p = p0 + 16 * cb + r;
    wl[cb] = p < npts ? w[p] : 0.0;
  }
  // P's rows and columns beyond M are zero: row blocks up to the last real
  // row, K up to the last real column (a multiple of 4)
  const int mr = (m + 63) / 64 * 64, mk = (m + 3) / 4 * 4;
  for (int i0 = 64 * half; i0 < mr; i0 += 128) {
    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    const double* aptr = P + (int64_t)q * mp + i0 + r;
    double a[4], b[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) a[rb] = aptr[16 * rb];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) b[cb] = bptr[cb][0];
    for (int k0 = 0; k0 < mk; k0 += 4) {
      // prefetch the next k-step (clamped on the last one; the value is unused)
      const int kn = k0 + 4 < mk ? k0 + 4 : k0;
      double an[4], bn[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) an[rb] = aptr[(int64_t)kn * mp + 16 * rb];
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
    // row sums over the pair's 64 points: lane (q, r) holds rows i0+16rb+q+4j
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
      for (int j0 = 0; j0 < 4; j0 += 2) {
        double v[2][4][3];
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int64_t i = i0 + 16 * rb + q + 4 * (j0 + jj);
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int k = 0; k < 3; ++k) v[jj][cb][k] = ao[(1 + k) * sd + i * ld + p0 + 16 * cb + r];
        }
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int64_t i = i0 + 16 * rb + q + 4 * (j0 + jj);
          double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
          for (int cb = 0; cb < 4; ++cb) {
            const double y = acc[rb][cb][j0 + jj] * wl[cb];
#pragma unroll
            for (int k = 0; k < 3; ++k) s[k] += v[jj][cb][k] * y;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double t = s[k];
            t += __shfl_xor(t, 1);
            t += __shfl_xor(t, 2);
            t += __shfl_xor(t, 4);
            t += __shfl_xor(t, 8);
            s[k] = t;
          }
          if (r == 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) atomicAdd(&T[i * 3 + k], s[k]);
          }
        }
      }
    }
  }
}

// dL/dR_A = -2 sum of T[i] over the functions i on atom A (T[i][k] = sum_g
// w_g d_k phi_i (P phi)_i), in function order: thread 3 A + k.
__global__ __launch_bounds__(64) void atom_gradient_kernel(const double* __restrict__ T,
                                                           const int32_t* __restrict__ func_atom, int m, int natom,
                                                           double* __restrict__ dR) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * natom) return;
  const int a = t / 3, k = t % 3;
  double s = 0.0;
  for (int i = 0; i < m; ++i)
    if (func_atom[i] == a) s -= 2.0 * T[3 * i + k];
  dR[t] = s;
}

}  // namespace

int gram_slices(int64_t npts) { return (int)((npts + kGramSlice - 1) / kGramSlice); }

void launch_gram_f64(const double* ao, int mp, int m, int64_t ld, const double* w, int64_t npts, double* part,
                     hipStream_t s) {
  CT_REQUIRE(m > 0 && m <= mp, "basis size");
  const int ntile = (m + 63) / 64;  // tiles over the row blocks with a real function
  const int nslice = gram_slices(npts);
  hipLaunchKernelGGL(gram_f64_kernel, dim3((unsigned)(ntile * (ntile + 1) / 2 * nslice)), dim3(64), 0, s, ao, mp, ld,
                     w, npts, part, ntile);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_gram_reduce(const double* part, int nslice, int64_t count, double* G, hipStream_t s) {
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, part, nslice, count,
                     G);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_atom_gradient(const double* T, const int32_t* func_atom, int m, int natom, double* dR, hipStream_t s) {
  if (natom <= 0) return;
  hipLaunchKernelGGL(atom_gradient_kernel, dim3((unsigned)((3 * natom + 63) / 64)), dim3(64), 0, s, T, func_atom, m,
                     natom, dR);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_contract_adjoint_f64(const double* P, int mp, int m, const double* ao, int64_t ld, const double* w,
                                 int64_t npts, double* T, hipStream_t s) {
  CT_REQUIRE(m > 0 && m <= mp, "basis size");
  // 128 points per workgroup (two wave pairs); ld is a multiple of 256
  hipLaunchKernelGGL(contract_adjoint_f64_kernel, dim3((unsigned)(ld / 128)), dim3(256), 0, s, P, mp, m, ao, ld, w,
                     npts, T);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
