// FP64 per-point epilogue: turns the contraction dots into the requested
// descriptors in one pass (the reference runs separate host round trips and
// elementwise kernels per descriptor, src/evaluate_densbased.cu:44-223,
// src/cuda_utils.cu:11-136).  The arithmetic follows the reference exactly,
// including its masking conventions:
//   * masked divide: num/den when den > 1e-20, else 1e-20 (src/cuda_utils.cu:71-84)
//   * RDG  = 1/2 (3 pi^2)^{-1/3} |grad| / rho^{4/3} (src/evaluate_densbased.cu:44-89)
//   * W    = |grad|^2 / (8 rho), masked (src/evaluate_densbased.cu:91-140)
//   * TF   = 0.3 (3 pi^2)^{2/3} rho^{5/3} (src/evaluate_densbased.cu:142-173)
//   * GGE  = TF + a W + b lap (src/evaluate_densbased.cu:175-193)
//   * general(ish) KED = tau_PD + (a-1)/4 lap (src/evaluate_densbased.cu:195-208)
//   * t_alpha = tau_PD + alpha lap (src/evaluate_kinetic_dens.cu:195-214)
//   * Shannon = rho ln rho, unguarded (src/evaluate_densbased.cu:210-223)
//
// The mixed path's finaliser also sums the per-row-block partials, forms the
// error estimate gamma * |phi|^T |P| |phi| and appends the points whose
// estimate exceeds tol * |rho| to a list for FP64 re-evaluation.

#include "ct_internal.h"

namespace ct {

namespace {

__device__ __forceinline__ double masked_div(double num, double den) { return den > 1e-20 ? num / den : 1e-20; }

// dot[] layout: see mode_ndot() in ct_internal.h
__device__ void emit_outputs(const FinalizeArgs& a, int64_t o, const double* dot) {
  const double rho = dot[0];
  double g[3] = {0.0, 0.0, 0.0};
  if (a.mode != MODE_RHO)
    for (int k = 0; k < 3; ++k) g[k] = 2.0 * dot[1 + k];
  double lap = 0.0, tau = 0.0;
  if (a.mode == MODE_LAP) {
    const double t = dot[5] + dot[6] + dot[7];
    lap = 2.0 * (dot[4] + t);
    tau = 0.5 * t;
  } else if (a.mode == MODE_HESS) {
    const double t = dot[10] + dot[12] + dot[15];
    lap = 2.0 * (dot[4] + dot[7] + dot[9] + t);
    tau = 0.5 * t;
  }
  const uint32_t q = a.q;
  if (q & CT_Q_RHO) a.out.rho[o] = rho;
  // (n, 3) / (n, 3, 3) rows, or components grad_ld / hess_ld apart
  if (q & CT_Q_GRAD)
    for (int k = 0; k < 3; ++k) a.out.grad[a.grad_ld ? k * a.grad_ld + o : 3 * o + k] = g[k];
  if (q & CT_Q_HESS) {
    // h_kl slots 4..9 = xx,xy,xz,yy,yz,zz ; t_kl slots 10..15 = xx,xy,yy,xz,yz,zz
    const int hs[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};
    const int ts[3][3] = {{10, 11, 13}, {11, 12, 14}, {13, 14, 15}};
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) {
        const int c = 3 * k + l;
        a.out.hess[a.hess_ld ? c * a.hess_ld + o : 9 * o + c] = 2.0 * (dot[hs[k][l]] + dot[ts[k][l]]);
      }
  }
  if (q & CT_Q_LAP) a.out.lap[o] = lap;
  if (q & CT_Q_TAU_PD) a.out.tau_pd[o] = tau;
  const double gnorm2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  const double pi2x3 = 3.0 * M_PI * M_PI;
  if (q & CT_Q_RDG) {
    const double pref = 0.5 / pow(pi2x3, 1.0 / 3.0);
    a.out.rdg[o] = pref * masked_div(sqrt(gnorm2), pow(rho, 4.0 / 3.0));
  }
  if (q & CT_Q_SHANNON) a.out.shannon[o] = rho * log(rho);
  if (q & (CT_Q_WEIZSACKER | CT_Q_THOMAS_FERMI | CT_Q_GGE)) {
    const double tf = 0.3 * pow(pi2x3, 2.0 / 3.0) * pow(rho, 5.0 / 3.0);
    const double w = masked_div(gnorm2, 8.0 * rho);
    if (q & CT_Q_WEIZSACKER) a.out.weizsacker[o] = w;
    if (q & CT_Q_THOMAS_FERMI) a.out.thomas_fermi[o] = tf;
    if (q & CT_Q_GGE) a.out.gge[o] = tf + a.par.gge_a * w + a.par.gge_b * lap;
  }
  if (q & CT_Q_GENERAL_KED) a.out.general_ked[o] = tau + (a.par.ked_a - 1.0) / 4.0 * lap;
  if (q & CT_Q_KED_ALPHA) a.out.ked_alpha[o] = tau + a.par.ked_alpha * lap;
}

__global__ __launch_bounds__(256) void finalize_kernel(FinalizeArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const int nd = mode_ndot(a.mode);
  // gather (screened layout): output point p reads its dots at its sorted
  // position; else the dots of p go to output index[p] (or p)
  const int64_t src = a.gather ? (int64_t)a.gather[p] : p;
  double dot[16];
  for (int d = 0; d < 16; ++d) dot[d] = d < nd ? a.dots[(int64_t)d * a.ld + src] : 0.0;
  const int64_t o = a.out_off + (a.gather ? p : a.index ? (int64_t)a.index[p] : p);
  emit_outputs(a, o, dot);
}

__global__ __launch_bounds__(256) void finalize_mixed_kernel(FinalizeArgs a, FlagArgs f) {
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // lanes past the end run on the last point (no stores) so that the whole
  // wave takes part in the reduction of the largest estimate
  const bool live = p0 < a.n;
  const int64_t p = live ? p0 : a.n - 1;
  const int nd = mode_ndot(a.mode);
  const int nd1 = nd + 1;
  double dot[16];
  for (int d = 0; d < 16; ++d) dot[d] = 0.0;
  double bound = 0.0;
  // gather: p is an output point (coalesced stores), its dots lie at the
  // sorted position gather[p] (reads from the contraction's fresh output);
  // else the dots of p go to output index[p]
  const int64_t src = a.gather ? (int64_t)a.gather[p] : p;
  for (int rt = 0; rt < a.nrt; ++rt) {
    const double* base = a.dots + (int64_t)rt * nd1 * a.ld + src;
    for (int d = 0; d < 16; ++d)
      if (d < nd) dot[d] += base[(int64_t)d * a.ld];
    bound += base[(int64_t)nd * a.ld];
  }
  // the chunk position of the output
  const int64_t pp = a.gather ? p : a.index ? (int64_t)a.index[p] : p;
  if (live) emit_outputs(a, a.out_off + pp, dot);
  const double arho = fabs(dot[0]);
  // what the error is held against: |rho|, or the residual |rho - target|
  const double ref = f.target ? fabs(dot[0] - f.target[pp]) : arho;
  double rel = 0.0;  // estimate / ref of an accepted point (non-negative: the bits order like the values)
  bool flag = false;
  if (live && !(arho < f.rho_floor)) {
    const double est = f.gamma * bound + f.trunc * sqrt(arho);
    if (!(est <= f.tol * ref)) {  // also catches NaN
      flag = true;
    } else if (ref > 0.0) {
      rel = est / ref;
    }
  }
  // flagged points: one atomic per wave for its count, each lane at its rank
  const uint64_t fm = __ballot(flag);
  if (fm) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)fm) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(f.count, __popcll(fm));
    base = __shfl(base, leader);
    if (flag) f.index[base + __popcll(fm & ((1ull << lane) - 1ull))] = (int32_t)(f.index_base + pp);
  }
  // largest estimate: reduced per wave, and the wave's atomic only when it
  // exceeds the value already stored (read first).  Atomics on one address
  // serialise at the L2; most waves' maxima stay below the running maximum
  // and end with the read.  A stale read only costs an unneeded atomic.
  for (int off = 32; off >= 1; off >>= 1) rel = fmax(rel, __shfl_xor(rel, off));
  if ((threadIdx.x & 63) == 0 && rel > 0.0) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(rel);
    if (bits > __hip_atomic_load(f.max_est, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(f.max_est, bits);
  }
}

__global__ __launch_bounds__(256) void invert_permutation_kernel(const int32_t* __restrict__ perm, int64_t n,
                                                                 int32_t* __restrict__ inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[perm[i]] = (int32_t)i;
}

// col_stride = 0: pts is (m, 3) row-major; else x, y, z of point p are at
// pts[p], pts[col_stride + p], pts[2 col_stride + p]
__global__ __launch_bounds__(256) void gather_points_kernel(const double* __restrict__ pts, int64_t col_stride,
                                                            const int32_t* __restrict__ index, int64_t n,
                                                            double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t p = index ? (int64_t)index[k] : k;
#pragma unroll
  for (int c = 0; c < 3; ++c) out[3 * k + c] = col_stride ? pts[c * col_stride + p] : pts[3 * p + c];
}

__global__ __launch_bounds__(256) void cubic_points_kernel(CubicGrid g, int64_t start, int64_t n,
                                                           double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t lin = start + k;
  const double iz = (double)(lin % g.shape[2]);
  const double iy = (double)((lin / g.shape[2]) % g.shape[1]);
  const double ix = (double)(lin / (g.shape[2] * g.shape[1]));
#pragma unroll
  for (int c = 0; c < 3; ++c) out[3 * k + c] = g.origin[c] + ix * g.axes[c] + iy * g.axes[3 + c] + iz * g.axes[6 + c];
}

__global__ __launch_bounds__(256) void norm_rows_kernel(const double* __restrict__ v, int64_t n,
                                                        double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x = v[3 * p], y = v[3 * p + 1], z = v[3 * p + 2];
  out[p] = sqrt(x * x + y * y + z * z);
}

__global__ __launch_bounds__(256) void fill_zero_kernel(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

void launch_fill_zero(unsigned long long* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_finalize(const FinalizeArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_finalize_mixed(const FinalizeArgs& a, const FlagArgs& f, hipStream_t s) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(finalize_mixed_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a, f);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_gather_points(const double* pts, const int32_t* index, int64_t n, double* out, hipStream_t s,
                          int64_t col_stride) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gather_points_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pts, col_stride, index,
                     n, out);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_invert_permutation(const int32_t* perm, int64_t n, int32_t* inv, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(invert_permutation_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, n, inv);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_cubic_points(const CubicGrid& g, int64_t start, int64_t n, double* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(cubic_points_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, start, n, out);
  CT_HIP_CHECK(hipGetLastError());
}

void launch_norm_rows(const double* v, int64_t n, double* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(norm_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, n, out);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
