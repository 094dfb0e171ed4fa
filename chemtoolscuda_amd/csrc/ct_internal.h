// Internal declarations shared by the host orchestration and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "chemtools_amd.h"

namespace ct {

// ---------------------------------------------------------------------------
// Errors: every failure becomes a ct::Error carrying a CT_E_* status; the C
// ABI converts it to a return code and a thread-local message.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define CT_HIP_CHECK(expr)                                                                     \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::ct::Error(CT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + \
                                      __FILE__ + ":" + std::to_string(__LINE__) + ")");        \
  } while (0)

#define CT_REQUIRE(cond, msg)                                  \
  do {                                                         \
    if (!(cond)) throw ::ct::Error(CT_E_INVALID, (msg));       \
  } while (0)

// ---------------------------------------------------------------------------
// Device shell table.  Shell type codes index the generated SHELL_TYPES list
// (0,1,2,3,4 Cartesian; 5,6,7 = pure d,f,g).
int shell_code(int l);
int shell_nfunc(int l);

struct ShellMeta {  // read with wave-uniform (scalar) loads
  int32_t code;
  int32_t nprim;
  int32_t prim_off;
  int32_t first;  // first basis function of the shell
};

struct ShellGeom {
  double cx, cy, cz;
  double amin;  // smallest exponent of the shell (screening)
};

// Point-independent primitive data: exponent a and c_p * N_p(L), where
// N_p(L) = sqrt((2a/pi)^{3/2} (4a)^L) is the radial part of the L2
// normalisation (the angular part is folded into the generated polynomials).
struct PrimData {
  double a;
  double c;
};

// Error model of the mixed path: |rho_mixed - rho| <= gamma * |phi|^T |P| |phi|
// (u = 2^-24).  Rounding the FP64-computed AOs and P to FP32 contributes
// ~2.1u |phi|^T|P||phi|, the terms the three-piece bf16 split drops <= 2u in
// the worst case (~0 in practice), FP32 accumulation the rest.  Largest ratio
// error / (u |phi|^T|P||phi|) measured on the 25-, 624- and 5400-function
// workloads: 3.3 with FP64 running sums (rho / grad / lap modes), 6.8 with the
// two FP32 accumulators of the Hessian mode; gamma leaves a margin of 1.8 and
// 1.2 over them.
constexpr double kMixedGammaF64Sums = 3.5762786865234375e-07;  // 6u
constexpr double kMixedGammaF32Sums = 4.76837158203125e-07;    // 8u
// Factor route (P = L L^T, psi = L^T phi, rho = psi . psi): the bound is
// B' = 2 sum_j |psi_j| (|L|^T |phi|)_j >= 2 rho, the first-order propagation
// of the per-column error of psi; largest ratio error / (u B') measured on the
// same workloads: 2.0, so gamma = 3u leaves a margin of 1.5.
constexpr double kMixedGammaFactor = 1.78813934326171875e-07;  // 3u


// Granularities.
constexpr int kRowBlock = 128;     // basis padding: rows of P per contraction block
constexpr int kPointAlign = 256;   // chunk stride alignment (points)

// Evaluation modes of the fused contraction (which GEMM columns and dots).
enum Mode : int { MODE_RHO = 0, MODE_GRAD = 1, MODE_LAP = 2, MODE_HESS = 3 };
inline double mixed_gamma(int mode) { return mode == MODE_HESS ? kMixedGammaF32Sums : kMixedGammaF64Sums; }
// AO arrays evaluated per mode: val; val+grad; val+grad+lap; val+grad+hess.
constexpr int mode_nder(int m) { return m == MODE_RHO ? 1 : m == MODE_GRAD ? 4 : m == MODE_LAP ? 5 : 10; }
// Per-point dot products produced by the contraction epilogue.
//   RHO : [rho]
//   GRAD: [rho, gx, gy, gz]                          g_k = d_k phi . P phi
//   LAP : [rho, gx, gy, gz, L0, txx, tyy, tzz]       L0 = lap phi . P phi,  t_kk = d_k phi . P d_k phi
//   HESS: [rho, gx, gy, gz, hxx,hxy,hxz,hyy,hyz,hzz, txx,txy,tyy,txz,tyz,tzz]
//         h_kl = d_kl phi . P phi,  t_kl = d_k phi . P d_l phi
constexpr int mode_ndot(int m) { return m == MODE_RHO ? 1 : m == MODE_GRAD ? 4 : m == MODE_LAP ? 8 : 16; }

// AO layouts.  FP64 path: ao[(d * mp + i) * ld + p] (points contiguous).
// Mixed path (FP32, "row quads"): four consecutive basis functions of one
// point are 16 contiguous bytes, quads of one array and row block follow each
// other point by point, so a 16-byte load yields four rows of one point.
__host__ __device__ inline int64_t ao_quad_index(int d, int i, int64_t p, int mp, int64_t ld) {
  return (((int64_t)d * (mp >> 2) + (i >> 2)) * ld + p) * 4 + (i & 3);
}
// Mixed path, screened factor route ("tile-packed" row quads): the chunk's
// sorted points in tiles of tp, and per (tile, 32-function k-step) one
// contiguous block of nder arrays x 8 quads x tp points x 4 floats, so the
// block a contraction tile reads per k-step (and the AO kernel writes for it)
// is one region (lap mode: 5 x 8 x 16 x 16 B = 10 KiB).  Offset in floats of
// quad qq (0..7) of k-step k, array d, point pp of tile `tile`; nk = mp / 32.
__host__ __device__ inline int64_t ao_packed_index(int64_t tile, int k, int d, int qq, int pp, int nk, int nder,
                                                   int tp) {
  return ((((tile * nk + k) * nder + d) * 8 + qq) * (int64_t)tp + pp) * 4;
}

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files).
struct FinalizeArgs {
  const double* dots;     // FP64 path: [ndot][ld];  mixed path: partials [nrt][ndot+1][ld]
  const int32_t* index;   // optional: output index of point p is out_off + index[p] (FP64: the
                          // re-evaluated points)
  const int32_t* gather;  // mixed, optional (index null): the dots of output point p are at
                          // position gather[p] (the points in Morton order: the inverse
                          // of their permutation; outputs written in order)
  int nrt;                // mixed path: number of 128-row blocks to sum
  int64_t ld;
  int64_t n;        // points in this chunk
  int64_t out_off;  // index of the chunk's first point in the output arrays
  int mode;
  uint32_t q;
  ct_eval_params par;
  ct_outputs out;   // device pointers (full-length arrays, indexed with out_off)
  int64_t grad_ld;  // 0: grad (n, 3) rows; else component k of output point o at grad[k * grad_ld + o]
  int64_t hess_ld;  // 0: hess (n, 3, 3) rows; else component 3k+l at hess[(3k + l) * hess_ld + o]
};

// kmask (optional): as for launch_ao_f32 (a wave skips, stores included, the
// shells no tile of its points reads).
void launch_ao_f64(int mode, int lmax, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                   int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, double* ao,
                   double screen, hipStream_t s, const uint32_t* kmask = nullptr, int nwords = 0,
                   int tile_points = 64, int rows_out = 0);
void launch_contract_f64(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld, double* dots,
                         hipStream_t s);
void launch_finalize(const FinalizeArgs& a, hipStream_t s);
// The same FP64 dots from a factor P = L L^T (L: nbasis_pad x rows_pad,
// k-major, rows_pad a multiple of 64): rank(P) columns instead of nbasis.
// kmask (optional, screened layout): per tile of
// tile_points points the active 32-function k-steps (launch_tile_activity);
// only those are contracted.  psi (lap and Hessian modes): rows_pad x ld
// doubles, psi = L^T phi of the mode's first launch, read by the others.
void launch_contract_f64_factor(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao,
                                int64_t ld, double* dots, hipStream_t s, const uint32_t* kmask = nullptr,
                                int nwords = 0, int tile_points = 64, double* psi = nullptr);
// The same with each 64-column pass of the factor in its own workgroups and
// an ordered sum of the passes' partial dots (for small point counts); part
// holds contract_f64_factor_split_workspace() doubles.
int64_t contract_f64_factor_split_workspace(int mode, int rows_pad, int64_t ld);
void launch_contract_f64_factor_split(int mode, const double* L, int rows_pad, int nbasis_pad, const double* ao,
                                      int64_t ld, double* part, double* dots, hipStream_t s,
                                      const uint32_t* kmask = nullptr, int nwords = 0, int tile_points = 64,
                                      double* psi = nullptr);
// The same contraction with one 64-row block of P per wave and an ordered
// sum of the blocks' partial dots (for small point counts); part holds
// contract_f64_split_workspace() doubles.
int64_t contract_f64_split_workspace(int mode, int nbasis_pad, int64_t ld);
void launch_contract_f64_split(int mode, const double* P, int nbasis_pad, const double* ao, int64_t ld,
                               double* part, double* dots, hipStream_t s);

// Mixed path.
// kmask (optional, screened layout): the activity masks of the point tiles of
// tile_points points (launch_tile_activity); a wave skips a shell, stores
// included, when no k-step the shell touches is active for its points.
void launch_ao_f32(int mode, int lmax, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                   int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, float* ao, double screen,
                   const uint32_t* kmask, int nwords, int tile_points, hipStream_t s, int rows_out = 0);
// Reorders three row-major bf16 planes of a rows x cols matrix
// ([3][rows][cols]; rows a multiple of 16, cols of 32) into the MFMA fragment
// order of the mixed tile engine's A operand (host side, same size).
void mixed_fragment_order(const uint16_t* planes, int rows, int cols, uint16_t* out);
void launch_contract_mixed(int mode, const uint16_t* Pp, int nbasis_pad, const float* ao, int64_t ld, double* part,
                           hipStream_t s);
// The same partial dots from a factor P = L L^T (ct_eval_params.route): Lf
// holds the bf16 pieces of L^T (rows_pad x nbasis_pad, fragment order), and
// the partials run over rows_pad / 128 blocks of the factor's columns.
// Modes rho, grad and lap.
// k-steps: screened, klist / kcount from launch_tile_activity at
// factor_tile_points(mode), only the listed k-steps are contracted; else
// kcount = nullptr and klist = 0 .. nbasis_pad / 32 - 1 for every tile.
// psi (Hessian mode only): rows_pad x ld doubles, psi = L^T phi of the
// first launch, read by the second-derivative launches.
void launch_contract_mixed_factor(int mode, const uint16_t* Lf, int rows_pad, int nbasis_pad, const float* ao,
                                  int64_t ld, double* part, const int32_t* klist, const int32_t* kcount,
                                  hipStream_t s, double* psi = nullptr);
// Pivoted Cholesky factorisation of the symmetric n x n matrix A (row stride
// lda) on the host: A = L L^T with L (n x rank, column k at L[k * n ...]),
// columns in pivot order, stopping when the largest remaining diagonal
// element is <= rel_tol times the largest diagonal element of A.  Returns
// false when the remaining diagonal is negative beyond that tolerance or the
// rank exceeds rmax.  A remainder with a small diagonal can still be
// indefinite (e.g. a zero diagonal with non-zero off-diagonal elements): the
// caller checks the off-diagonal residual (launch_factor_residual).
bool pivoted_cholesky(const double* A, int n, int lda, int rmax, double rel_tol, std::vector<double>& L, int& rank);
// *out = max(*out, largest violation of |R_ij| <= sqrt(R_ii R_jj) over
// i, j < m), R = A - L L^T (A: mp x mp row-major; row i of L at L[i * rows]
// with rank columns; diag[i] = R_ii), as the bits of a non-negative double
// (NaN -> 1e300).
void launch_factor_residual(const double* A, int mp, const double* L, int rows, int rank, int m, const double* diag,
                            unsigned long long* out, hipStream_t s);

// ---------------------------------------------------------------------------
// Spatial screening of the factor route (screen_host.cpp, screen.hip).
// sqrt((2a/pi)^{3/2} (4a)^L): radial part of a primitive's L2 normalisation.
double prim_radial_norm(double a, int L);
// Distance from the centre beyond which the value, every first and second
// derivative and the Laplacian of every function of the shell are at most
// eps in magnitude (cn[p] = c_p N_p; rigorous bound, see screen_host.cpp).
double shell_screen_radius(int l, int nprim, const double* a, const double* cn, double eps);
struct ScreenSpan {  // one shell of a k-step: centre and squared cutoff radius
  double cx, cy, cz, r2;
};
// Shell order of the screened AO layout and what the activity kernel reads.
struct ScreenLayout {
  std::vector<int> order;     // screened position -> shell
  std::vector<int> first;     // shell -> first function in the screened layout
  std::vector<int> func_src;  // screened function -> function of the natural layout
  std::vector<int> span_off;  // k-step kb -> spans[span_off[kb] .. span_off[kb + 1])
  std::vector<ScreenSpan> spans;
};
ScreenLayout build_screen_layout(const std::vector<int>& shell_l, const std::vector<int>& shell_atom,
                                 const std::vector<double>& radius, const double* xyz, int nbasis_pad);
// Refills span_off / spans of a layout (its order and function offsets kept)
// from the given centres and radii: after the centres move by less than what
// warrants a new order.
void screen_layout_spans(ScreenLayout& out, const std::vector<int>& shell_l, const std::vector<int>& shell_atom,
                         const std::vector<double>& radius, const double* xyz, int nbasis_pad);
// Lattice of the Morton keys: cell (floor((x - origin) * inv_cell)), 8 bits
// per axis, clamped.
struct MortonFrame {
  double origin[3];
  double inv_cell;
};
size_t screen_sort_workspace(int64_t n);
// ascending sort of n non-negative int32 below 2^bits (the flagged points of
// a call)
size_t sort_indices_workspace(int64_t n, int bits);
void launch_sort_indices(const int32_t* in, int32_t* out, int64_t n, int bits, void* temp, size_t temp_bytes,
                         hipStream_t s);
// perm[i] = index of the i-th point in Morton order (stable)
void launch_screen_sort(const double* pts, int64_t n, const MortonFrame& f, uint32_t* keys_in, uint32_t* keys_out,
                        int32_t* idx_in, int32_t* perm, void* temp, size_t temp_bytes, hipStream_t s);
// Per tile of tile_points consecutive (sorted) points: active k-steps
// klist[tile * nk ..] (kcount[tile] of them, increasing), mask
// kmask[tile * nwords ..], nwords = ceil(nk / 32); the counts are added to
// active_count[16 c], c = tile % kScreenCounters (sum them for the total).
constexpr int kScreenCounters = 64;
void launch_tile_activity(const double* sorted_pts, int64_t n, int tile_points, int64_t ntile, const int32_t* span_off,
                          const ScreenSpan* spans, int nk, int32_t* klist, int32_t* kcount, uint32_t* kmask,
                          int nwords, unsigned long long* active_count, hipStream_t s);
// Points per tile of the factor route's contraction in this mode (the
// granularity of the activity lists).
int factor_tile_points(int mode);
struct FlagArgs {
  double gamma;          // error estimate = gamma * B (B = |phi|^T |P| |phi|)
  double tol;            // relative tolerance on rho
  double rho_floor;      // points with |rho| below this are not re-evaluated
  int32_t* count;        // number of flagged points (appended to, not reset, across chunks)
  int32_t* index;        // flagged point indices
  int64_t index_base;    // added to the chunk-relative index of a flagged point
  unsigned long long* max_est;  // max estimate / |rho| over accepted points (bit pattern of a double)
  double trunc;          // screened factor route: + trunc * sqrt|rho| (bound of the skipped AO blocks)
  const double* target;  // optional, chunk positions: the estimate is held against tol * |rho - target|
};
void launch_finalize_mixed(const FinalizeArgs& a, const FlagArgs& f, hipStream_t s);
// out (n, 3) row-major: the points index[k] (index == nullptr: k) of pts,
// which is (m, 3) row-major (col_stride = 0) or column-major with x, y, z of
// point p at pts[p], pts[col_stride + p], pts[2 col_stride + p]
void launch_gather_points(const double* pts, const int32_t* index, int64_t n, double* out, hipStream_t s,
                          int64_t col_stride = 0);
// inv[perm[i]] = i for i < n
void launch_invert_permutation(const int32_t* perm, int64_t n, int32_t* inv, hipStream_t s);

// Uniform (possibly skewed) grid: point (i, j, k) = origin + i a0 + j a1 + k a2,
// a_r = axes[3r .. 3r+2]; linear index (i * n1 + j) * n2 + k (k fastest).
struct CubicGrid {
  double origin[3];
  double axes[9];
  int64_t shape[3];
};
// Writes the points of linear indices start .. start+n-1 as (n, 3) row-major.
void launch_cubic_points(const CubicGrid& g, int64_t start, int64_t n, double* out, hipStream_t s);
void launch_norm_rows(const double* v, int64_t n, double* out, hipStream_t s);
// zeroes n 64-bit words (a kernel: no memset nodes in work a caller may capture)
void launch_fill_zero(unsigned long long* p, int64_t n, hipStream_t s);

// Adjoint.
int gram_slices(int64_t npts);
// tiles over the row blocks holding the m real functions only (partials of
// G's other entries are left unwritten)
void launch_gram_f64(const double* ao, int mp, int m, int64_t ld, const double* w, int64_t npts, double* part,
                     hipStream_t s);
void launch_gram_reduce(const double* part, int nslice, int64_t count, double* G, hipStream_t s);
void launch_contract_adjoint_f64(const double* P, int mp, int m, const double* ao, int64_t ld, const double* w,
                                 int64_t npts, double* T, hipStream_t s);
// dR (3 natom doubles, device) = -2 sum of T[i] (3 m doubles) over the
// functions i with func_atom[i] = A
void launch_atom_gradient(const double* T, const int32_t* func_atom, int m, int natom, double* dR, hipStream_t s);
// Mixed adjoint (FP32 row-quad AOs, split-bf16 MFMA, FP64 partials).
// Sustained rates of the device (peak.hip): bf16 and FP64 matrix cores,
// FP64 vector FMA (TFLOP/s), HBM copy (GB/s, read + write).
void measure_peaks(double* bf16_tflops, double* f64_mfma_tflops, double* f64_valu_tflops, double* hbm_gbs);
// Point slices of the mixed Gram kernel for npts points (at most max_slices):
// the number of slices and the points per slice (a multiple of 32).
void gram_mixed_plan(int64_t npts, int mp, int* nslice, int64_t* slice_len, int max_slices = 64);
void launch_gram_mixed(const float* ao, int mp, int m, int64_t ld, const double* w, int64_t npts, const int32_t* tiles,
                       int nslice, int64_t slice_len, double* part, hipStream_t s);
void launch_contract_adjoint_mixed(const uint16_t* Pf, int mp, int m, const float* ao, int64_t ld, const double* w,
                                   int64_t npts, double* T, hipStream_t s);

}  // namespace ct
