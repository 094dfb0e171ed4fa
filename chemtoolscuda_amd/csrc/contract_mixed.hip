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
// columns (the variants below).  The epilogue dots each wave's rows with the
// matching AO row quads (one 16-byte load per quad) in FP64, reduces over
// lane groups and the four waves in a fixed order, and writes one partial per
// (row block, dot, point); the finaliser sums the row blocks.

#include <type_traits>

#include "mixed_tile.h"

namespace ct {

namespace {

using namespace mixed;

// Tile variants.  A workgroup owns a 128-row block of A and TC columns.
// Density route, A = P:
//   V_RHO:         64 points of phi, triangle of P only (y_tile's TRI: rho and
//                  the bound are quadratic forms);
//   V_GRAD:        64 points of phi, full K;
//   V_LAPPHI:      64 points of phi, full K -- the lap mode's rho, grad and
//                  lap phi . P phi dots (and the bound);
//   V_LAPDER<d>:   128 points of one derivative column type d (1..3), triangle
//                  of P only -- the lap mode's tau_dd = d_d phi^T P d_d phi;
//   V_HESS:        32 points x {phi, d_x phi, d_y phi, d_z phi}, full K (the
//                  t_kl with k != l are not quadratic forms).
// Factor route, A = L^T with P = L L^T, so that the tile holds psi = L^T phi
// and its derivatives for 128 columns of L, and every dot is a lane-local
// sum over those columns (phi^T P x = psi^T (L^T x)):
//   V_F1:          64 points of phi: rho = psi . psi and the bound;
//   V_F4<mode>:    32 points x {phi, d_x phi, d_y phi, d_z phi}: rho,
//                  g_k = psi . d_k psi and the bound (grad mode), and
//                  t_kl = d_k psi . d_l psi (Hessian mode);
//   V_F5:          16 points x {phi, d_x phi, d_y phi, d_z phi, lap phi}: the
//                  whole lap mode in one launch (rho, g_k, t_kk = d_k psi .
//                  d_k psi, psi . lap psi and the bound);
//   V_FH<part>:    32 points x {d_kl phi}: the Hessian mode's h_kl =
//                  psi . d_kl psi (part 0: xx, xy, xz, yy; part 1: yz, zz)
//                  with psi as V_F4<hess> stored it (the same FP64 running
//                  sums these launches would form again); with V_F4<hess>,
//                  whose t_kl = d_k psi . d_l psi cover all six kl.
// The error bound of the factor route is 2 sum_j |psi_j| (|L|^T |phi|)_j,
// the first-order propagation of the per-column bound of psi into
// rho = sum_j psi_j^2.
// Launches of one mode write disjoint dot slots of the same partials array.
enum Var : int { V_RHO, V_GRAD, V_LAPPHI, V_LAPDER, V_HESS, V_F1, V_F4, V_FH, V_F5 };

// DT: the derivative type of a V_LAPDER launch, the mode of a V_F4 launch,
// the part of a V_FH launch.
template <int V, int DT>
struct Cols {
  static constexpr bool FAC = V >= V_F1;
  static constexpr int MODE = V == V_RHO || V == V_F1 ? MODE_RHO
                              : V == V_GRAD           ? MODE_GRAD
                              : V == V_HESS || V == V_FH ? MODE_HESS
                              : V == V_F4             ? DT
                                                      : MODE_LAP;
  static constexpr int NT = V == V_F5 ? 5 : V == V_FH ? (DT == 0 ? 4 : 2) : V == V_HESS || V == V_F4 ? 4 : 1;  // column types per tile
  static constexpr int TC = V == V_F5 ? 80 : V == V_FH ? 32 * NT : V == V_HESS || V == V_LAPDER || V == V_F4 ? 128 : 64;
  static constexpr int NCB = TC / 16;                     // 16-column MFMA blocks
  static constexpr int PTS = TC / NT;                     // points per workgroup
  static constexpr int NBC = V == V_LAPDER || V == V_FH ? 0
                             : V == V_F5                           ? 1
                             : V == V_HESS || V == V_F4            ? 2
                                                                   : NCB;  // bound blocks (the phi columns)
  // column blocks with FP64 running sums: none in the Hessian tile (short of
  // registers); in the factor route's lap and gradient/Hessian tiles only the
  // phi blocks (psi, which rho and its error bound come from), the
  // derivative blocks in FP32 pairs (no VALU work per k-step)
  static constexpr int NF = V == V_HESS ? 0 : V == V_F5 || V == V_F4 ? PTS / 16 : NCB;
  static constexpr bool TRI = V == V_RHO || V == V_LAPDER;  // quadratic forms only: triangle of P
  // AO array of column c: 0 = phi, 1..3 = d_x, d_y, d_z phi, then 4 = lap phi
  // (lap mode) or 4..9 = d_xx, d_xy, d_xz, d_yy, d_yz, d_zz phi (Hessian mode)
  __host__ __device__ static constexpr int type(int c) {
    return V == V_FH   ? 4 + 4 * DT + c / PTS
           : NT >= 4   ? c / PTS
           : V == V_LAPDER ? DT
                           : 0;
  }
  __host__ __device__ static constexpr int point(int c) { return NT > 1 ? c % PTS : c; }
};

// Dots produced from a GEMM column of type t: how many, and for the k-th one
// its slot in the mode's dot layout and the AO array it pairs the column with.
template <int V>
__host__ __device__ constexpr int ndots_of(int t) {
  if (V == V_RHO) return 1;
  if (V == V_GRAD) return 4;
  if (V == V_LAPPHI) return 5;
  if (V == V_LAPDER) return 1;
  return t == 0 ? 10 : t;  // HESS: t_xy.. lists grow with t
}
template <int V>
__host__ __device__ constexpr int dot_slot(int t, int k) {
  if (V == V_RHO || V == V_GRAD || V == V_LAPPHI || t == 0) return k;
  if (V == V_LAPDER) return 4 + t;
  return (t == 1 ? 10 : t == 2 ? 11 : 13) + k;  // t_xx | t_xy t_yy | t_xz t_yz t_zz
}
template <int V>
__host__ __device__ constexpr int dot_array(int t, int k) {
  if (V == V_RHO || V == V_GRAD || V == V_LAPPHI || t == 0) return k;
  if (V == V_LAPDER) return t;
  return 1 + k;
}

template <int N>
using IC = std::integral_constant<int, N>;

constexpr int kMaxDots = 11;  // 10 Hessian-mode dots of the phi column + the bound

// Dots of a factor-route tile per point (see factor_epilogue) and the slot
// of the k-th in the mode's partials layout (mode_ndot, the bound after the
// dots).
template <int V, int DT>
__host__ __device__ constexpr int factor_ndf() {
  using C = Cols<V, DT>;
  return V == V_F1 ? 2 : V == V_FH ? C::NT : V == V_F5 ? 9 : C::MODE == MODE_HESS ? 11 : 5;
}
template <int V, int DT>
__host__ __device__ constexpr int factor_slot(int k) {
  using C = Cols<V, DT>;
  return V == V_FH                        ? 4 + 4 * DT + k
         : V == V_F5 && k == 7            ? 4
         : k == factor_ndf<V, DT>() - 1   ? mode_ndot(C::MODE)
         : k < 4                          ? k
         : C::MODE == MODE_HESS           ? k + 6
                                          : k + 1;
}

// Epilogue of the factor route: per point, lane-local FP64 sums over the
// tile's rows (columns of L) of products of the point's psi columns, then over
// the four lane groups and the four waves in a fixed order; one partial per
// (row block, dot, point).  psi (Hessian mode): psi of the tile's rows and
// points at psi[row * ld + point], written by V_F4 and read by V_FH.
template <int V, int DT, class Tile>
__device__ __forceinline__ void factor_epilogue(const Tile& yt, uint16_t* lds, double* __restrict__ part,
                                                int64_t ld, int rt, int64_t pbase, int tid, double* __restrict__ psi) {
  using C = Cols<V, DT>;
  constexpr int ND = mode_ndot(C::MODE), ND1 = ND + 1;
  constexpr int NPB = C::PTS / 16;  // 16-point column blocks per column type
  constexpr bool LAP = C::MODE == MODE_LAP, HES = C::MODE == MODE_HESS;
  // dots per point: V_F1 [rho, bound]; V_FH [psi.d_kl psi x NT]; V_F4 [rho,
  // g_x, g_y, g_z, (Hessian: t_xx, t_xy, t_yy, t_xz, t_yz, t_zz,) bound];
  // V_F5 [rho, g_x, g_y, g_z, t_xx, t_yy, t_zz, psi.lap psi, bound]
  static_assert(V != V_F4 || !LAP, "the lap mode runs as V_F5");
  constexpr int NDF = factor_ndf<V, DT>();
  const int lane = tid & 63, w = tid >> 6, r = lane & 15, q = lane >> 4;
  double* red = reinterpret_cast<double*>(lds);  // [NW][PTS][NDF]
#pragma unroll
  for (int pb = 0; pb < NPB; ++pb) {
    double d[NDF];
#pragma unroll
    for (int k = 0; k < NDF; ++k) d[k] = 0.0;
    // lane row 4 q + j of the wave's 16-row block rb, point 16 pb + r
    auto psi_at = [&](int rb, int j) {
      return psi + (int64_t)(rt * TR + RW * w + 16 * rb + 4 * q + j) * ld + pbase + 16 * pb + r;
    };
    double ps[NRB][4];
    if constexpr (V == V_FH) {
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int j = 0; j < 4; ++j) ps[rb][j] = *psi_at(rb, j);
    }
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double y0 = V == V_FH ? ps[rb][j] : yt.y(rb, pb, j);
        if constexpr (V == V_FH) {
#pragma unroll
          for (int k = 0; k < C::NT; ++k) d[k] += y0 * yt.y(rb, k * NPB + pb, j);
        } else {
          if constexpr (HES) *psi_at(rb, j) = y0;
          d[0] += y0 * y0;
          d[NDF - 1] += 2.0 * fabs(y0) * (double)yt.bound[rb][pb][j];
          if constexpr (V == V_F4 || V == V_F5) {
            double yk[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              yk[k] = yt.y(rb, (k + 1) * NPB + pb, j);
              d[1 + k] += y0 * yk[k];
              if constexpr (LAP) d[4 + k] += yk[k] * yk[k];
            }
            if constexpr (V == V_F5) d[7] += y0 * yt.y(rb, 4 * NPB + pb, j);  // psi . lap psi
            if constexpr (HES) {  // t_xx, t_xy, t_yy, t_xz, t_yz, t_zz
              d[4] += yk[0] * yk[0];
              d[5] += yk[0] * yk[1];
              d[6] += yk[1] * yk[1];
              d[7] += yk[0] * yk[2];
              d[8] += yk[1] * yk[2];
              d[9] += yk[2] * yk[2];
            }
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      double x = d[k];
      x += __shfl_xor(x, 16);
      x += __shfl_xor(x, 32);
      d[k] = x;
    }
    if (q == 0) {
#pragma unroll
      for (int k = 0; k < NDF; ++k) red[((int64_t)w * C::PTS + 16 * pb + r) * NDF + k] = d[k];
    }
  }
  __syncthreads();
  if (tid < C::PTS) {
#pragma unroll
    for (int k = 0; k < NDF; ++k) {
      double sum = 0.0;
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) sum += red[((int64_t)ww * C::PTS + tid) * NDF + k];
      // dot slots of the mode (mode_ndot): lap [rho, g, L0, t_xx, t_yy, t_zz],
      // Hessian [rho, g, h_xx..h_zz, t_xx, t_xy, t_yy, t_xz, t_yz, t_zz]; the
      // bound after them
      part[((int64_t)rt * ND1 + factor_slot<V, DT>(k)) * ld + pbase + tid] = sum;
    }
  }
}

template <int V, int DT>
__global__ __launch_bounds__(256, 2) void contract_mixed_kernel(const u4* __restrict__ Pf, int rows, int mp,
                                                                const float* __restrict__ ao, int64_t ld,
                                                                double* __restrict__ part, int nrt,
                                                                int64_t nptile, const int32_t* __restrict__ klist,
                                                                int64_t kstride, const int32_t* __restrict__ kcount,
                                                                int kcount_all, double* __restrict__ psi) {
  using C = Cols<V, DT>;
  constexpr int TC = C::TC, NCB = C::NCB, NBC = C::NBC;
  constexpr int ND = mode_ndot(C::MODE);
  constexpr int ND1 = ND + 1;
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][NPL][TC][LROW];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar branches
  const int r = lane & 15, q = lane >> 4;
  // XCD-aware tile order: workgroup g runs on XCD g % 8 (round-robin
  // dispatch); the nrt row blocks of one point tile get consecutive slots on
  // the same XCD so its AO tile is fetched from HBM once and re-read from L2.
  // V_LAPDER: DT is the derivative type (1..3).
  const int g = blockIdx.x;
  constexpr int dtype = DT;
  const int xcd = g & 7, slot = g >> 3;
  const int rt = slot % nrt;
  const int64_t ptile = (int64_t)(slot / nrt) * 8 + xcd;
  if (ptile >= nptile) return;
  const int64_t pbase = ptile * C::PTS;
  const int i0 = rt * TR;
  const int64_t mq = mp / 4;  // row quads per AO array

  YTile<TC, NBC, C::NF> yt;
  // screened factor route (per-tile lists): the AOs are tile-packed
  // (ao_packed_index), this tile's k-step k at a fixed stride from k-step 0
  constexpr int NDER = mode_nder(C::MODE);
  const bool packed = C::FAC && kcount != nullptr;
  const int64_t tile_floats = (int64_t)(mp / KB) * NDER * 8 * C::PTS * 4;
  // this thread's B column (TC a multiple of 64), or any column's (bcol)
  auto bcol = [&](int c) {
    return packed ? ao + ptile * tile_floats + ((int64_t)C::type(c) * 8 * C::PTS + C::point(c)) * 4
                  : ao + ((int64_t)C::type(c) * mq * ld + pbase + C::point(c)) * 4;
  };
  const float* bsrc = TC % 64 == 0 ? bcol(tid & (TC - 1)) : nullptr;
  // factor route: the point tile's list of k-steps (screened: its own list
  // and length; else one shared list of all k-steps, kstride 0)
  const int32_t* kl = C::FAC ? klist + ptile * kstride : nullptr;
  const int kc = C::FAC ? (kcount ? kcount[ptile] : kcount_all) : 0;
  y_tile<TC, NBC, C::NF, C::TRI, C::FAC>(Pf, rows, mp, ld, i0, tid, bsrc, Bs, yt, C::TRI ? i0 / KB : 0, kl, kc,
                                          bcol, packed ? 4 * C::PTS : 0, packed ? (int64_t)NDER * 32 * C::PTS : 0);
  if constexpr (C::FAC) {
    factor_epilogue<V, DT>(yt, &Bs[0][0][0][0], part, ld, rt, pbase, tid, psi);
    return;
  } else {

  // ---------------------------------------------------------------- epilogue
  double* red = reinterpret_cast<double*>(&Bs[0][0][0][0]);  // [NW][TC][kMaxDots]
  const int64_t quad0 = (i0 + RW * w) / 4 + q;
  auto per_cb = [&](auto cbtag) {
    constexpr int CB = decltype(cbtag)::value;
    constexpr int tc = V == V_LAPDER ? 0 : C::type(16 * CB);  // compile-time column type (0 for V_LAPDER)
    constexpr int nd = ndots_of<V>(tc);
    constexpr bool bnd = CB < NBC;
    constexpr int t = V == V_LAPDER ? dtype : tc;
    const int64_t p = pbase + C::point(16 * CB) + r;
    double d[kMaxDots];
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k) d[k] = 0.0;
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
      f4 v[nd];
#pragma unroll
      for (int k = 0; k < nd; ++k)
        v[k] = *reinterpret_cast<const f4*>(ao + (((int64_t)dot_array<V>(t, k) * mq + quad0 + 4 * rb) * ld + p) * 4);
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
    const int tc = V == V_LAPDER ? 0 : C::type(c);
    const int t = V == V_LAPDER ? dtype : tc;
    const int nd = ndots_of<V>(tc);
    const bool has_bound = NBC > 0 && tc == 0;
    const int64_t p = pbase + C::point(c);
#pragma unroll
    for (int k = 0; k < kMaxDots; ++k) {
      const bool bound = (k == kMaxDots - 1) && has_bound;
      if (k < nd || bound) {
        double sum = 0.0;
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) sum += red[((int64_t)ww * TC + c) * kMaxDots + k];
        const int slot_k = bound ? ND : dot_slot<V>(t, k);
        part[((int64_t)rt * ND1 + slot_k) * ld + p] = sum;
      }
    }
  }
  }  // density route
}

// A: rows x mp (P: rows = mp); the workgroups cover rows / 128 row blocks
// times ld / PTS point tiles.
template <int V, int DT = 0>
void launch_var(const u4* A, int rows, int mp, const float* ao, int64_t ld, double* part, hipStream_t s,
                const int32_t* klist = nullptr, int64_t kstride = 0, const int32_t* kcount = nullptr,
                int kcount_all = 0, double* psi = nullptr) {
  using C = Cols<V, DT>;
  const int nrt = rows / TR;
  const int64_t nptile = ld / C::PTS;  // ld is a multiple of 256
  const dim3 block(256), grid((unsigned)(((nptile + 7) / 8) * 8 * nrt));
  hipLaunchKernelGGL((contract_mixed_kernel<V, DT>), grid, block, 0, s, A, rows, mp, ao, ld, part, nrt, nptile, klist,
                     kstride, kcount, kcount_all, psi);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace

void mixed_fragment_order(const uint16_t* planes, int rows, int cols, uint16_t* out) {
  // out[((pl * rows/16 + b16) * nk + kb) * 64 + lane][e] = plane pl of
  // A[16 b16 + lane % 16][32 kb + 8 (lane / 16) + e]
  const int nk = cols / KB, nb = rows / 16;
  const size_t plane = (size_t)rows * cols;
  size_t o = 0;
  for (int pl = 0; pl < NPL; ++pl)
    for (int b16 = 0; b16 < nb; ++b16)
      for (int kb = 0; kb < nk; ++kb)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e)
            out[o++] = planes[pl * plane + (size_t)(16 * b16 + (l & 15)) * cols + 32 * kb + 8 * (l >> 4) + e];
}

void launch_contract_mixed(int mode, const uint16_t* Pf, int nbasis_pad, const float* ao, int64_t ld, double* part,
                           hipStream_t s) {
  CT_REQUIRE(nbasis_pad % TR == 0, "padded basis size must be a multiple of 128");
  CT_REQUIRE(ld % 256 == 0, "point stride must be a multiple of 256");
  const u4* P = reinterpret_cast<const u4*>(Pf);
  const int mp = nbasis_pad;
  switch (mode) {
    case MODE_RHO:
      launch_var<V_RHO>(P, mp, mp, ao, ld, part, s);
      break;
    case MODE_GRAD:
      launch_var<V_GRAD>(P, mp, mp, ao, ld, part, s);
      break;
    case MODE_LAP:
      launch_var<V_LAPPHI>(P, mp, mp, ao, ld, part, s);
      launch_var<V_LAPDER, 1>(P, mp, mp, ao, ld, part, s);
      launch_var<V_LAPDER, 2>(P, mp, mp, ao, ld, part, s);
      launch_var<V_LAPDER, 3>(P, mp, mp, ao, ld, part, s);
      break;
    default:
      launch_var<V_HESS>(P, mp, mp, ao, ld, part, s);
      break;
  }
}

int factor_tile_points(int mode) {
  switch (mode) {
    case MODE_RHO: return Cols<V_F1, 0>::PTS;
    case MODE_GRAD: return Cols<V_F4, MODE_GRAD>::PTS;
    case MODE_LAP:
      return Cols<V_F5, 0>::PTS;
    default:
      static_assert(Cols<V_F4, MODE_HESS>::PTS == Cols<V_FH, 0>::PTS && Cols<V_FH, 0>::PTS == Cols<V_FH, 1>::PTS,
                    "one tile size per mode");
      return Cols<V_F4, MODE_HESS>::PTS;
  }
}

void launch_contract_mixed_factor(int mode, const uint16_t* Lf, int rows_pad, int nbasis_pad, const float* ao,
                                  int64_t ld, double* part, const int32_t* klist, const int32_t* kcount,
                                  hipStream_t s, double* psi) {
  CT_REQUIRE(klist != nullptr, "the factor route needs a k-step list");
  CT_REQUIRE(mode != MODE_HESS || psi != nullptr, "the factor route's Hessian mode needs the psi buffer");
  // screened: per-tile lists of stride nk and their lengths; else one list
  // of all nk k-steps shared by every tile
  const int nk = nbasis_pad / KB;
  const int64_t kstride = kcount ? nk : 0;
  CT_REQUIRE(nbasis_pad % TR == 0 && rows_pad % TR == 0 && rows_pad > 0, "factor / basis padding must be multiples of 128");
  CT_REQUIRE(ld % 256 == 0, "point stride must be a multiple of 256");
  const u4* L = reinterpret_cast<const u4*>(Lf);
  switch (mode) {
    case MODE_RHO:
      launch_var<V_F1>(L, rows_pad, nbasis_pad, ao, ld, part, s, klist, kstride, kcount, nk);
      break;
    case MODE_GRAD:
      launch_var<V_F4, MODE_GRAD>(L, rows_pad, nbasis_pad, ao, ld, part, s, klist, kstride, kcount, nk);
      break;
    case MODE_LAP:
      launch_var<V_F5>(L, rows_pad, nbasis_pad, ao, ld, part, s, klist, kstride, kcount, nk);
      break;
    default:
      launch_var<V_F4, MODE_HESS>(L, rows_pad, nbasis_pad, ao, ld, part, s, klist, kstride, kcount, nk, psi);
      launch_var<V_FH, 0>(L, rows_pad, nbasis_pad, ao, ld, part, s, klist, kstride, kcount, nk, psi);
      launch_var<V_FH, 1>(L, rows_pad, nbasis_pad, ao, ld, part, s, klist, kstride, kcount, nk, psi);
      break;
  }
}

}  // namespace ct
