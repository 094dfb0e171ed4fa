// Fused AO evaluation: values and first/second derivatives of every basis
// function on a chunk of points, one lane per point.
//
// Replaces the reference AO device functions (src/evaluate_density.cu:25-411,
// src/evaluate_gradient.cu:18-1126, src/evaluate_laplacian.cu:15-421,
// src/evaluate_hessian.cu:17-1836).  Differences that matter on CDNA4:
//   * the shell table is a device array read at a wave-uniform index, so the
//     compiler issues scalar (SMEM) loads into SGPRs -- no constant-bank size
//     limit and no per-chunk re-upload;
//   * each primitive's exponential is evaluated once per (point, shell) and
//     reused by every component and derivative through the three radial sums
//     S0, S1, S2 (the reference re-evaluates normalisation sqrt()s and does a
//     global read-modify-write per primitive and component);
//   * normalisation is folded into the coefficients on the host;
//   * every output element is written exactly once (no += on uninitialised
//     memory), rows beyond M are written as zeros for the padded contraction;
//   * a shell, and within a shell a primitive, is skipped for the whole wave
//     when every lane has a r^2 > screen (exp underflow region in parity
//     mode; below FP32 resolution of any density above the re-evaluation
//     floor in the mixed path);
//   * in the screened layout of the factor route (points in Morton
//     order, per-tile activity masks) a wave skips, stores included, every
//     shell that lies only in k-steps the contraction skips for its points;
//   * values are computed in FP64 for both paths; the mixed path rounds them
//     once to FP32 on the store (0.5 ulp per value: the FP32 AO error that
//     enters the mixed path's error estimate is only this rounding).
//
// Output layouts (d = derivative slot, i = basis function, p = point in the
// chunk): FP64 ao[(d * Mp + i) * ld + p], points contiguous (coalesced stores,
// the B operand layout of the FP64 contraction); FP32 in row quads
// (ao_quad_index): the four values of a quad are assembled in registers and
// stored as 16 contiguous bytes per point, which the mixed contraction reads
// with one 16-byte load.

#include <type_traits>

#include <algorithm>

#include "ct_internal.h"
#include "generated/ao_poly.h"

namespace ct {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// TS: storage type (double: FP64 path, row-major; float: mixed path, row quads).
// LMAX: largest |l| of the basis (the code of larger shells is compiled out).
template <typename TS, int MODE, int LMAX>
__global__ __launch_bounds__(256) void ao_eval_kernel(const ShellMeta* __restrict__ meta,
                                                      const ShellGeom* __restrict__ geom,
                                                      const PrimData* __restrict__ prim, int nshell, int nbasis,
                                                      int nbasis_pad, const double* __restrict__ pts, int64_t npts,
                                                      int64_t ld, TS* __restrict__ ao, double screen,
                                                      const uint32_t* __restrict__ kmask, int nwords,
                                                      int tile_points, int rows_out) {
  using T = double;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // p < ld by construction
  const int64_t pc = p < npts ? p : (npts - 1);
  const double px = pts[3 * pc + 0], py = pts[3 * pc + 1], pz = pts[3 * pc + 2];
  constexpr int nder = mode_nder(MODE);
  constexpr bool quads = sizeof(TS) == 4;
  const int64_t mq = nbasis_pad >> 2;
  // FP32 (row quads): the four rows of a quad are assembled in registers and
  // stored with one 16-byte store per lane once the quad is complete (shell
  // boundaries need not align with quads: the partial quad carries over).
  TS qb[quads ? nder : 1][4];
#pragma unroll
  for (int d = 0; d < (quads ? nder : 1); ++d)
#pragma unroll
    for (int e = 0; e < 4; ++e) qb[d][e] = TS(0);
  // screened layout, up to 64 k-steps: this lane's own tile mask; a quad is
  // stored only where the lane's tile reads its k-step (the wave evaluates
  // the union of its tiles' k-steps, ~10 % more blocks than any one tile
  // reads with 16-point tiles).  Tiles past the chunk's last point hold no
  // mask, and the padding lanes in them store everything.
  uint32_t lm0 = 0xFFFFFFFFu, lm1 = 0xFFFFFFFFu;
  if (kmask && nwords <= 2 && p < npts) {
    const uint32_t* lmask = kmask + (p / tile_points) * nwords;
    lm0 = lmask[0];
    if (nwords == 2) lm1 = lmask[1];
  }
  // Screened layout: tile-packed blocks (ao_packed_index), so that the quads
  // one wave stores for a (tile, k-step) land in one 10 KiB region (lap mode,
  // 16-point tiles) instead of 256-byte pieces a whole array apart.
  const int64_t tile = kmask ? p / tile_points : 0;
  const int pp = kmask ? (int)(p - tile * tile_points) : 0;
  const int nk = nbasis_pad >> 5;
  auto store_quad = [&](int d, int quad) {
    if constexpr (quads) {
      const int k = quad >> 3;  // 32-function k-step of the quad
      if ((((k >> 5) ? lm1 : lm0) >> (k & 31)) & 1u) {
        const int64_t off = kmask ? ao_packed_index(tile, k, d, quad & 7, pp, nk, nder, tile_points)
                                  : (((int64_t)d * mq + quad) * ld + p) * 4;
        *reinterpret_cast<f4*>(ao + off) = f4{qb[d][0], qb[d][1], qb[d][2], qb[d][3]};
      }
    }
  };
  auto dispatch = [&](int code, T x, T y, T z, T s0, T s1, T s2, auto&& st) {
    if constexpr (MODE == MODE_RHO) ao_shell_dispatch_val<T, LMAX>(code, x, y, z, s0, s1, s2, st);
    if constexpr (MODE == MODE_GRAD) ao_shell_dispatch_grad<T, LMAX>(code, x, y, z, s0, s1, s2, st);
    if constexpr (MODE == MODE_LAP) ao_shell_dispatch_lap<T, LMAX>(code, x, y, z, s0, s1, s2, st);
    if constexpr (MODE == MODE_HESS) ao_shell_dispatch_hess<T, LMAX>(code, x, y, z, s0, s1, s2, st);
  };

  // screened layout: the activity masks of the 64 / tile_points point tiles
  // of this wave's 64 points
  const uint32_t* mask0 = nullptr;
  int ntw = 0;
  if (kmask) {
    const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + 64 * __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    mask0 = kmask + (p0 / tile_points) * nwords;
    ntw = 64 / tile_points;
  }
  // whether k-step k is active for any tile of the wave; up to 64 k-steps
  // (2048 functions) the wave's union mask is formed once, in SGPRs
  uint32_t wm0 = 0, wm1 = 0;
  if (kmask && nwords <= 2)
    for (int t = 0; t < ntw; ++t) {
      wm0 |= mask0[t * nwords];
      if (nwords == 2) wm1 |= mask0[t * nwords + 1];
    }
  auto active = [&](int k) {
    uint32_t m = 0;
    if (nwords <= 2) {
      m = (k >> 5) ? wm1 : wm0;
    } else {
      for (int t = 0; t < ntw; ++t) m |= mask0[t * nwords + (k >> 5)];
    }
    return (m >> (k & 31)) & 1u;
  };
  constexpr int kNfunc[8] = {1, 3, 6, 10, 15, 5, 7, 9};  // per shell code

  // FP64 rows (no quads shared between shells): gridDim.y > 1 splits the
  // shells into ranges, one per workgroup row, so that few points still fill
  // the GPU (the re-evaluation of a mixed call's flagged points)
  const int s_begin = (int)((int64_t)blockIdx.y * nshell / gridDim.y);
  const int s_end = (int)((int64_t)(blockIdx.y + 1) * nshell / gridDim.y);
  for (int s = s_begin; s < s_end; ++s) {
    const ShellMeta m = meta[s];
    if (kmask) {
      // skipped: no k-step that the shell touches is read by the contraction
      // for these points (a quad it shares with a neighbouring shell lies in
      // one of those k-steps, so storing it or not does not matter either)
      const int k0 = m.first >> 5, k1 = (m.first + kNfunc[m.code] - 1) >> 5;
      if (!active(k0) && !active(k1)) continue;
    }
    const ShellGeom g = geom[s];
    const double dx = px - g.cx, dy = py - g.cy, dz = pz - g.cz;
    const double r2d = dx * dx + dy * dy + dz * dz;
    const T x = dx, y = dy, z = dz;
    T s0 = 0, s1 = 0, s2 = 0;
    const bool far = g.amin * r2d > screen;
    if (!__all(far)) {
      for (int k = 0; k < m.nprim; ++k) {
        const PrimData pd = prim[m.prim_off + k];
        const T a = pd.a;
        const T ar2 = a * r2d;
        if (__all(ar2 > screen)) continue;
        const T e = pd.c * exp(-ar2);
        s0 += e;
        if constexpr (MODE != MODE_RHO) {
          const T m2a = T(-2) * a;
          s1 += m2a * e;
          if constexpr (MODE == MODE_LAP || MODE == MODE_HESS) s2 += m2a * m2a * e;
        }
      }
    }
    const int first = m.first;
    if constexpr (!quads) {
      auto st = [&](int d, int ic, T v) { ao[((int64_t)d * nbasis_pad + first + ic) * ld + p] = (TS)v; };
      dispatch(m.code, x, y, z, s0, s1, s2, st);
    } else {
      // the slot of component ic in its quad is (first + ic) & 3: compile-time
      // per alignment, so qb stays in registers
      auto run = [&](auto atag) {
        constexpr int A = decltype(atag)::value;
        auto st = [&](int d, int ic, T v) {
          const int sl = (A + ic) & 3;
          qb[d][sl] = (TS)v;
          if (sl == 3) store_quad(d, (first + ic) >> 2);
        };
        dispatch(m.code, x, y, z, s0, s1, s2, st);
      };
      switch (first & 3) {
        case 0: run(std::integral_constant<int, 0>{}); break;
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        default: run(std::integral_constant<int, 3>{}); break;
      }
    }
  }
  // zero rows of the padded basis so the contraction never reads garbage
  // (up to rows_out: a caller whose kernels stop short of nbasis_pad leaves
  // the rest unwritten)
  if constexpr (!quads) {
    if (blockIdx.y + 1 == gridDim.y)
      for (int d = 0; d < nder; ++d)
        for (int i = nbasis; i < rows_out; ++i) ao[((int64_t)d * nbasis_pad + i) * ld + p] = TS(0);
  } else {
    const int a = nbasis & 3;
    if (a != 0) {
#pragma unroll
      for (int d = 0; d < nder; ++d) {
        qb[d][3] = TS(0);
        if (a < 3) qb[d][2] = TS(0);
        if (a < 2) qb[d][1] = TS(0);
        store_quad(d, nbasis >> 2);
      }
    }
#pragma unroll
    for (int d = 0; d < nder; ++d) {
#pragma unroll
      for (int e = 0; e < 4; ++e) qb[d][e] = TS(0);
      for (int quad = (nbasis + 3) >> 2; quad < (rows_out >> 2); ++quad) store_quad(d, quad);
    }
  }
}

template <typename TS, int MODE>
void launch_lmax(int lmax, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell, int nbasis,
                 int nbasis_pad, const double* pts, int64_t npts, int64_t ld, TS* ao, double screen,
                 const uint32_t* kmask, int nwords, int tile_points, int rows_out, hipStream_t s) {
  // FP64 rows: shell ranges over grid.y until ~2048 waves (a few per SIMD)
  // are in flight; the FP32 row quads carry partial quads from shell to shell
  const int64_t waves = ld / 64;
  const int nsplit = sizeof(TS) == 8 ? (int)std::max<int64_t>(1, std::min<int64_t>({(2048 + waves - 1) / waves, 8,
                                                                                      (int64_t)nshell}))
                                     : 1;
  const dim3 block(256), grid((unsigned)(ld / 256), (unsigned)nsplit);
  if (lmax <= 2)
    hipLaunchKernelGGL((ao_eval_kernel<TS, MODE, 2>), grid, block, 0, s, meta, geom, prim, nshell, nbasis, nbasis_pad,
                       pts, npts, ld, ao, screen, kmask, nwords, tile_points, rows_out);
  else if (lmax == 3)
    hipLaunchKernelGGL((ao_eval_kernel<TS, MODE, 3>), grid, block, 0, s, meta, geom, prim, nshell, nbasis, nbasis_pad,
                       pts, npts, ld, ao, screen, kmask, nwords, tile_points, rows_out);
  else
    hipLaunchKernelGGL((ao_eval_kernel<TS, MODE, 4>), grid, block, 0, s, meta, geom, prim, nshell, nbasis, nbasis_pad,
                       pts, npts, ld, ao, screen, kmask, nwords, tile_points, rows_out);
  CT_HIP_CHECK(hipGetLastError());
}

template <typename TS>
void launch_mode(int mode, int lmax, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                 int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, TS* ao, double screen,
                 const uint32_t* kmask, int nwords, int tile_points, int rows_out, hipStream_t s) {
  switch (mode) {
    case MODE_RHO:
      launch_lmax<TS, MODE_RHO>(lmax, meta, geom, prim, nshell, nbasis, nbasis_pad, pts, npts, ld, ao, screen, kmask,
                          nwords, tile_points, rows_out, s);
      break;
    case MODE_GRAD:
      launch_lmax<TS, MODE_GRAD>(lmax, meta, geom, prim, nshell, nbasis, nbasis_pad, pts, npts, ld, ao, screen, kmask,
                          nwords, tile_points, rows_out, s);
      break;
    case MODE_LAP:
      launch_lmax<TS, MODE_LAP>(lmax, meta, geom, prim, nshell, nbasis, nbasis_pad, pts, npts, ld, ao, screen, kmask,
                          nwords, tile_points, rows_out, s);
      break;
    default:
      launch_lmax<TS, MODE_HESS>(lmax, meta, geom, prim, nshell, nbasis, nbasis_pad, pts, npts, ld, ao, screen, kmask,
                          nwords, tile_points, rows_out, s);
      break;
  }
}

}  // namespace

void launch_ao_f64(int mode, int lmax, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                   int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, double* ao,
                   double screen, hipStream_t s, const uint32_t* kmask, int nwords, int tile_points, int rows_out) {
  CT_REQUIRE(!kmask || (tile_points > 0 && 64 % tile_points == 0 && nwords == (nbasis_pad / 32 + 31) / 32),
             "activity masks do not match the AO kernel's waves");
  if (rows_out <= 0) rows_out = nbasis_pad;
  CT_REQUIRE(rows_out >= nbasis && rows_out <= nbasis_pad && rows_out % 4 == 0, "rows written by the AO kernel");
  launch_mode<double>(mode, lmax, meta, geom, prim, nshell, nbasis, nbasis_pad, pts, npts, ld, ao, screen, kmask,
                      nwords, tile_points, rows_out, s);
}

void launch_ao_f32(int mode, int lmax, const ShellMeta* meta, const ShellGeom* geom, const PrimData* prim, int nshell,
                   int nbasis, int nbasis_pad, const double* pts, int64_t npts, int64_t ld, float* ao, double screen,
                   const uint32_t* kmask, int nwords, int tile_points, hipStream_t s, int rows_out) {
  CT_REQUIRE(!kmask || (tile_points > 0 && 64 % tile_points == 0 && nwords == (nbasis_pad / 32 + 31) / 32),
             "activity masks do not match the AO kernel's waves");
  if (rows_out <= 0) rows_out = nbasis_pad;
  CT_REQUIRE(rows_out >= nbasis && rows_out <= nbasis_pad && rows_out % 4 == 0, "rows written by the AO kernel");
  launch_mode<float>(mode, lmax, meta, geom, prim, nshell, nbasis, nbasis_pad, pts, npts, ld, ao, screen, kmask,
                     nwords, tile_points, rows_out, s);
}

}  // namespace ct
