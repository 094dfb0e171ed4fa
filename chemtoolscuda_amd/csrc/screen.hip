// Spatial screening of the factor route, device side.
//
//   1. points are put in Morton order (24-bit keys on a lattice around the
//      molecule; stable radix sort of (key, index) pairs), so that a point
//      tile of the contraction is a compact box in space;
//   2. per point tile (16, 32 or 64 points by mode; a wave takes 64 points)
//      the tile's bounding box marks the
//      32-function k-steps holding a shell within reach of the box (distance
//      below the shell's cutoff radius); the others are skipped by the AO
//      kernel (no evaluation, no stores) and by the contraction (no loads, no
//      MFMAs).  Per tile: the list of active k-steps in increasing order, its
//      length and a bit mask.
//
// The reference evaluates every basis function at every point
// (src/evaluate_density.cu:25-411); screening is this library's own.

#include <rocprim/device/device_radix_sort.hpp>

#include "ct_internal.h"

namespace ct {

namespace {

// bits of an 8-bit cell index spread to every third bit
__device__ __forceinline__ uint32_t spread8(uint32_t v) {
  v &= 0xFFu;
  v = (v | (v << 8)) & 0x0000F00Fu;
  v = (v | (v << 4)) & 0x000C30C3u;
  v = (v | (v << 2)) & 0x00249249u;
  return v;
}

__global__ __launch_bounds__(256) void morton_keys_kernel(const double* __restrict__ pts, int64_t n, MortonFrame f,
                                                          uint32_t* __restrict__ keys, int32_t* __restrict__ idx) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double t = (pts[3 * p + k] - f.origin[k]) * f.inv_cell;
    q[k] = t > 0.0 ? (t < 255.0 ? (uint32_t)t : 255u) : 0u;  // NaN -> 0
  }
  keys[p] = (spread8(q[0]) << 2) | (spread8(q[1]) << 1) | spread8(q[2]);
  idx[p] = (int32_t)p;
}

// One wave per 64 points = 64 / tp tiles of tp points (tp in 16, 32, 64):
// the tiles' bounding boxes by reductions within groups of tp lanes, then
// each group the active k-steps of its tile, one lane per k-step, tp k-steps
// per round (all lanes busy: one tile at a time left three quarters of them
// idle with 16-point tiles and 20 k-steps).
__global__ __launch_bounds__(256) void tile_activity_kernel(const double* __restrict__ pts, int64_t n, int tp,
                                                            int64_t ntile, const int32_t* __restrict__ span_off,
                                                            const ScreenSpan* __restrict__ spans, int nk,
                                                            int32_t* __restrict__ klist, int32_t* __restrict__ kcount,
                                                            uint32_t* __restrict__ kmask, int nwords,
                                                            unsigned long long* __restrict__ active_count) {
  const int lane = threadIdx.x & 63;
  const int tpw = 64 / tp;
  const int64_t t0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * tpw;  // the wave's first tile
  if (t0 >= ntile) return;
  const int64_t p = t0 * tp + lane;
  const bool valid = t0 + lane / tp < ntile && p < n;
  double lo[3], hi[3];
  bool finite = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double v = valid ? pts[3 * p + k] : 0.0;
    finite = finite && isfinite(v);
    lo[k] = valid ? v : 1e300;
    hi[k] = valid ? v : -1e300;
  }
  // a point with a non-finite coordinate has no place in space: its tile
  // keeps every k-step (and its NaN results)
  const uint64_t nonfinite = __ballot(!finite);
  for (int off = tp >> 1; off >= 1; off >>= 1)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], off));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], off));
    }
  const uint64_t gmask = tp == 64 ? ~0ull : ((1ull << tp) - 1ull);
  // the wave's tiles side by side: lane group g = lane / tp holds tile t0 + g
  // (its box is already in every lane of the group) and tests tp k-steps per
  // round, one per lane
  const int g = lane / tp, j = lane % tp;
  const int64_t tile = t0 + g;
  const bool tvalid = tile < ntile;
  const bool keep_all = ((nonfinite >> (g * tp)) & gmask) != 0;
  const bool any = lo[0] <= hi[0];  // the tile holds at least one point
  int total = 0;                    // the tile's active k-steps so far (same in its group)
  uint32_t word = 0;                // mask word in assembly (group leader)
  for (int kb0 = 0; kb0 < nk; kb0 += tp) {
    const int kb = kb0 + j;
    bool act = tvalid && keep_all && kb < nk && span_off[kb + 1] > span_off[kb];
    if (tvalid && any && !keep_all && kb < nk) {
      for (int e = span_off[kb]; e < span_off[kb + 1] && !act; ++e) {
        const ScreenSpan sp = spans[e];
        const double dx = fmax(0.0, fmax(lo[0] - sp.cx, sp.cx - hi[0]));
        const double dy = fmax(0.0, fmax(lo[1] - sp.cy, sp.cy - hi[1]));
        const double dz = fmax(0.0, fmax(lo[2] - sp.cz, sp.cz - hi[2]));
        act = dx * dx + dy * dy + dz * dz < sp.r2;
      }
    }
    const uint64_t gbits = (__ballot(act) >> (g * tp)) & gmask;
    if (act) klist[tile * nk + total + __popcll(gbits & ((1ull << j) - 1ull))] = kb;
    if (j == 0 && tvalid) {
      // this round's tp mask bits start at bit kb0 of the tile's mask
      if (tp == 64) {
        kmask[tile * nwords + (kb0 >> 5)] = (uint32_t)gbits;
        if ((kb0 >> 5) + 1 < nwords) kmask[tile * nwords + (kb0 >> 5) + 1] = (uint32_t)(gbits >> 32);
      } else {
        word |= (uint32_t)gbits << (kb0 & 31);
        if (((kb0 + tp) & 31) == 0 || kb0 + tp >= nk) {
          kmask[tile * nwords + (kb0 >> 5)] = word;
          word = 0;
        }
      }
    }
    total += __popcll(gbits);
  }
  if (j == 0 && tvalid) kcount[tile] = total;
  // the wave's active blocks, one atomic per wave over kScreenCounters
  // counters on separate 128-byte lines (one address for every wave would
  // serialise the atomics)
  int wsum = 0;
  for (int t = 0; t < 64 / tp; ++t) wsum += __shfl(total, t * tp);  // 0 for tiles past the last
  if (lane == 0)
    atomicAdd(active_count + ((t0 / (64 / tp)) % kScreenCounters) * 16, (unsigned long long)wsum);
}

}  // namespace

// rocPRIM sorts up to 2^20 items by a merge sort (a block sort and two
// launches per merge pass: 17 launches for 663k pairs); this configuration
// takes the radix passes for every size beyond one block instead: a
// histogram, its scan and one pass per 8 key bits.
using RadixOnly = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                             rocprim::default_config, 0>;

size_t screen_sort_workspace(int64_t n) {
  size_t bytes = 0;
  CT_HIP_CHECK(rocprim::radix_sort_pairs<RadixOnly>(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                    (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n, 0, 24));
  return bytes;
}

void launch_screen_sort(const double* pts, int64_t n, const MortonFrame& f, uint32_t* keys_in, uint32_t* keys_out,
                        int32_t* idx_in, int32_t* perm, void* temp, size_t temp_bytes, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(morton_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pts, n, f, keys_in,
                     idx_in);
  CT_HIP_CHECK(hipGetLastError());
  size_t bytes = temp_bytes;
  CT_HIP_CHECK(rocprim::radix_sort_pairs<RadixOnly>(temp, bytes, (const uint32_t*)keys_in, keys_out,
                                                    (const int32_t*)idx_in, perm, (size_t)n, 0, 24, s));
}

size_t sort_indices_workspace(int64_t n, int bits) {
  size_t bytes = 0;
  CT_HIP_CHECK(rocprim::radix_sort_keys<RadixOnly>(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                   (size_t)n, 0, (unsigned)bits));
  return bytes;
}

void launch_sort_indices(const int32_t* in, int32_t* out, int64_t n, int bits, void* temp, size_t temp_bytes,
                         hipStream_t s) {
  if (n <= 0) return;
  size_t bytes = temp_bytes;
  CT_HIP_CHECK(rocprim::radix_sort_keys<RadixOnly>(temp, bytes, reinterpret_cast<const uint32_t*>(in),
                                                   reinterpret_cast<uint32_t*>(out), (size_t)n, 0, (unsigned)bits,
                                                   s));
}

void launch_tile_activity(const double* sorted_pts, int64_t n, int tile_points, int64_t ntile, const int32_t* span_off,
                          const ScreenSpan* spans, int nk, int32_t* klist, int32_t* kcount, uint32_t* kmask,
                          int nwords, unsigned long long* active_count, hipStream_t s) {
  CT_REQUIRE(tile_points == 16 || tile_points == 32 || tile_points == 64, "tiles of 16, 32 or 64 points");
  CT_REQUIRE(nwords == (nk + 31) / 32, "mask width does not match the number of k-steps");
  if (ntile <= 0) return;
  const int64_t waves = (ntile + 64 / tile_points - 1) / (64 / tile_points);
  hipLaunchKernelGGL(tile_activity_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, sorted_pts, n,
                     tile_points, ntile, span_off, spans, nk, klist, kcount, kmask, nwords, active_count);
  CT_HIP_CHECK(hipGetLastError());
}

}  // namespace ct
