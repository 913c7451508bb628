// pagerank.hip -- pull-direction PageRank sweep for gfx950.
//
// Replaces the sweep of nw::graph::page_rank (include/nwgraph/algorithms/page_rank.hpp:121-147)
// with the Jacobi schedule of its twin page_rank_v10
// (include/nwgraph/experimental/algorithms/page_rank.hpp:397-430):
//   z_i  = sum_{u in in(i)} contrib[u]                 (gather over CSC row i)
//   x_i' = fma(d, z_i, base)                           (page_rank.hpp:133)
//   err += |x_i' - x_i|  (float difference, double sum) (page_rank.hpp:134)
//   contrib'_i = x_i' / (float)outdeg_i                 (page_rank.hpp:116,135)
// contrib is double-buffered (read contrib[k&1], write contrib[(k+1)&1]); x is
// updated in place (each row owns its entry).  One sweep is a short sequence of
// launches, one per row class (degree-binned schedule):
//   pr_tile_kernel    256 consecutive rows, one lane per row with in-degree <= short_max
//   pr_group_kernel   rows listed per class, 4/8/16/32/64 lanes per row, butterfly reduction
//   pr_sgroup_kernel  source-sliced rows (slice_min <= d < hub_min): workgroup b gathers
//                     only sources u with (u / 32) % 8 == b % 8 -> one partial per (row, slice)
//   pr_hubseg_kernel  hub_segment nonzeros of one hub row per workgroup -> partial sum
//   pr_finish_kernel  adds each hub row's segment partials and each sliced row's slice
//                     partials in order, updates those rows and folds the per-workgroup
//                     L1 partials
//
// Convergence stays on the device: the L1 change of sweep k is accumulated in
// fixed point in acc[k] (integer atomics: independent of block order); every
// workgroup of sweep k+1 first reads acc[k] and exits if it is below threshold
// (page_rank.hpp:144-146), so a captured loop of max_iters sweeps stops doing work
// after convergence without a host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kThreads = 256;
// rounding slack allowed above 1 in the sum of a teleport vector
constexpr double kTeleportSlack = 1e-4;

struct SweepParams {
  const uint32_t* rp;
  const uint32_t* col;
  const uint32_t* deg;
  const float* cin;
  float* cout;
  float* x;
  const float* basev;  // personalised base per row, or null
  float base;
  float damp;
  uint32_t n;
  uint32_t short_max;
  uint32_t n_tiles;
  uint32_t cls_blk[kGroupClasses + 1];  // first workgroup of each group class (relative to n_tiles)
  uint32_t cls_off[kGroupClasses + 1];
  const uint32_t* group_rows;
  uint32_t hub_blk;  // workgroups before the hub segments = entries of blk_err
  uint32_t n_hub_segs;
  uint32_t n_hub_rows;
  uint32_t hub_segment;
  const uint32_t* hub_rows;
  const uint32_t* seg_ptr;
  const uint32_t* seg_row;
  float* hub_part;
  double* blk_err;  // L1 partial of every tile / group workgroup of the sweep
  // source-sliced rows
  const uint32_t* srows;
  const uint32_t* sseg;
  const uint32_t* slist;
  uint32_t scls_off[kGroupClasses + 1];
  uint32_t n_srows;
  float* spart;  // kSlices x n_srows
  unsigned long long* acc;
  uint32_t iter;
  double threshold;
  // delta schedule (page_rank_v13): cin holds the deltas, cout the residuals, x the scores;
  // acc[k] is nonzero when some vertex pushed in iteration k
  uint32_t delta;
};

// entries of blk_err reduced by one workgroup of the finishing kernel
constexpr uint32_t kFinishChunk = 2048;
// gathers in flight per lane: group, hub and sliced rows / one-lane tile rows
constexpr uint32_t kUnroll = 4;
constexpr uint32_t kTileUnroll = 4;

// Streamed arrays (ids, row pointers, degrees, scores) may be read and written with
// non-temporal hints so that they do not displace the gathered contributions from
// L2 / MALL; the gathers themselves always use ordinary cached loads.
// NT is the hint level: 0 none, 1 the column-id stream only, 2 every streamed array
// (loads and stores).
template <int NT, class T>
__device__ __forceinline__ T ld(const T* p) {
  if constexpr (NT >= 2) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int NT, class T>
__device__ __forceinline__ T ld_col(const T* p) {
  if constexpr (NT >= 1) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int NT, class T>
__device__ __forceinline__ void st(T* p, T v) {
  if constexpr (NT >= 2) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// z + sum of cin[col[q]] over q = p, p + S, p + 2S, ... < p1, added in that order (one
// sequential sum), with kU gathers issued before they are added: a lane keeps kU
// misses in flight instead of one.
template <int NT, uint32_t S, uint32_t kU>
__device__ __forceinline__ float gather_sum(const SweepParams& P, uint32_t p0, uint32_t p1, float z) {
  for (uint64_t p = p0; p < p1; p += S * kU) {  // 64-bit: no wrap near 2^32 nonzeros
    float v[kU];
#pragma unroll
    for (uint32_t r = 0; r < kU; ++r) {
      const uint64_t q = p + r * S;
      v[r] = q < p1 ? P.cin[ld_col<NT>(P.col + q)] : 0.f;
    }
#pragma unroll
    for (uint32_t r = 0; r < kU; ++r)
      if (p + r * S < p1) z += v[r];
  }
  return z;
}

__device__ __forceinline__ bool converged_before(const SweepParams& P) {
  if (P.iter == 0) return false;
  // delta schedule: the loop ends after an iteration in which nothing was pushed
  // (experimental/algorithms/page_rank.hpp:616-620)
  if (P.delta) return P.acc[P.iter - 1] == 0ull;
  return fixed_to_err(P.acc[P.iter - 1]) < P.threshold;
}

template <int NT>
__device__ __forceinline__ double update_row(const SweepParams& P, uint32_t i, float z) {
  if (P.delta) {
    // delta schedule, pull phase (experimental/algorithms/page_rank.hpp:597-612): the row's
    // residual becomes the sum of its in-neighbours' deltas (all >= 0), or 0
    st<NT>(P.cout + i, z > 0.f ? z : 0.f);
    return 0.0;
  }
  const float b = P.basev ? ld<NT>(P.basev + i) : P.base;
  const float xo = ld<NT>(P.x + i);
  const float xn = __builtin_fmaf(P.damp, z, b);
  st<NT>(P.x + i, xn);
  st<NT>(P.cout + i, xn / (float)ld<NT>(P.deg + i));
  return (double)fabsf(xn - xo);
}

// Tiles: 256 consecutive rows, one lane per row with in-degree <= short_max; the lane
// adds its row's contributions in row order.
template <int NT>
__global__ __launch_bounds__(kThreads) void pr_tile_kernel(SweepParams P) {
  if (converged_before(P)) return;
  __shared__ double red_d[kThreads / 64];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  double e = 0.0;
  if (i < P.n) {
    const uint32_t p0 = ld<NT>(P.rp + i), p1 = ld<NT>(P.rp + i + 1);
    if (p1 - p0 <= P.short_max) {
      e = update_row<NT>(P, i, gather_sum<NT, 1, kTileUnroll>(P, p0, p1, 0.f));
    }
  }
  e = block_sum<kThreads>(e, red_d);
  if (threadIdx.x == 0) P.blk_err[blockIdx.x] = e;
}

// Group classes: G lanes per listed row stride over the row, butterfly reduction.
template <int G, int NT>
__global__ __launch_bounds__(kThreads) void pr_group_kernel(SweepParams P, uint32_t cls) {
  if (converged_before(P)) return;
  __shared__ double red_d[kThreads / 64];
  constexpr uint32_t kRowsPerBlock = kThreads / G;
  const uint32_t slot = blockIdx.x * kRowsPerBlock + threadIdx.x / G;
  const uint32_t gl = threadIdx.x % G;
  const uint32_t cnt = P.cls_off[cls + 1] - P.cls_off[cls];
  const bool active = slot < cnt;
  float z = 0.f;
  uint32_t i = 0;
  if (active) {
    i = P.group_rows[P.cls_off[cls] + slot];
    z = gather_sum<NT, G, kUnroll>(P, ld<NT>(P.rp + i) + gl, ld<NT>(P.rp + i + 1), z);
  }
  z = group_sum<G>(z);
  double e = (active && gl == 0) ? update_row<NT>(P, i, z) : 0.0;
  e = block_sum<kThreads>(e, red_d);
  if (threadIdx.x == 0) P.blk_err[P.n_tiles + P.cls_blk[cls] + blockIdx.x] = e;
}

// Hub segments: hub_segment nonzeros of one hub row per workgroup -> partial sum.
template <int NT>
__global__ __launch_bounds__(kThreads) void pr_hubseg_kernel(SweepParams P) {
  if (converged_before(P)) return;
  __shared__ float red_f[kThreads / 64];
  const uint32_t s = blockIdx.x;
  const uint32_t h = P.seg_row[s];
  const uint32_t i = P.hub_rows[h];
  const uint32_t p0 = P.rp[i] + (s - P.seg_ptr[h]) * P.hub_segment;
  const uint32_t pe = P.rp[i + 1];
  const uint32_t p1 = (pe - p0 > P.hub_segment) ? p0 + P.hub_segment : pe;
  float z = gather_sum<NT, kThreads, kUnroll>(P, p0 + threadIdx.x, p1, 0.f);
  z = block_sum<kThreads>(z, red_f);
  if (threadIdx.x == 0) P.hub_part[s] = z;
}

// Source-sliced rows of lane class `cls`: workgroup b gathers only from source slice
// b % kSlices.  Workgroups are dispatched to the XCDs round-robin, so every slice stays
// on one XCD and that XCD's L2 caches the contributions of its own slice only: the eight
// L2s hold eight different parts of the hot set instead of eight copies of one.
template <int G, int NT>
__global__ __launch_bounds__(kThreads) void pr_sgroup_kernel(SweepParams P, uint32_t cls) {
  if (converged_before(P)) return;
  constexpr uint32_t kRowsPerBlock = kThreads / G;
  const uint32_t x = blockIdx.x % kSlices, lb = blockIdx.x / kSlices;
  const uint32_t slot = lb * kRowsPerBlock + threadIdx.x / G;
  const uint32_t gl = threadIdx.x % G;
  const bool active = slot < P.scls_off[cls + 1] - P.scls_off[cls];
  float z = 0.f;
  uint32_t k = 0;
  if (active) {
    k = P.slist[P.scls_off[cls] + slot];
    const uint32_t* sg = P.sseg + (uint64_t)k * (kSlices + 1) + x;
    z = gather_sum<NT, G, kUnroll>(P, sg[0] + gl, sg[1], z);
  }
  z = group_sum<G>(z);
  if (active && gl == 0) P.spart[(uint64_t)x * P.n_srows + k] = z;
}

// Last launch of a sweep: finishes hub rows (segment partials added in segment order)
// and source-sliced rows (slice partials added in slice order), and folds the workgroup
// L1 partials of the sweep into acc[iter].  Only this kernel's few workgroups touch acc,
// so the atomics never queue up on one address.
template <int NT>
__global__ __launch_bounds__(kThreads) void pr_finish_kernel(SweepParams P) {
  if (converged_before(P)) return;
  __shared__ double red_d[kThreads / 64];
  const uint32_t h = blockIdx.x * kThreads + threadIdx.x;
  double e = 0.0;
  if (h < P.n_hub_rows) {
    float z = 0.f;
    const uint32_t s1 = P.seg_ptr[h + 1];
    for (uint32_t s = P.seg_ptr[h]; s < s1; ++s) z += P.hub_part[s];
    e = update_row<NT>(P, P.hub_rows[h], z);
  } else if (h < P.n_hub_rows + P.n_srows) {
    const uint32_t k = h - P.n_hub_rows;
    float z = 0.f;
#pragma unroll
    for (uint32_t x = 0; x < kSlices; ++x) z += P.spart[(uint64_t)x * P.n_srows + k];
    e = update_row<NT>(P, P.srows[k], z);
  }
  const uint32_t c0 = blockIdx.x * kFinishChunk;
  for (uint32_t q = c0 + threadIdx.x; q < c0 + kFinishChunk && q < P.hub_blk; q += kThreads) e += P.blk_err[q];
  e = block_sum<kThreads>(e, red_d);
  if (threadIdx.x == 0 && e > 0.0) atomicAdd(P.acc + P.iter, err_to_fixed(e));
}

// x = init, and with c0 contrib = init / outdeg (page_rank.hpp:100-104,114-118); acc cleared.
__global__ __launch_bounds__(kThreads) void pr_init_kernel(uint32_t n, const uint32_t* __restrict__ deg, float init,
                                                           float* __restrict__ x, float* __restrict__ c0,
                                                           unsigned long long* __restrict__ acc, uint64_t iters) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    x[i] = init;
    if (c0) c0[i] = init / (float)deg[i];
  }
  if (i < iters) acc[i] = 0ull;
}

// contrib = x / outdeg (page_rank.hpp:114-118): the contribution init on its own, when the
// two inits are timed separately
__global__ __launch_bounds__(kThreads) void pr_contrib_init_kernel(uint32_t n, const uint32_t* __restrict__ deg,
                                                                   const float* __restrict__ x,
                                                                   float* __restrict__ c0) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) c0[i] = x[i] / (float)deg[i];
}

// delta schedule (page_rank_v13, experimental/algorithms/page_rank.hpp:570-577): scores 0,
// deltas 0, every residual (float)(1.0 - d)
__global__ __launch_bounds__(kThreads) void pr_delta_init_kernel(uint32_t n, float r0, float* __restrict__ x,
                                                                 float* __restrict__ dlt, float* __restrict__ res,
                                                                 unsigned long long* __restrict__ acc,
                                                                 uint64_t iters) {
  const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    x[i] = 0.f;
    dlt[i] = 0.f;
    res[i] = r0;
  }
  if (i < iters) acc[i] = 0ull;
}

// delta schedule, outgoing phase (experimental/algorithms/page_rank.hpp:580-594): a residual
// above the threshold is absorbed into the score and, for a vertex with out-edges, turned
// into the delta r * d / deg its out-neighbours receive.  A workgroup in which some vertex
// pushed stores 1 in flags[iter * stride] (plain stores of one value: only zero / nonzero
// matters); the kernel does nothing after an iteration whose flag stayed 0.
__global__ __launch_bounds__(kThreads) void pr_delta_push_kernel(SweepParams P, float* __restrict__ res,
                                                                 float* __restrict__ dlt,
                                                                 unsigned long long* __restrict__ flags,
                                                                 uint32_t stride) {
  if (P.iter > 0 && flags[(uint64_t)(P.iter - 1) * stride] == 0ull) return;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const float thr = (float)P.threshold;
  bool pushed = false;
  if (i < P.n) {
    float dv = 0.f;
    const float r = res[i];
    if (r > thr) {
      P.x[i] += r;
      res[i] = 0.f;
      const uint32_t dg = P.deg[i];
      if (dg > 0) {
        dv = r * P.damp / (float)dg;
        pushed = true;
      }
    }
    dlt[i] = dv;
  }
  if (__syncthreads_or(pushed) && threadIdx.x == 0) flags[(uint64_t)P.iter * stride] = 1ull;
}

// Teleport check: out[0] += sum of v (double), out[1] = 1 when some entry is negative, above
// 1 or not finite.  The sweeps assume scores stay in [0, 1] (the binned sweep's exact row
// sums cover [0, 16), the fixed-point L1 change of a sweep [0, 32)), which holds for every
// teleport with entries >= 0 summing to at most 1.
__global__ __launch_bounds__(kThreads) void pr_teleport_check_kernel(uint32_t n, const float* __restrict__ v,
                                                                     double* __restrict__ out) {
  __shared__ double red_d[kThreads / 64];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  double t = 0.0;
  bool bad = false;
  if (i < n) {
    const float x = v[i];
    bad = !(x >= 0.f && x <= 1.f);  // also catches NaN
    t = bad ? 0.0 : (double)x;
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) out[1] = 1.0;
  t = block_sum<kThreads>(t, red_d);
  if (threadIdx.x == 0 && t > 0.0) atomicAdd(out, t);
}

// personalised base (1-d) * v rounded once from double; with a relabelled CSC row i
// is original vertex order[i]
__global__ __launch_bounds__(kThreads) void pr_base_kernel(uint32_t n, const float* __restrict__ v,
                                                           const uint32_t* __restrict__ order, double omd,
                                                           float* __restrict__ base) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) base[i] = (float)(omd * (double)v[order ? order[i] : i]);
}

// scores back to the original numbering: out[v] = x[rank[v]] (rank: original -> new id).
// A gather: the writes are whole lines, and the reads of x come mostly from MALL.
__global__ __launch_bounds__(kThreads) void pr_unpermute_kernel(uint32_t n, const uint32_t* __restrict__ rank,
                                                                const float* __restrict__ x, float* __restrict__ out) {
  const uint32_t v = blockIdx.x * kThreads + threadIdx.x;
  if (v < n) out[v] = x[rank[v]];
}

// ---------------------------------------------------------------------------
// plan construction
// ---------------------------------------------------------------------------
// class id of a row: 0 tile, 1..5 group classes (4..64 lanes), 6 hub, 7 source-sliced
// (slice_min <= d < hub_min; slice_min > short_max is a plan invariant)
__host__ __device__ __forceinline__ uint32_t row_class(uint32_t d, const nwg_plan_params& pp) {
  if (d <= pp.short_max) return 0;
  if (d >= pp.hub_min) return 6;
  if (pp.slice_min && d >= pp.slice_min) return 7;
  if (d <= pp.group_max) return 1u + (d > 16) + (d > 32) + (d > 64);
  return 5;
}

__global__ __launch_bounds__(kThreads) void class_flag_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                              nwg_plan_params pp, uint32_t cls,
                                                              uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) flag[i] = row_class(rp[i + 1] - rp[i], pp) == cls ? 1u : 0u;
}

__global__ __launch_bounds__(kThreads) void class_scatter_kernel(uint32_t n, const uint32_t* __restrict__ flag,
                                                                 const uint32_t* __restrict__ pos,
                                                                 uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = i;
}

__global__ __launch_bounds__(kThreads) void hub_nseg_kernel(uint32_t nh, const uint32_t* __restrict__ rows,
                                                            const uint32_t* __restrict__ rp, uint32_t seg,
                                                            uint32_t* __restrict__ nseg) {
  const uint32_t h = blockIdx.x * kThreads + threadIdx.x;
  if (h < nh) {
    const uint32_t i = rows[h];
    const uint32_t d = rp[i + 1] - rp[i];
    nseg[h] = (d + seg - 1) / seg;
  }
}

__global__ __launch_bounds__(kThreads) void hub_segrow_kernel(uint32_t nh, const uint32_t* __restrict__ seg_ptr,
                                                              uint32_t* __restrict__ seg_row) {
  const uint32_t h = blockIdx.x * kThreads + threadIdx.x;
  if (h < nh)
    for (uint32_t s = seg_ptr[h]; s < seg_ptr[h + 1]; ++s) seg_row[s] = h;
}

// lane class of a sliced row by its mean slice length (4, 8, 16, 32 or 64 lanes)
__global__ __launch_bounds__(kThreads) void slabel_kernel(uint32_t ns, const uint32_t* __restrict__ srows,
                                                          const uint32_t* __restrict__ rp,
                                                          uint32_t* __restrict__ label) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= ns) return;
  const uint32_t i = srows[k];
  const uint32_t len = (rp[i + 1] - rp[i] + kSlices - 1) / kSlices;
  label[k] = (len > 16) + (len > 32) + (len > 64) + (len > 128);
}

__global__ __launch_bounds__(kThreads) void label_flag_kernel(uint32_t m, const uint32_t* __restrict__ label,
                                                              uint32_t c, uint32_t* __restrict__ flag) {
  const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
  if (k < m) flag[k] = label[k] == c ? 1u : 0u;
}

// Sort key of CSC position p: its row, then (sliced rows only) the slice of its source.
// A stable sort of these keys carrying the column id keeps every row's sources
// ascending inside each slice and leaves the other rows as they are.
__global__ __launch_bounds__(kThreads) void slice_key_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ rp,
                                                             const uint32_t* __restrict__ col, nwg_plan_params pp,
                                                             uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= nnz) return;
  uint32_t lo = 0, hi = n;  // row of p: first r with rp[r + 1] > p
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rp[mid + 1] <= p) lo = mid + 1;
    else hi = mid;
  }
  const uint32_t u = col[p];
  const uint32_t x = row_class(rp[lo + 1] - rp[lo], pp) == 7 ? slice_of(u) : 0u;
  keys[p] = ((uint64_t)lo << 3) | x;
  vals[p] = u;
}

// sseg[k][x]: first position of slice x in sliced row k (binary search on the slice field)
__global__ __launch_bounds__(kThreads) void sseg_kernel(uint32_t ns, const uint32_t* __restrict__ srows,
                                                        const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ scol,
                                                        uint32_t* __restrict__ sseg) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (uint64_t)ns * (kSlices + 1)) return;
  const uint32_t k = (uint32_t)(t / (kSlices + 1)), x = (uint32_t)(t % (kSlices + 1));
  const uint32_t i = srows[k];
  uint32_t lo = rp[i], hi = rp[i + 1];
  if (x == kSlices) {
    lo = hi;
  } else {
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (slice_of(scol[mid]) < x) lo = mid + 1;
      else hi = mid;
    }
  }
  sseg[t] = lo;
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)ceil_div(n ? n : 1, kThreads); }

uint32_t d2h_u32(hipStream_t s, const uint32_t* p) {
  uint32_t v = 0;
  NWG_HIP(hipMemcpyAsync(&v, p, sizeof v, hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  return v;
}

// rows of class `cls` in ascending order -> out (allocated); returns count
uint32_t select_class(hipStream_t s, uint32_t n, const uint32_t* rp, const nwg_plan_params& pp, uint32_t cls,
                      DevBuf<uint32_t>& flag, DevBuf<uint32_t>& pos, uint32_t* out /*may be null to count*/) {
  class_flag_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, rp, pp, cls, flag.p);
  NWG_LAUNCH_CHECK();
  exclusive_scan_u32(s, flag.p, pos.p, n);
  const uint32_t cnt = d2h_u32(s, pos.p + n);
  if (out && cnt) {
    class_scatter_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, flag.p, pos.p, out);
    NWG_LAUNCH_CHECK();
  }
  return cnt;
}

}  // namespace

void build_segments(hipStream_t s, const uint32_t* rp, uint32_t n, uint32_t min_deg, uint32_t seg,
                    DevBuf<uint32_t>& rows, DevBuf<uint32_t>& seg_ptr, DevBuf<uint32_t>& seg_row, uint32_t* nrows,
                    uint32_t* nsegs) {
  // rows with in-degree >= min_deg form the hub class of a plan whose other cutoffs are below it
  nwg_plan_params pp{};
  pp.short_max = 0;
  pp.group_max = 0;
  pp.hub_min = min_deg;
  DevBuf<uint32_t> flag(n ? n : 1), pos((uint64_t)n + 1);
  const uint32_t nh = select_class(s, n, rp, pp, 6, flag, pos, nullptr);
  rows.alloc(nh ? nh : 1);
  seg_ptr.alloc((uint64_t)nh + 1);
  uint32_t ns = 0;
  if (nh) {
    select_class(s, n, rp, pp, 6, flag, pos, rows.p);
    hub_nseg_kernel<<<blocks_for(nh), kThreads, 0, s>>>(nh, rows.p, rp, seg, flag.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, flag.p, seg_ptr.p, nh);
    ns = d2h_u32(s, seg_ptr.p + nh);
    seg_row.alloc(ns);
    hub_segrow_kernel<<<blocks_for(nh), kThreads, 0, s>>>(nh, seg_ptr.p, seg_row.p);
    NWG_LAUNCH_CHECK();
  } else {
    const uint32_t zero = 0;
    NWG_HIP(hipMemcpyAsync(seg_ptr.p, &zero, 4, hipMemcpyHostToDevice, s));
    seg_row.alloc(1);
  }
  NWG_HIP(hipStreamSynchronize(s));
  *nrows = nh;
  *nsegs = ns;
}

nwg_plan_params default_plan_params() {
  nwg_plan_params p{};
  p.short_max = 8;
  p.group_max = 128;
  p.hub_min = 16384;
  p.hub_segment = 4096;
  p.slice_min = 0;
  return p;
}

// The plan a PageRank solve builds when none was set or tuned.  Measured on MI355X (RMAT,
// edge factor 16): the binned sweep beats the pull sweep from 2^16 vertices (1M edges: 0.023
// vs 0.048 ms per sweep) up (RMAT-20: 0.049 vs 0.164; RMAT-25: 0.94 vs 3.65), and on the
// 4096^2 grid; degree ordering pays from RMAT-19 (8M edges) on and costs a little below.
// The binned layout indexes entries and pairs with 32 bits (below 2^31 entries and 2^30
// pairs with pads), so larger graphs keep the pull sweep.
nwg_plan_params default_plan_for(const nwg_graph* g) {
  nwg_plan_params p = default_plan_params();
  if (g->nnz >= (1ull << 18) && g->nnz < (1ull << 30) - (1ull << 26)) p.binned = 1;
  if (g->nnz >= (1ull << 23)) p.relabel = 1;
  // nontemporal streams: the pair values the chunk pass stores are read once by the bin pass
  // (as cached stores, their write-back landed in the bin pass: RMAT-25 0.81 -> 0.72 ms)
  if (p.binned && g->nnz >= (1ull << 23)) p.stream_hints = 2;
  return p;
}

static void validate_params(const nwg_plan_params& p) {
  NWG_REQUIRE(p.short_max <= p.group_max, "plan: short_max must not exceed group_max");
  NWG_REQUIRE(p.group_max < p.hub_min, "plan: group_max must be below hub_min");
  NWG_REQUIRE(p.hub_segment >= 256 && p.hub_segment % 256 == 0, "plan: hub_segment must be a multiple of 256");
  NWG_REQUIRE(p.stream_hints <= 2, "plan: stream_hints must be 0, 1 or 2");
  NWG_REQUIRE(p.ppr_tile == 0 || p.ppr_tile == 8 || p.ppr_tile == 16 || p.ppr_tile == 32 || p.ppr_tile == 64,
              "plan: ppr_tile must be 0, 8, 16, 32 or 64");
  NWG_REQUIRE(p.relabel <= 1, "plan: relabel must be 0 or 1");
  NWG_REQUIRE(p.binned <= 1, "plan: binned must be 0 or 1");
  NWG_REQUIRE(p.bin_bits == 0 || p.bin_bits == 13 || p.bin_bits == 14, "plan: bin_bits must be 0, 13 or 14");
  NWG_REQUIRE(p.slice_min == 0 || (p.slice_min >= 2 && p.slice_min > p.short_max),
              "plan: slice_min must be 0 or above both 1 and short_max");
}

void plan_build(hipStream_t s, nwg_graph* g, const nwg_plan_params& pp) {
  validate_params(pp);
  const uint32_t n = g->n;
  if (pp.relabel) relabel_build(s, g);
  const uint32_t* rp = g->active(pp.relabel != 0).ptr;
  Plan pl;
  pl.params = pp;
  DevBuf<uint32_t> flag(n), pos((uint64_t)n + 1);
  uint32_t counts[kGroupClasses];
  uint32_t total = 0;
  for (int c = 0; c < kGroupClasses; ++c) {
    counts[c] = select_class(s, n, rp, pp, 1 + c, flag, pos, nullptr);
    total += counts[c];
  }
  pl.group_rows.alloc(total ? total : 1);
  pl.cls_off[0] = 0;
  for (int c = 0; c < kGroupClasses; ++c) {
    pl.cls_count[c] = counts[c];
    pl.cls_off[c + 1] = pl.cls_off[c] + counts[c];
    if (counts[c]) select_class(s, n, rp, pp, 1 + c, flag, pos, pl.group_rows.p + pl.cls_off[c]);
  }
  const uint32_t nh = select_class(s, n, rp, pp, 6, flag, pos, nullptr);
  pl.n_hub_rows = nh;
  pl.hub_rows.alloc(nh ? nh : 1);
  pl.hub_seg_ptr.alloc((uint64_t)nh + 1);
  pl.n_hub_segs = 0;
  if (nh) {
    select_class(s, n, rp, pp, 6, flag, pos, pl.hub_rows.p);
    hub_nseg_kernel<<<blocks_for(nh), kThreads, 0, s>>>(nh, pl.hub_rows.p, rp, pp.hub_segment, flag.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, flag.p, pl.hub_seg_ptr.p, nh);
    pl.n_hub_segs = d2h_u32(s, pl.hub_seg_ptr.p + nh);
    pl.hub_seg_row.alloc(pl.n_hub_segs);
    hub_segrow_kernel<<<blocks_for(nh), kThreads, 0, s>>>(nh, pl.hub_seg_ptr.p, pl.hub_seg_row.p);
    NWG_LAUNCH_CHECK();
  } else {
    pl.hub_seg_row.alloc(1);
  }
  const uint32_t ns = pp.slice_min ? select_class(s, n, rp, pp, 7, flag, pos, nullptr) : 0;
  pl.n_srows = ns;
  pl.srows.alloc(ns ? ns : 1);
  pl.sseg.alloc((uint64_t)(ns ? ns : 1) * (kSlices + 1));
  pl.slist.alloc(ns ? ns : 1);
  if (ns) {
    select_class(s, n, rp, pp, 7, flag, pos, pl.srows.p);
    const ActiveCsc a = g->active(pp.relabel != 0);
    const uint64_t nnz = g->nnz;
    pl.scol.alloc(nnz);
    {
      DevBuf<uint64_t> keys(nnz), keys_tmp(nnz);
      DevBuf<uint32_t> vals_tmp(nnz);
      slice_key_kernel<<<blocks_for(nnz), kThreads, 0, s>>>(n, nnz, rp, a.idx, pp, keys.p, pl.scol.p);
      NWG_LAUNCH_CHECK();
      radix_sort_pairs(s, keys.p, keys_tmp.p, pl.scol.p, vals_tmp.p, nnz, key_bits_for(n) + 3);
    }
    sseg_kernel<<<blocks_for((uint64_t)ns * (kSlices + 1)), kThreads, 0, s>>>(ns, pl.srows.p, rp, pl.scol.p,
                                                                              pl.sseg.p);
    NWG_LAUNCH_CHECK();
    DevBuf<uint32_t> label(ns);
    slabel_kernel<<<blocks_for(ns), kThreads, 0, s>>>(ns, pl.srows.p, rp, label.p);
    NWG_LAUNCH_CHECK();
    for (uint32_t c = 0; c < (uint32_t)kGroupClasses; ++c) {
      label_flag_kernel<<<blocks_for(ns), kThreads, 0, s>>>(ns, label.p, c, flag.p);
      NWG_LAUNCH_CHECK();
      exclusive_scan_u32(s, flag.p, pos.p, ns);
      const uint32_t cnt = d2h_u32(s, pos.p + ns);
      pl.scls_off[c + 1] = pl.scls_off[c] + cnt;
      if (cnt) {
        class_scatter_kernel<<<blocks_for(ns), kThreads, 0, s>>>(ns, flag.p, pos.p, pl.slist.p + pl.scls_off[c]);
        NWG_LAUNCH_CHECK();
      }
    }
  } else {
    pl.scol.release();
  }
  pl.n_tiles = (uint32_t)ceil_div(n, kThreads);
  if (pp.binned) binned_build(s, g->active(pp.relabel != 0), n, g->nnz, pl.bl, true, nullptr, (int)pp.bin_bits);
  NWG_HIP(hipStreamSynchronize(s));
  pl.built = true;
  g->plan = std::move(pl);
  g->plan_version++;
}

static constexpr uint32_t kLanes[kGroupClasses] = {4, 8, 16, 32, 64};

// workgroups of a sweep that precede the hub segments (tiles + group classes)
static uint64_t sweep_workgroups_before_hubs(const Plan& pl) {
  uint64_t b = pl.n_tiles;
  for (int c = 0; c < kGroupClasses; ++c) b += ceil_div(pl.cls_count[c], kThreads / kLanes[c]);
  return b;
}

static uint32_t finish_blocks(const SweepParams& P) {
  const uint64_t a = ceil_div((uint64_t)P.n_hub_rows + P.n_srows, kThreads), b = ceil_div(P.hub_blk, kFinishChunk);
  return (uint32_t)std::max<uint64_t>(std::max(a, b), 1);
}

static SweepParams make_params(nwg_graph* g, PrState& st, bool personalised, float damping, float threshold) {
  const Plan& pl = g->plan;
  SweepParams P{};
  const ActiveCsc a = g->active(pl.params.relabel != 0);
  P.rp = a.ptr;
  P.col = pl.n_srows ? pl.scol.p : a.idx;  // scol differs from the CSC only inside sliced rows
  P.deg = a.deg;
  P.x = st.x.p;
  P.basev = personalised ? st.base.p : nullptr;
  P.base = (float)((1.0 - (double)damping) / g->n);  // page_rank.hpp:94
  P.damp = damping;
  P.n = g->n;
  P.short_max = pl.params.short_max;
  P.n_tiles = pl.n_tiles;
  P.cls_blk[0] = 0;
  for (int c = 0; c < kGroupClasses; ++c) {
    const uint32_t rows_per_block = kThreads / kLanes[c];
    P.cls_blk[c + 1] = P.cls_blk[c] + (uint32_t)ceil_div(pl.cls_count[c], rows_per_block);
    P.cls_off[c] = pl.cls_off[c];
  }
  P.cls_off[kGroupClasses] = pl.cls_off[kGroupClasses];
  P.group_rows = pl.group_rows.p;
  P.hub_blk = P.n_tiles + P.cls_blk[kGroupClasses];
  P.n_hub_segs = pl.n_hub_segs;
  P.n_hub_rows = pl.n_hub_rows;
  P.hub_segment = pl.params.hub_segment;
  P.hub_rows = pl.hub_rows.p;
  P.seg_ptr = pl.hub_seg_ptr.p;
  P.seg_row = pl.hub_seg_row.p;
  P.hub_part = st.hub_part.p;
  P.blk_err = st.blk_err.p;
  P.srows = pl.srows.p;
  P.sseg = pl.sseg.p;
  P.slist = pl.slist.p;
  for (int c = 0; c <= kGroupClasses; ++c) P.scls_off[c] = pl.scls_off[c];
  P.n_srows = pl.n_srows;
  P.spart = st.spart.p;
  P.acc = st.acc.p;
  P.threshold = (double)threshold;  // page_rank.hpp:144 compares double error with (double)threshold
  return P;
}

// Enqueue init + max_iters sweeps on stream s.  With ev (2*iters + 4 events) every sweep
// is bracketed by an event pair, and the score and contribution inits (then two launches)
// by the last two pairs.
// one sweep: tiles, each non-empty group class, hub segments, finish
template <int NT>
static void enqueue_sweep(hipStream_t s, const SweepParams& P) {
  if (P.n_tiles) pr_tile_kernel<NT><<<P.n_tiles, kThreads, 0, s>>>(P);
  for (uint32_t c = 0; c < (uint32_t)kGroupClasses; ++c) {
    const uint32_t nb = P.cls_blk[c + 1] - P.cls_blk[c];
    if (!nb) continue;
    switch (c) {
      case 0: pr_group_kernel<4, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      case 1: pr_group_kernel<8, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      case 2: pr_group_kernel<16, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      case 3: pr_group_kernel<32, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      default: pr_group_kernel<64, NT><<<nb, kThreads, 0, s>>>(P, c); break;
    }
  }
  for (uint32_t c = 0; c < (uint32_t)kGroupClasses; ++c) {
    const uint32_t cnt = P.scls_off[c + 1] - P.scls_off[c];
    if (!cnt) continue;
    const uint32_t nb = kSlices * (uint32_t)ceil_div(cnt, kThreads / kLanes[c]);
    switch (c) {
      case 0: pr_sgroup_kernel<4, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      case 1: pr_sgroup_kernel<8, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      case 2: pr_sgroup_kernel<16, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      case 3: pr_sgroup_kernel<32, NT><<<nb, kThreads, 0, s>>>(P, c); break;
      default: pr_sgroup_kernel<64, NT><<<nb, kThreads, 0, s>>>(P, c); break;
    }
  }
  if (P.n_hub_segs) pr_hubseg_kernel<NT><<<P.n_hub_segs, kThreads, 0, s>>>(P);
  pr_finish_kernel<NT><<<finish_blocks(P), kThreads, 0, s>>>(P);
  NWG_LAUNCH_CHECK();
}

// Device parameters of the binned sweep over g's plan (contribution arrays set per sweep).
static BinParams make_bin_params(nwg_graph* g, PrState& st, const SweepParams& P) {
  const BinnedLayout& L = g->plan.bl;
  BinParams B{};
  B.brow = L.brow.p;
  B.bsrc = L.bsrc.p;
  B.bpid = L.bpid.p;
  B.vals = st.bvals.p;
  B.items2 = L.items2.p;
  B.items1 = L.items1.p;
  B.seg_start = L.seg_start.p;
  B.seg_flat = L.seg_flat.p;
  B.part = st.bpart.p;
  B.items3 = L.items3.p;
  B.n_items3 = L.n_items3;
  B.deg = P.deg;
  B.x = P.x;
  B.item_he = L.item_he.p;
  B.item_hd = L.item_hd.p;
  B.hsrc = L.hsrc.p;
  B.n_hot = L.n_hot;
  B.basev = P.basev;
  B.base = P.base;
  B.damp = P.damp;
  B.n = g->n;
  B.n_items2 = L.n_items2;
  B.n_items2_live = L.n_items2_live;
  B.bin_bits = L.bin_bits;
  B.n_items1 = L.n_items1;
  B.acc = st.bacc.p;
  B.threshold = P.threshold;
  B.hints = g->plan.params.stream_hints;
  return B;
}

// Sweeps [k0, k1) of an iters-sweep solve: the inits when k0 = 0, and for the binned sweep
// the fold of these sweeps' residual lanes into acc.
static void enqueue_solve(hipStream_t s, nwg_graph* g, PrState& st, SweepParams P, uint64_t iters,
                          hipEvent_t* ev, PrSchedule sched, uint64_t k0, uint64_t k1) {
  const uint32_t n = g->n;
  const uint32_t nt = g->plan.params.stream_hints;
  // In place, the binned sweep reads and writes one contribution array: its chunk pass stages
  // every non-hot contribution before any row of the sweep is finished, so updates made in
  // the sweep reach other rows only through the hot sources, which each bin-pass work item
  // stages when it starts.  The delta schedule keeps the pull kernels: the binned sweep's
  // exact fixed-point row sums cover [0, 16), enough for PageRank (a row sum is at most the
  // total score, 1) but not for un-normalised residuals, which reach thousands on hub rows.
  const bool binned = g->plan.params.binned != 0 && sched != PrSchedule::kDelta;
  BinParams B{};
  if (binned) {
    B = make_bin_params(g, st, P);
    if (k0 == 0) binned_clear_residuals(s, st.bacc.p, iters);
  }
  auto sweep = [&](const float* cin, float* cout) {
    if (binned) {
      B.iter = P.iter;
      B.cin = cin;
      B.cout = cout;
      binned_enqueue_sweep(s, B);
      return;
    }
    P.cin = cin;
    P.cout = cout;
    if (nt == 2) enqueue_sweep<2>(s, P);
    else if (nt == 1) enqueue_sweep<1>(s, P);
    else enqueue_sweep<0>(s, P);
  };
  if (sched == PrSchedule::kDelta) {
    // page_rank_v13: contrib0 holds the deltas, contrib1 the residuals, x the scores; the
    // push kernel's flags go to acc, where the pull kernels' stopping test reads them
    P.delta = 1;
    if (k0 == 0) {
      if (ev) NWG_HIP(hipEventRecord(ev[2 * iters], s));
      pr_delta_init_kernel<<<blocks_for(std::max<uint64_t>(n, iters)), kThreads, 0, s>>>(
          n, (float)(1.0 - (double)P.damp), st.x.p, st.contrib0.p, st.contrib1.p, st.acc.p, iters);
      NWG_LAUNCH_CHECK();
      if (ev) {
        NWG_HIP(hipEventRecord(ev[2 * iters + 1], s));
        NWG_HIP(hipEventRecord(ev[2 * iters + 2], s));
        NWG_HIP(hipEventRecord(ev[2 * iters + 3], s));
      }
    }
    for (uint64_t k = k0; k < k1; ++k) {
      P.iter = (uint32_t)k;
      if (ev) NWG_HIP(hipEventRecord(ev[2 * k], s));
      pr_delta_push_kernel<<<blocks_for(n), kThreads, 0, s>>>(P, st.contrib1.p, st.contrib0.p, st.acc.p, 1u);
      NWG_LAUNCH_CHECK();
      sweep(st.contrib0.p, st.contrib1.p);
      if (ev) NWG_HIP(hipEventRecord(ev[2 * k + 1], s));
    }
    return;
  }
  const bool inplace = sched == PrSchedule::kInPlace;
  const float init = (float)(1.0 / n);  // page_rank.hpp:93
  if (k0 == 0) {
    if (ev) {
      NWG_HIP(hipEventRecord(ev[2 * iters], s));
      pr_init_kernel<<<blocks_for(std::max<uint64_t>(n, iters)), kThreads, 0, s>>>(n, P.deg, init, st.x.p, nullptr,
                                                                                   st.acc.p, iters);
      NWG_HIP(hipEventRecord(ev[2 * iters + 1], s));
      NWG_HIP(hipEventRecord(ev[2 * iters + 2], s));
      pr_contrib_init_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, P.deg, st.x.p, st.contrib0.p);
      NWG_HIP(hipEventRecord(ev[2 * iters + 3], s));
    } else {
      pr_init_kernel<<<blocks_for(std::max<uint64_t>(n, iters)), kThreads, 0, s>>>(n, P.deg, init, st.x.p,
                                                                                   st.contrib0.p, st.acc.p, iters);
    }
    NWG_LAUNCH_CHECK();
  }
  for (uint64_t k = k0; k < k1; ++k) {
    P.iter = (uint32_t)k;
    if (ev) NWG_HIP(hipEventRecord(ev[2 * k], s));
    // in place (page_rank.hpp:135): one contribution array, each row's new contribution is
    // visible to the rows gathered after it; else double-buffered (Jacobi)
    if (inplace) sweep(st.contrib0.p, st.contrib0.p);
    else if (k & 1) sweep(st.contrib1.p, st.contrib0.p);
    else sweep(st.contrib0.p, st.contrib1.p);
    if (ev) NWG_HIP(hipEventRecord(ev[2 * k + 1], s));
  }
  if (binned && k1 > k0)
    binned_fold_residuals(s, st.bacc.p + k0 * kBinAccStride, st.acc.p + k0, k1 - k0);
}

void pagerank_run(nwg_ctx* ctx, nwg_graph* g, const float* teleport_dev, float damping, float threshold,
                  uint64_t max_iters, nwg_pr_info* info, float* scores_dev, PrSchedule sched) {
  NWG_REQUIRE(sched != PrSchedule::kDelta || teleport_dev == nullptr, "the delta schedule takes no teleport vector");
  hipStream_t s = ctx->stream;
  NWG_REQUIRE(max_iters < (1ull << 31), "max_iters too large");
  if (!g->plan.built) plan_build(s, g, default_plan_for(g));
  if (!g->pr) g->pr.reset(new PrState());
  PrState& st = *g->pr;
  const uint32_t n = g->n;
  st.x.ensure(n);
  st.contrib0.ensure(n);
  st.contrib1.ensure(n);
  st.acc.ensure(std::max<uint64_t>(max_iters, 1));
  st.hub_part.ensure(std::max<uint32_t>(g->plan.n_hub_segs, 1));
  st.blk_err.ensure(std::max<uint64_t>(sweep_workgroups_before_hubs(g->plan), 1));
  st.spart.ensure(std::max<uint64_t>((uint64_t)kSlices * g->plan.n_srows, 1));
  if (g->plan.params.binned && sched != PrSchedule::kDelta) {
    const BinnedLayout& L = g->plan.bl;
    st.bvals.ensure(L.pairs);
    st.bpart.ensure(std::max<uint64_t>((uint64_t)L.n_slots << L.bin_bits, 1));
    st.bacc.ensure(std::max<uint64_t>(max_iters, 1) * kBinAccStride);
  }
  const bool personalised = teleport_dev != nullptr;
  if (personalised) {
    st.tcheck.ensure(2);
    NWG_HIP(hipMemsetAsync(st.tcheck.p, 0, 2 * sizeof(double), s));
    pr_teleport_check_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, teleport_dev, st.tcheck.p);
    NWG_LAUNCH_CHECK();
    double chk[2] = {0.0, 0.0};
    NWG_HIP(hipMemcpyAsync(chk, st.tcheck.p, sizeof chk, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    NWG_REQUIRE(chk[1] == 0.0, "teleport entries must be finite and in [0, 1]");
    NWG_REQUIRE(chk[0] <= 1.0 + kTeleportSlack, "teleport entries must sum to at most 1");
    st.base.ensure(n);
    pr_base_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, teleport_dev, g->active(g->plan.params.relabel != 0).order,
                                                      1.0 - (double)damping, st.base.p);
    NWG_LAUNCH_CHECK();
  }
  const SweepParams P = make_params(g, st, personalised, damping, threshold);
  const bool timing = info && (info->flags & NWG_INFO_KERNEL_TIMING);
  float ms_total = 0.f;
  double sweep_ms = 0.0;
  std::vector<hipEvent_t> evs;
  if (!timing) {
    CaptureKey key;
    key.ptrs[0] = st.x.p;
    key.ptrs[1] = st.contrib0.p;
    key.ptrs[2] = st.contrib1.p;
    key.ptrs[3] = st.acc.p;
    key.ptrs[4] = st.hub_part.p;
    key.ptrs[5] = personalised ? st.base.p : nullptr;
    key.ptrs[6] = P.rp;
    key.ptrs[7] = st.blk_err.p;
    key.ptrs[8] = st.spart.p;
    key.ptrs[9] = st.bvals.p;
    key.ptrs[10] = st.bpart.p;
    key.ptrs[11] = st.bacc.p;
    key.plan_version = g->plan_version;
    key.iters = max_iters;
    key.damping = damping;
    key.threshold = threshold;
    key.extra = (uint32_t)sched;
    // the solve's segments (kSegFirst sweeps, then kSegNext at a time), each one captured
    // graph; between segments the host reads the last sweep's L1 change and stops once the
    // solve has converged, so sweeps after convergence launch only up to a segment's end
    uint64_t nseg = 0;
    for (uint64_t k0 = 0;; k0 = segment_end(k0, max_iters)) {
      const uint64_t k1 = segment_end(k0, max_iters);
      if (st.caps.size() <= nseg) st.caps.emplace_back(new Captured());
      Captured& cap = *st.caps[nseg++];
      if (!cap.valid || !(cap.key == key)) {
        cap.reset();
        hipStream_t cs = ctx->capture_stream;
        NWG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_solve(cs, g, st, P, max_iters, nullptr, sched, k0, k1);
        } catch (...) {
          hipGraph_t tmp = nullptr;
          (void)hipStreamEndCapture(cs, &tmp);
          if (tmp) (void)hipGraphDestroy(tmp);
          throw;
        }
        NWG_HIP(hipStreamEndCapture(cs, &cap.graph));
        NWG_HIP(hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0));
        cap.key = key;
        cap.valid = true;
      }
      if (k1 >= max_iters) break;
    }
    NWG_HIP(hipEventRecord(ctx->ev0, s));
    uint64_t seg = 0;
    for (uint64_t k0 = 0;; k0 = segment_end(k0, max_iters)) {
      const uint64_t k1 = segment_end(k0, max_iters);
      NWG_HIP(hipGraphLaunch(st.caps[seg++]->exec, s));
      if (k1 >= max_iters) break;
      unsigned long long last = 0ull;
      NWG_HIP(hipMemcpyAsync(&last, st.acc.p + (k1 - 1), sizeof last, hipMemcpyDeviceToHost, s));
      NWG_HIP(hipStreamSynchronize(s));
      // stopped: delta schedule, nothing pushed (page_rank.hpp:616-620); else the L1 change
      // below the threshold (page_rank.hpp:144) -- or the solve stopped earlier and this
      // sweep's accumulator kept the 0 it was cleared to
      if (sched == PrSchedule::kDelta ? last == 0ull : fixed_to_err(last) < (double)threshold) break;
    }
    NWG_HIP(hipEventRecord(ctx->ev1, s));
  } else {
    evs.resize(2 * max_iters + 4);
    for (auto& e : evs) NWG_HIP(hipEventCreate(&e));
    NWG_HIP(hipEventRecord(ctx->ev0, s));
    try {
      enqueue_solve(s, g, st, P, max_iters, evs.data(), sched, 0, max_iters);
    } catch (...) {
      for (auto& e : evs) (void)hipEventDestroy(e);
      throw;
    }
    NWG_HIP(hipEventRecord(ctx->ev1, s));
  }
  std::vector<unsigned long long> acc(std::max<uint64_t>(max_iters, 1));
  NWG_HIP(hipMemcpyAsync(acc.data(), st.acc.p, max_iters * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  if (scores_dev) {
    const uint32_t* order = g->active(g->plan.params.relabel != 0).order;
    if (order) {
      pr_unpermute_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, g->rl->rank.p, st.x.p, scores_dev);
      NWG_LAUNCH_CHECK();
    } else {
      NWG_HIP(hipMemcpyAsync(scores_dev, st.x.p, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
  }
  NWG_HIP(hipStreamSynchronize(s));
  NWG_HIP(hipEventElapsedTime(&ms_total, ctx->ev0, ctx->ev1));
  uint64_t iters = max_iters;
  double last = 0.0;
  for (uint64_t k = 0; k < max_iters; ++k) {
    if (sched == PrSchedule::kDelta) {
      // the reference traces error 0 for this schedule (page_rank.hpp:615); the loop ends
      // after the first iteration that pushed nothing
      if (acc[k] == 0ull) {
        iters = k + 1;
        break;
      }
      continue;
    }
    const double e = fixed_to_err(acc[k]);
    last = e;
    if (e < (double)threshold) {
      iters = k + 1;
      break;
    }
  }
  double init_ms[2] = {0.0, 0.0};
  if (timing) {
    for (uint64_t k = 0; k < iters; ++k) {
      float t = 0.f;
      NWG_HIP(hipEventElapsedTime(&t, evs[2 * k], evs[2 * k + 1]));
      sweep_ms += t;
      if (info->time_trace && k < info->time_trace_len) info->time_trace[k] = t;
    }
    for (int q = 0; q < 2; ++q) {
      float t = 0.f;
      NWG_HIP(hipEventElapsedTime(&t, evs[2 * max_iters + 2 * q], evs[2 * max_iters + 2 * q + 1]));
      init_ms[q] = t;
    }
    for (auto& e : evs) (void)hipEventDestroy(e);
    sweep_ms = iters ? sweep_ms / (double)iters : 0.0;
  }
  if (info) {
    info->iterations = iters;
    info->final_error = max_iters ? last : 0.0;
    info->ms_total = ms_total;
    info->ms_per_iter = iters ? ms_total / (double)iters : 0.0;
    info->sweep_ms_avg = sweep_ms;
    info->init_ms[0] = init_ms[0];
    info->init_ms[1] = init_ms[1];
    if (info->error_trace)
      for (uint64_t k = 0; k < iters && k < info->error_trace_len; ++k)
        info->error_trace[k] = sched == PrSchedule::kDelta ? 0.0 : fixed_to_err(acc[k]);
  }
}

}  // namespace nwg
