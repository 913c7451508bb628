// ppr.hip -- k personalised PageRank vectors as one SpMM per sweep.
//
// Each vector j is the reference's PageRank loop (page_rank_v10 Jacobi schedule,
// include/nwgraph/experimental/algorithms/page_rank.hpp:377-432) with the uniform
// base (1-d)/N replaced by (1-d) v_j, v_j uniform over a small seed set, and its own
// stopping rule (first sweep with L1 change < threshold).  Contributions are tile-major,
// so one in-edge u of row i gathers the contiguous T-float slice C[tile][u][0..T) of the
// k-tile being served (T = 8/16/32/64): a T-lane slice of a wavefront loads it in one
// coalesced access.  Workgroup b serves tile b % tiles, so with tiles dividing 8 each
// XCD's L2 caches one tile's contributions (T floats per source) instead of all k.
// Scores are not kept during the solve: a row's old score is read back from its previous
// contribution (x / deg, or x itself for a row without out-edges), and the row-major N x k
// scores are written once at the end.
//
// Work split per sweep (one launch + one finishing launch):
//   row workgroups   (256/T) lane groups x 8 (T = 8: 4) consecutive rows x one k-tile; a
//                    group walks the in-edges of its rows as one flat list over a compacted
//                    copy of the light rows (in-degree <= 256), T ids per coalesced load
//   heavy segments   (256/T)*32 in-edges of one heavier row x one k-tile; lane group g
//                    takes 32 contiguous edges with all 32 gathers in flight -> partial
//                    vectors
//   finish           hub rows (partials added in segment order) and per-vector L1
//                    partials folded into acc[iter][j] (fixed point, order independent)
// A vector whose previous sweep converged is skipped (its lanes idle); a workgroup
// whose k-tile is entirely converged exits at once.  From the third sweep on, the rows
// after the last row with in-edges (their values fixed after one sweep) are not launched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kSeedFlag = 0x80000000u;
// gathers a lane issues before it adds them (in edge order): memory-level parallelism
constexpr uint32_t kUnroll = 8;
// (heavy row, vector) items per finishing workgroup, and the vectors whose L1 changes that
// workgroup first adds in LDS
constexpr uint32_t kFinItems = 2048;
constexpr uint32_t kFinMaxK = 256;
constexpr uint32_t kFoldRows = 128;  // row-workgroup L1 partials folded per finishing workgroup
constexpr uint32_t kFinU = 4;        // heavy-row items a finishing thread has in flight
// edges of a heavy-row segment per lane group (all their gathers are in flight at once)
constexpr uint32_t kSegPerGroup = 32;
// consecutive light rows per lane group (lanes 0..kR of the group hold their row pointers)
template <int T>
struct PprRows {
  static constexpr uint32_t value = T >= 16 ? 8u : 4u;
};

struct PprParams {
  const uint32_t* rp;
  const uint32_t* col;
  const uint32_t* lrp;   // light CSC: rows with in-degree <= light_max, heavy rows empty
  const uint32_t* lcol;
  const uint32_t* degflag;   // out-degree | kSeedFlag when the row is a seed of some vector
  const uint32_t* seed_row;  // sorted (row, vector) seed entries
  const uint32_t* seed_vec;
  const float* seed_base;    // (1-d) v_j[row] for the entry
  uint32_t n_seed;
  const float* cin;
  float* cout;
  uint32_t n, k;
  uint32_t tiles, kt_shift;  // k-tiles and log2 of their width
  float damp;
  uint32_t light_max, hub_segment;  // rows with in-degree <= light_max take the row path
  uint32_t row_blocks;       // workgroups of rows per k-tile
  uint32_t n_hub_rows, n_hub_segs;
  const uint32_t* hub_rows;
  const uint32_t* seg_ptr;
  const uint32_t* seg_row;
  const uint32_t* hcol;        // heavy rows' in-edges regrouped by (row, source slice)
  const uint32_t* seg_start;   // segment s covers hcol[seg_start[s], seg_end[s])
  const uint32_t* seg_end;
  const uint32_t* slice_segs;  // segment ids of slice x: slice_segs[slice_off[x] .. slice_off[x + 1])
  uint32_t slice_off[9];
  uint32_t n_slices;           // source slices per k-tile; tiles * n_slices <= 8 workgroup classes
  uint32_t heavy_blocks;       // heavy-segment workgroups (8 per slice segment slot), launched first
  uint32_t row_blocks_run;     // row workgroups per k-tile this sweep runs (and whose L1 partials are folded)
  float* hub_part;           // n_hub_segs x k
  double* blk_err;           // (row_blocks * tiles) x T
  unsigned long long* acc;   // iters x k
  uint32_t iter;
  double threshold;
};

__device__ __forceinline__ bool vec_done(const PprParams& P, uint32_t j) {
  return P.iter > 0 && fixed_to_err(P.acc[(uint64_t)(P.iter - 1) * P.k + j]) < P.threshold;
}

__device__ __forceinline__ float seed_base(const PprParams& P, uint32_t i, uint32_t j) {
  // lower_bound of i in seed_row, then scan the row's entries for vector j
  uint32_t lo = 0, hi = P.n_seed;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (P.seed_row[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  float b = 0.f;
  for (uint32_t q = lo; q < P.n_seed && P.seed_row[q] == i; ++q)
    if (P.seed_vec[q] == j) b = P.seed_base[q];
  return b;
}

// Contributions are stored tile-major, [tile][row][T] (T = 1 << kt_shift vectors per tile),
// so that the gathers of one k-tile touch only that tile's array; scores stay row-major.
__device__ __forceinline__ uint64_t cidx(const PprParams& P, uint32_t i, uint32_t j) {
  return (((uint64_t)(j >> P.kt_shift) * P.n + i) << P.kt_shift) + (j & ((1u << P.kt_shift) - 1));
}

// Scores are not stored during the solve: the contribution array holds x / deg for a row with
// out-edges and x itself for a row without (no in-edge ever gathers it), and a row's old score
// is read back from the contributions of the previous sweep.
__device__ __forceinline__ float ppr_score(float c, uint32_t df) {
  const uint32_t d = df & ~kSeedFlag;
  return d ? c * (float)d : c;
}

// update of row i, vector j from its sum z, its degree/seed word df and x_old xo (the score
// read back from the previous sweep's contribution, within an ulp of the score)
__device__ __forceinline__ double ppr_finish_row(const PprParams& P, uint32_t i, uint32_t j, float z, uint32_t df,
                                                 float xo) {
  const float b = (df & kSeedFlag) ? seed_base(P, i, j) : 0.f;
  const float xn = __builtin_fmaf(P.damp, z, b);
  const uint32_t d = df & ~kSeedFlag;
  P.cout[cidx(P, i, j)] = d ? xn / (float)d : xn;
  return (double)fabsf(xn - xo);
}

// per-vector sum over the 64/T row slots of a wave, then over the 4 waves; result for
// vector lane vl (< T) in threads [0, T)
template <int T>
__device__ __forceinline__ double slot_reduce(double e, double* lds /* 4*T */) {
#pragma unroll
  for (int o = T; o < 64; o <<= 1) e += __shfl_xor(e, o, 64);
  const uint32_t w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l < (uint32_t)T) lds[w * T + l] = e;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < (uint32_t)T)
    for (int q = 0; q < 4; ++q) r += lds[q * T + threadIdx.x];
  __syncthreads();
  return r;
}

template <int T>
__global__ __launch_bounds__(kThreads) void ppr_sweep_kernel(PprParams P) {
  __shared__ double red[4 * T];
  __shared__ float hub_red[kThreads];
  __shared__ float zbuf[kThreads * PprRows<T>::value];
  __shared__ int any_live;
  // Heavy-segment workgroups come first (a multiple of 8 of them); workgroup class c = b % 8
  // (its XCD: workgroup b runs on XCD b % 8) serves k-tile c % tiles and source slice
  // c / tiles, so that each of the XCDs of one k-tile caches only its slice of the sources.
  // Row workgroup lb = b - heavy_blocks serves k-tile lb % tiles: with tiles dividing 8,
  // every XCD serves one k-tile, and its L2 caches that tile's contributions only.
  const bool heavy = blockIdx.x < P.heavy_blocks;
  uint32_t tile, b, seg = 0;
  if (!heavy) {
    const uint32_t lb = blockIdx.x - P.heavy_blocks;
    tile = lb % P.tiles;
    b = lb / P.tiles;
    if (b >= P.row_blocks_run) return;
  } else {
    const uint32_t hb = blockIdx.x, cls = hb % 8u;
    if (cls >= P.tiles * P.n_slices) return;
    tile = cls % P.tiles;
    const uint32_t x = cls / P.tiles, q = hb / 8u;
    if (q >= P.slice_off[x + 1] - P.slice_off[x]) return;
    seg = P.slice_segs[P.slice_off[x] + q];
    b = 0;
  }
  const uint32_t vl = threadIdx.x % T;
  const uint32_t j = tile * T + vl;
  const bool live = j < P.k && !vec_done(P, j);
  if (threadIdx.x == 0) any_live = 0;
  __syncthreads();
  if (live) any_live = 1;
  __syncthreads();
  if (!any_live) return;
  const float* __restrict__ ctile = P.cin + (uint64_t)tile * P.n * T;
  if (!heavy) {
    // Light rows: each T-lane group takes kR consecutive rows and walks their in-edges as
    // one flat list over the compacted light CSC (heavy rows are empty there), T column
    // ids per coalesced load and kUnroll gathers in flight; each row's sum (in edge order)
    // goes to LDS, and the kR row updates are issued together at the end.  Every test
    // that steers the loops is uniform over the group, so all its lanes take part in the
    // shuffles (converged vectors gather with their group and only skip their updates).
    constexpr uint32_t kR = PprRows<T>::value;
    constexpr uint32_t kSlots = kThreads / T;
    const uint32_t slot = threadIdx.x / T;
    const uint32_t g0 = (b * kSlots + slot) * kR;
    double e = 0.0;
    if (g0 < P.n) {
      const uint32_t nr = min(kR, P.n - g0);
      const uint32_t q = min(vl, nr);  // lanes 0..nr hold the row pointers of rows g0..g0+nr
      const uint32_t my_lrp = P.lrp[g0 + q];
      const uint32_t my_rp = P.rp[g0 + q];
      const uint32_t pend = (uint32_t)__shfl((int)my_lrp, (int)nr, T);
      uint32_t r = 0;
      uint32_t rend = (uint32_t)__shfl((int)my_lrp, 1, T);
      float z = 0.f;
      for (uint32_t p = (uint32_t)__shfl((int)my_lrp, 0, T); p < pend; p += T) {
        const uint32_t m = min((uint32_t)T, pend - p);
        const uint32_t c = vl < m ? P.lcol[p + vl] : 0u;
        // Every lane gathers: lanes of converged vectors share their group's line, and edges
        // past m read source 0 (c is 0 there).  Loads that a lane condition skipped made the
        // compiler drain every load in flight before each add (4 % of the solve at RMAT-22).
        // (Issuing the next round's gathers before adding this one's took 84 VGPRs instead of
        // 54 and was 17 % slower.)
        for (uint32_t q0 = 0; q0 < m; q0 += kUnroll) {  // q0 + t < T: q0 is a multiple of 8 below m <= T
          float v[kUnroll];
#pragma unroll
          for (uint32_t t = 0; t < kUnroll; ++t) {
            const uint32_t u = (uint32_t)__shfl((int)c, (int)(q0 + t), T);
            v[t] = ctile[(uint64_t)u * T + vl];
          }
#pragma unroll
          for (uint32_t t = 0; t < kUnroll; ++t) {
            if (q0 + t >= m) break;
            while (p + q0 + t >= rend) {  // row r ends before this edge (r + 1 < nr here)
              zbuf[(slot * kR + r) * T + vl] = z;
              z = 0.f;
              ++r;
              rend = (uint32_t)__shfl((int)my_lrp, (int)(r + 1), T);
            }
            z += v[t];  // a converged vector's sums are never used
          }
        }
      }
      for (; r < nr; ++r) {
        zbuf[(slot * kR + r) * T + vl] = z;
        z = 0.f;
      }
      float xo[kR], zz[kR];
      uint32_t df[kR];
      bool light[kR];
#pragma unroll
      for (uint32_t t = 0; t < kR; ++t) {
        const uint32_t d = (uint32_t)__shfl((int)my_rp, (int)(t + 1), T) - (uint32_t)__shfl((int)my_rp, (int)t, T);
        light[t] = t < nr && live && d <= P.light_max;  // heavy rows are finished by the segment path
        if (light[t]) {
          const uint32_t i = g0 + t;
          df[t] = P.degflag[i];
          xo[t] = P.cin[cidx(P, i, j)];
          zz[t] = zbuf[(slot * kR + t) * T + vl];
        }
      }
#pragma unroll
      for (uint32_t t = 0; t < kR; ++t)
        if (light[t]) e += ppr_finish_row(P, g0 + t, j, zz[t], df[t], ppr_score(xo[t], df[t]));
    }
    e = slot_reduce<T>(e, red);
    if (threadIdx.x < (uint32_t)T) P.blk_err[((uint64_t)tile * P.row_blocks + b) * T + threadIdx.x] = e;
  } else {
    // heavy segment s of one k-tile (at most hub_segment edges of one row and one source
    // slice): edge group g takes the contiguous kSegPerGroup edges [p0 + g*kSegPerGroup, ...),
    // loads their ids in one pass and issues all the gathers before adding them in edge
    // order; the group partials are added in group order
    const uint32_t s = seg;
    const uint32_t ps = P.seg_start[s], p1 = P.seg_end[s];
    const uint32_t p0 = ps + (threadIdx.x / T) * kSegPerGroup;
    constexpr uint32_t kIdLoads = (kSegPerGroup + T - 1) / T;
    uint32_t c[kIdLoads];
#pragma unroll
    for (uint32_t t = 0; t < kIdLoads; ++t) {
      const uint32_t pp = p0 + t * T + vl;
      c[t] = (t * T + vl < kSegPerGroup && pp < p1) ? P.hcol[pp] : 0u;
    }
    float v[kSegPerGroup];
#pragma unroll
    for (uint32_t t = 0; t < kSegPerGroup; ++t) {
      const uint32_t u = (uint32_t)__shfl((int)c[t / T], (int)(t % T), T);
      v[t] = ctile[(uint64_t)u * T + vl];  // every lane: past p1, c is 0 (source 0); see the light rows
    }
    float z = 0.f;
#pragma unroll
    for (uint32_t t = 0; t < kSegPerGroup; ++t)
      if (p0 + t < p1) z += v[t];
    hub_red[threadIdx.x] = z;
    __syncthreads();
    if (threadIdx.x < (uint32_t)T) {
      constexpr uint32_t kGroups = kThreads / T;
      float t = 0.f;
      for (uint32_t g = 0; g < kGroups; ++g) t += hub_red[g * T + threadIdx.x];
      if (j < P.k) P.hub_part[(uint64_t)s * P.k + j] = t;
    }
  }
}

// finishing launch: hub rows + fold of the per-workgroup, per-vector L1 partials.
// grid.x covers max(hub rows * k / 256, partial chunks); each thread owns one (item, vector).
template <int T>
__global__ __launch_bounds__(kThreads) void ppr_finish_kernel(PprParams P, uint32_t tiles) {
  __shared__ double red[kThreads];
  __shared__ unsigned long long errj[kFinMaxK];
  // part 1: heavy rows, kFinItems (heavy row, vector) items per workgroup; their L1 changes
  // are added per vector in LDS (fixed point: order independent), then one atomic per vector
  for (uint32_t q = threadIdx.x; q < kFinMaxK; q += kThreads) errj[q] = 0ull;
  __syncthreads();
  const uint64_t hub_items = (uint64_t)P.n_hub_rows * P.k;
  // kFinU items per thread at a time, every item's first loads (segment bounds, then the
  // first partial, degree word and old score) issued before any is used
  for (uint32_t m0 = 0; m0 < kFinItems; m0 += kFinU * kThreads) {
    bool ok[kFinU];
    uint32_t jv[kFinU], row[kFinU], s0[kFinU], s1[kFinU], df[kFinU];
    float z[kFinU], xo[kFinU];
#pragma unroll
    for (uint32_t u = 0; u < kFinU; ++u) {
      const uint64_t t = (uint64_t)blockIdx.x * kFinItems + m0 + u * kThreads + threadIdx.x;
      jv[u] = (uint32_t)(t % P.k);
      ok[u] = t < hub_items && !vec_done(P, jv[u]);
      if (ok[u]) {
        const uint32_t h = (uint32_t)(t / P.k);
        s0[u] = P.seg_ptr[h];
        s1[u] = P.seg_ptr[h + 1];  // > s0: a heavy row has at least one segment
        row[u] = P.hub_rows[h];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kFinU; ++u) {
      if (ok[u]) {
        z[u] = P.hub_part[(uint64_t)s0[u] * P.k + jv[u]];
        df[u] = P.degflag[row[u]];
        xo[u] = P.cin[cidx(P, row[u], jv[u])];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kFinU; ++u) {
      if (!ok[u]) continue;
      // the other segment partials (rows of many segments), added in segment order with 32
      // loads in flight
      for (uint32_t s = s0[u] + 1; s < s1[u]; s += 32) {
        float v[32];
#pragma unroll
        for (uint32_t r = 0; r < 32; ++r) v[r] = s + r < s1[u] ? P.hub_part[(uint64_t)(s + r) * P.k + jv[u]] : 0.f;
#pragma unroll
        for (uint32_t r = 0; r < 32; ++r)
          if (s + r < s1[u]) z[u] += v[r];
      }
      const double e = ppr_finish_row(P, row[u], jv[u], z[u], df[u], ppr_score(xo[u], df[u]));
      if (e > 0.0) {
        if (jv[u] < kFinMaxK) atomicAdd(&errj[jv[u]], err_to_fixed(e));
        else atomicAdd(P.acc + (uint64_t)P.iter * P.k + jv[u], err_to_fixed(e));
      }
    }
  }
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < kFinMaxK && q < P.k; q += kThreads)
    if (errj[q]) atomicAdd(P.acc + (uint64_t)P.iter * P.k + q, errj[q]);
  // part 2: fold blk_err rows; this workgroup folds kFoldRows workgroup rows of every tile
  // (few workgroups, so few atomics land on each vector's accumulator)
  const uint64_t rows_total = (uint64_t)P.row_blocks_run;
  const uint64_t r0 = (uint64_t)blockIdx.x * kFoldRows;
  for (uint32_t tl = 0; tl < tiles; ++tl) {
    const uint32_t vl = threadIdx.x % T;
    const uint32_t slot = threadIdx.x / T;  // kThreads/T slots stride over the kFoldRows rows
    constexpr uint32_t kS = kThreads / T, kPer = kFoldRows / kS;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      const uint64_t r = r0 + slot + (uint64_t)q * kS;
      if (r < rows_total) a[q & 3] += P.blk_err[((uint64_t)tl * P.row_blocks + r) * T + vl];
    }
    red[threadIdx.x] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (threadIdx.x < (uint32_t)T) {
      double sum = 0.0;
      for (uint32_t q = 0; q < kThreads / T; ++q) sum += red[q * T + threadIdx.x];
      const uint32_t jj = tl * T + threadIdx.x;
      if (jj < P.k && sum > 0.0 && !vec_done(P, jj))
        atomicAdd(P.acc + (uint64_t)P.iter * P.k + jj, err_to_fixed(sum));
    }
    __syncthreads();
  }
}

// c0 = init / outdeg (init itself for a row without out-edges) for every vector j < k, in
// the tile-major layout: blockIdx.y is the k-tile, and x runs over its n x T floats in
// address order, so (row, lane) come from shifts and the stores are contiguous.
__global__ __launch_bounds__(kThreads) void ppr_init_kernel(uint32_t n, uint32_t k, uint32_t kt_shift,
                                                            const uint32_t* __restrict__ deg, float init,
                                                            float* __restrict__ c0) {
  const uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x;  // offset inside the tile
  const uint32_t i = (uint32_t)(q >> kt_shift);
  const uint32_t j = (blockIdx.y << kt_shift) + (uint32_t)(q & ((1u << kt_shift) - 1));
  if (i < n && j < k) {
    const uint32_t d = deg[i];
    c0[((uint64_t)blockIdx.y * n << kt_shift) + q] = d ? init / (float)d : init;
  }
}

__global__ __launch_bounds__(kThreads) void ppr_zero_acc_kernel(unsigned long long* __restrict__ acc, uint64_t m) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < m) acc[t] = 0ull;
}

// out[i] = deg[i] | kSeedFlag when row i appears in the sorted seed rows (binary search,
// so no two threads ever write the same entry)
__global__ __launch_bounds__(kThreads) void ppr_degflag_kernel(uint32_t n, const uint32_t* __restrict__ deg,
                                                               const uint32_t* __restrict__ seed_row, uint32_t ns,
                                                               uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = ns;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (seed_row[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  out[i] = deg[i] | ((lo < ns && seed_row[lo] == i) ? kSeedFlag : 0u);
}

__global__ __launch_bounds__(kThreads) void gather_u32_kernel(uint32_t m, const uint32_t* __restrict__ idx,
                                                              const uint32_t* __restrict__ table,
                                                              uint32_t* __restrict__ out) {
  const uint32_t q = blockIdx.x * kThreads + threadIdx.x;
  if (q < m) out[q] = table[idx[q]];
}

// sweeps vector j executed: the first t with acc[t][j] below the threshold, plus one (or all)
__global__ __launch_bounds__(kThreads) void ppr_its_kernel(uint32_t k, const unsigned long long* __restrict__ acc,
                                                           uint64_t iters, double threshold,
                                                           uint32_t* __restrict__ its) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= k) return;
  uint64_t n_it = iters;
  for (uint64_t q = 0; q < iters; ++q)
    if (fixed_to_err(acc[q * k + j]) < threshold) {
      n_it = q + 1;
      break;
    }
  its[j] = (uint32_t)n_it;
}

// Final scores from the contributions: vector j's last sweep t = its[j] - 1 wrote c0 when t
// is odd, c1 when even (c0 before any sweep).  With a relabelled solve row r is original
// vertex order[r].
__global__ __launch_bounds__(kThreads) void ppr_scores_kernel(uint32_t n, uint32_t k, uint32_t kt_shift,
                                                              const uint32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ degflag,
                                                              const float* __restrict__ c0,
                                                              const float* __restrict__ c1,
                                                              const uint32_t* __restrict__ its,
                                                              float* __restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const uint64_t nk = (uint64_t)n * k;
  if (t >= nk) return;
  uint32_t r, j;
  if (nk <= 0xFFFFFFFFull) {  // uniform over the launch: a 32-bit division is far cheaper
    const uint32_t t32 = (uint32_t)t;
    r = t32 / k;
    j = t32 - r * k;
  } else {
    r = (uint32_t)(t / k);
    j = (uint32_t)(t % k);
  }
  const uint32_t nj = its[j];
  const float* c = nj == 0 ? c0 : ((nj - 1) & 1) ? c0 : c1;
  const float v = c[(((uint64_t)(j >> kt_shift) * n + r) << kt_shift) + (j & ((1u << kt_shift) - 1))];
  out[(uint64_t)(order ? order[r] : r) * k + j] = ppr_score(v, degflag[r]);
}

inline uint32_t grid1(uint64_t m) { return (uint32_t)ceil_div(m ? m : 1, kThreads); }

__global__ __launch_bounds__(kThreads) void heavy_deg_kernel(uint32_t nh, const uint32_t* __restrict__ rows,
                                                             const uint32_t* __restrict__ rp,
                                                             uint32_t* __restrict__ deg) {
  const uint32_t h = blockIdx.x * kThreads + threadIdx.x;
  if (h < nh) deg[h] = rp[rows[h] + 1] - rp[rows[h]];
}

// heavy edge e (row h's edges at hoff[h] ..): key (h, slice), payload the source.  A row
// of at least big edges is split by source slice; a shorter one keeps its edges together
// in slice h mod nsl (a slice of its own would leave its segments mostly empty).
__global__ __launch_bounds__(kThreads) void heavy_key_kernel(uint64_t eh, uint32_t nh,
                                                             const uint32_t* __restrict__ hoff,
                                                             const uint32_t* __restrict__ rows,
                                                             const uint32_t* __restrict__ rp,
                                                             const uint32_t* __restrict__ col, uint32_t sh,
                                                             uint32_t nsl, uint32_t big,
                                                             uint64_t* __restrict__ keys,
                                                             uint32_t* __restrict__ vals) {
  const uint64_t e = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= eh) return;
  uint32_t lo = 0, hi = nh - 1;  // last h with hoff[h] <= e
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (hoff[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t u = col[rp[rows[lo]] + (uint32_t)(e - hoff[lo])];
  const uint32_t x = hoff[lo + 1] - hoff[lo] >= big ? (u >> sh) % nsl : lo % nsl;
  keys[e] = ((uint64_t)lo << 3) | x;
  vals[e] = u;
}

// gp[t] = first sorted position with key >= t, t in [0, ng]
__global__ __launch_bounds__(kThreads) void group_ptr_kernel(uint64_t eh, const uint64_t* __restrict__ keys,
                                                             uint64_t ng, uint32_t* __restrict__ gp) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t > ng) return;
  uint64_t lo = 0, hi = eh;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (keys[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  gp[t] = (uint32_t)lo;
}

// light CSC: row i keeps its in-edges when in-degree <= light_max, else none
__global__ __launch_bounds__(kThreads) void light_len_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                             uint32_t light_max, uint32_t* __restrict__ len) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    const uint32_t d = rp[i + 1] - rp[i];
    len[i] = d <= light_max ? d : 0u;
  }
}

// *end = one past the last row with in-edges
__global__ __launch_bounds__(kThreads) void row_end_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                           uint32_t* __restrict__ end) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n && rp[i + 1] > rp[i]) atomicMax(end, i + 1);
}

__global__ __launch_bounds__(kThreads) void light_copy_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                              const uint32_t* __restrict__ col,
                                                              const uint32_t* __restrict__ lrp,
                                                              uint32_t* __restrict__ lcol) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint32_t d = lrp[i + 1] - lrp[i];
  for (uint32_t t = 0; t < d; ++t) lcol[lrp[i] + t] = col[rp[i] + t];
}

}  // namespace

// Heavy rows (in-degree >= min_deg): their in-edges regrouped by (row, source slice), the
// slice of source u being (u >> sh) mod nsl with sh chosen so that the T-float
// contributions of one 128-byte line fall in one slice (rows of fewer than nsl * seg
// edges stay whole, in slice h mod nsl); then cut into segments of at most seg edges of
// one (row, slice), numbered row by row (so a row's partials are added in a fixed order),
// and listed per slice.
static void build_heavy_slices(hipStream_t s, const ActiveCsc& a, uint32_t n, uint32_t min_deg, uint32_t seg,
                               uint32_t tile, uint32_t nsl, PprState& st) {
  DevBuf<uint32_t> unused_ptr, unused_row;
  uint32_t nh = 0, ns0 = 0;
  build_segments(s, a.ptr, n, min_deg, seg, st.heavy_rows, unused_ptr, unused_row, &nh, &ns0);
  const uint32_t sh = tile >= 32 ? 0u : tile == 16 ? 1u : 2u;
  std::vector<uint32_t> gp_h((uint64_t)nh * 8 + 1, 0);
  uint64_t eh = 0;
  if (nh) {
    DevBuf<uint32_t> deg(nh), hoff((uint64_t)nh + 1);
    heavy_deg_kernel<<<grid1(nh), kThreads, 0, s>>>(nh, st.heavy_rows.p, a.ptr, deg.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, deg.p, hoff.p, nh);
    uint32_t eh32 = 0;
    NWG_HIP(hipMemcpyAsync(&eh32, hoff.p + nh, 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    eh = eh32;
    st.hcol.alloc(std::max<uint64_t>(eh, 1));
    if (eh) {
      DevBuf<uint64_t> keys(eh), keys_tmp(eh);
      DevBuf<uint32_t> vals_tmp(eh);
      heavy_key_kernel<<<(uint32_t)ceil_div(eh, kThreads), kThreads, 0, s>>>(eh, nh, hoff.p, st.heavy_rows.p, a.ptr,
                                                                             a.idx, sh, nsl, nsl * seg, keys.p,
                                                                             st.hcol.p);
      NWG_LAUNCH_CHECK();
      int kb = 3;
      while (kb < 64 && (1ull << kb) < (uint64_t)nh * 8) ++kb;
      radix_sort_pairs(s, keys.p, keys_tmp.p, st.hcol.p, vals_tmp.p, eh, kb);  // stable: sources stay ascending
      DevBuf<uint32_t> gp((uint64_t)nh * 8 + 1);
      group_ptr_kernel<<<grid1((uint64_t)nh * 8 + 1), kThreads, 0, s>>>(eh, keys.p, (uint64_t)nh * 8, gp.p);
      NWG_LAUNCH_CHECK();
      NWG_HIP(hipMemcpyAsync(gp_h.data(), gp.p, gp_h.size() * 4, hipMemcpyDeviceToHost, s));
      NWG_HIP(hipStreamSynchronize(s));
    }
  } else {
    st.hcol.alloc(1);
  }
  std::vector<uint32_t> sptr((uint64_t)nh + 1), sstart, send, srow;
  std::vector<std::vector<uint32_t>> lists(nsl);
  for (uint32_t h = 0; h < nh; ++h) {
    sptr[h] = (uint32_t)sstart.size();
    for (uint32_t x = 0; x < nsl; ++x) {
      const uint32_t g0 = gp_h[(uint64_t)h * 8 + x], g1 = gp_h[(uint64_t)h * 8 + x + 1];
      for (uint32_t p = g0; p < g1; p += seg) {
        lists[x].push_back((uint32_t)sstart.size());
        sstart.push_back(p);
        send.push_back(std::min(p + seg, g1));
        srow.push_back(h);
      }
    }
  }
  sptr[nh] = (uint32_t)sstart.size();
  std::vector<uint32_t> flat;
  st.slice_off.assign(nsl + 1, 0);
  for (uint32_t x = 0; x < nsl; ++x) {
    st.slice_off[x] = (uint32_t)flat.size();
    flat.insert(flat.end(), lists[x].begin(), lists[x].end());
  }
  st.slice_off[nsl] = (uint32_t)flat.size();
  const uint32_t ns = (uint32_t)sstart.size();
  st.heavy_seg_ptr.alloc((uint64_t)nh + 1);
  st.heavy_seg_row.alloc(std::max<uint32_t>(ns, 1));
  st.seg_start.alloc(std::max<uint32_t>(ns, 1));
  st.seg_end.alloc(std::max<uint32_t>(ns, 1));
  st.slice_segs.alloc(std::max<uint32_t>(ns, 1));
  NWG_HIP(hipMemcpyAsync(st.heavy_seg_ptr.p, sptr.data(), sptr.size() * 4, hipMemcpyHostToDevice, s));
  if (ns) {
    NWG_HIP(hipMemcpyAsync(st.heavy_seg_row.p, srow.data(), ns * 4ull, hipMemcpyHostToDevice, s));
    NWG_HIP(hipMemcpyAsync(st.seg_start.p, sstart.data(), ns * 4ull, hipMemcpyHostToDevice, s));
    NWG_HIP(hipMemcpyAsync(st.seg_end.p, send.data(), ns * 4ull, hipMemcpyHostToDevice, s));
    NWG_HIP(hipMemcpyAsync(st.slice_segs.p, flat.data(), ns * 4ull, hipMemcpyHostToDevice, s));
  }
  NWG_HIP(hipStreamSynchronize(s));
  st.n_heavy = nh;
  st.n_segs = ns;
}

// Host driver; tile = k-tile width (8, 16, 32 or 64).
void ppr_run(nwg_ctx* ctx, nwg_graph* g, uint32_t k, const std::vector<uint32_t>& seed_ptr,
             const std::vector<uint32_t>& seeds, float damping, float threshold, uint64_t max_iters, uint32_t tile,
             float* scores_dev, uint64_t* iterations, nwg_pr_info* info) {
  hipStream_t s = ctx->stream;
  const uint32_t n = g->n;
  NWG_REQUIRE(k >= 1, "k must be >= 1");
  NWG_REQUIRE(tile == 8 || tile == 16 || tile == 32 || tile == 64, "k-tile must be 8, 16, 32 or 64");
  NWG_REQUIRE(max_iters < (1ull << 31), "max_iters too large");
  if (!g->plan.built) plan_build(s, g, default_plan_params());
  const Plan& pl = g->plan;
  const ActiveCsc a = g->active(pl.params.relabel != 0);
  if (!g->ppr) g->ppr.reset(new PprState());
  PprState& st = *g->ppr;
  // Heavy rows (in-degree > light_max) are split into segments of seg_edges edges so
  // that no lane group walks more than 32 edges of a row; rebuilt only when the plan
  // (hence the active CSC) or the k-tile changes.  Rows up to light_max in-edges stay on
  // the row path, which writes no partials: measured on MI355X (k=64, RMAT-22), a cut of
  // 256 solves in 51.0 ms against 53.3 ms with 64, and 1024 or 2048 are no faster.
  const uint32_t light_max = 256;
  const uint32_t seg_edges = (kThreads / tile) * kSegPerGroup;
  // source slices per k-tile: the 8 XCDs are shared out as tiles x slices
  const uint32_t tiles_k = (uint32_t)ceil_div(k, tile);
  const uint32_t nsl = (8 % tiles_k == 0) ? 8 / tiles_k : 1;
  if (st.seg_version != g->plan_version || st.seg_tile != tile || st.seg_slices != nsl) {
    st.seg_version = 0;
    build_heavy_slices(s, a, n, light_max + 1, seg_edges, tile, nsl, st);
    DevBuf<uint32_t> len(n);
    st.lrp.alloc((uint64_t)n + 1);
    light_len_kernel<<<grid1(n), kThreads, 0, s>>>(n, a.ptr, light_max, len.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, len.p, st.lrp.p, n);
    uint32_t n_light = 0;
    NWG_HIP(hipMemcpyAsync(&n_light, st.lrp.p + n, 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    st.lcol.alloc(std::max<uint32_t>(n_light, 1));
    light_copy_kernel<<<grid1(n), kThreads, 0, s>>>(n, a.ptr, a.idx, st.lrp.p, st.lcol.p);
    NWG_LAUNCH_CHECK();
    DevBuf<uint32_t> end(1);
    NWG_HIP(hipMemsetAsync(end.p, 0, 4, s));
    row_end_kernel<<<grid1(n), kThreads, 0, s>>>(n, a.ptr, end.p);
    NWG_LAUNCH_CHECK();
    NWG_HIP(hipMemcpyAsync(&st.row_end, end.p, 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    st.seg_version = g->plan_version;
    st.seg_tile = tile;
    st.seg_slices = nsl;
  }
  const uint32_t n_heavy = st.n_heavy, n_segs = st.n_segs;
  // seed entries sorted by (row, vector); rows in the active numbering
  std::vector<std::pair<uint32_t, uint32_t>> ent;
  std::vector<float> ebase;
  for (uint32_t j = 0; j < k; ++j) {
    const uint32_t c = seed_ptr[j + 1] - seed_ptr[j];
    NWG_REQUIRE(c >= 1, "every teleport vector needs at least one seed");
    for (uint32_t q = seed_ptr[j]; q < seed_ptr[j + 1]; ++q) {
      NWG_REQUIRE(seeds[q] < n, "seed id >= n");
      ent.emplace_back(seeds[q], j);
    }
  }
  std::sort(ent.begin(), ent.end());
  for (size_t q = 1; q < ent.size(); ++q) NWG_REQUIRE(ent[q] != ent[q - 1], "duplicate seed in a teleport vector");
  const uint32_t ns = (uint32_t)ent.size();
  if (a.order) {  // map seed ids through rank (original -> relabelled) and re-sort
    std::vector<uint32_t> ids(ns);
    for (uint32_t q = 0; q < ns; ++q) ids[q] = ent[q].first;
    st.ids.ensure(ns);
    st.ids_new.ensure(ns);
    NWG_HIP(hipMemcpyAsync(st.ids.p, ids.data(), ns * 4ull, hipMemcpyHostToDevice, s));
    gather_u32_kernel<<<grid1(ns), kThreads, 0, s>>>(ns, st.ids.p, g->rl->rank.p, st.ids_new.p);
    NWG_LAUNCH_CHECK();
    NWG_HIP(hipMemcpyAsync(ids.data(), st.ids_new.p, ns * 4ull, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    for (uint32_t q = 0; q < ns; ++q) ent[q].first = ids[q];
    std::sort(ent.begin(), ent.end());
  }
  std::vector<uint32_t> erow(ns), evec(ns);
  ebase.resize(ns);
  for (size_t q = 0; q < ns; ++q) {
    erow[q] = ent[q].first;
    evec[q] = ent[q].second;
    const uint32_t c = seed_ptr[ent[q].second + 1] - seed_ptr[ent[q].second];
    const float v = 1.0f / (float)c;                                  // teleport value v_j[row]
    ebase[q] = (float)((1.0 - (double)damping) * (double)v);          // same rounding as the plain path
  }
  st.seed_row.ensure(ns);
  st.seed_vec.ensure(ns);
  st.seed_base.ensure(ns);
  st.degflag.ensure(n);
  NWG_HIP(hipMemcpyAsync(st.seed_row.p, erow.data(), ns * 4ull, hipMemcpyHostToDevice, s));
  NWG_HIP(hipMemcpyAsync(st.seed_vec.p, evec.data(), ns * 4ull, hipMemcpyHostToDevice, s));
  NWG_HIP(hipMemcpyAsync(st.seed_base.p, ebase.data(), ns * 4ull, hipMemcpyHostToDevice, s));
  ppr_degflag_kernel<<<grid1(n), kThreads, 0, s>>>(n, a.deg, st.seed_row.p, ns, st.degflag.p);
  NWG_LAUNCH_CHECK();

  const uint32_t tiles = (uint32_t)ceil_div(k, tile);
  const uint32_t rows_per_group = tile >= 16 ? 8u : 4u;  // PprRows<tile>
  const uint32_t row_blocks = (uint32_t)ceil_div(n, (kThreads / tile) * rows_per_group);
  const uint64_t nk = (uint64_t)n * k;
  const uint64_t nc = (uint64_t)tiles * n * tile;  // tile-major contributions, last tile padded
  st.c0.ensure(nc);
  st.c1.ensure(nc);
  st.hub_part.ensure((uint64_t)std::max<uint32_t>(n_segs, 1) * k);
  st.blk_err.ensure((uint64_t)row_blocks * tiles * tile);
  st.acc.ensure(std::max<uint64_t>(max_iters, 1) * k);
  float* const c0p = st.c0.p;
  float* const c1p = st.c1.p;

  PprParams P{};
  P.rp = a.ptr;
  P.col = a.idx;
  P.lrp = st.lrp.p;
  P.lcol = st.lcol.p;
  P.degflag = st.degflag.p;
  P.seed_row = st.seed_row.p;
  P.seed_vec = st.seed_vec.p;
  P.seed_base = st.seed_base.p;
  P.n_seed = ns;
  P.n = n;
  P.k = k;
  P.tiles = tiles;
  P.kt_shift = (uint32_t)__builtin_ctz(tile);
  P.damp = damping;
  P.light_max = light_max;
  P.hub_segment = seg_edges;
  P.row_blocks = row_blocks;
  P.n_hub_rows = n_heavy;
  P.n_hub_segs = n_segs;
  P.hub_rows = st.heavy_rows.p;
  P.seg_ptr = st.heavy_seg_ptr.p;
  P.seg_row = st.heavy_seg_row.p;
  P.hub_part = st.hub_part.p;
  P.blk_err = st.blk_err.p;
  P.acc = st.acc.p;
  P.threshold = (double)threshold;
  P.hcol = st.hcol.p;
  P.seg_start = st.seg_start.p;
  P.seg_end = st.seg_end.p;
  P.slice_segs = st.slice_segs.p;
  P.n_slices = nsl;
  uint32_t max_slice = 0;
  for (uint32_t x = 0; x <= nsl; ++x) P.slice_off[x] = st.slice_off[x];
  for (uint32_t x = 0; x < nsl; ++x) max_slice = std::max(max_slice, st.slice_off[x + 1] - st.slice_off[x]);
  P.heavy_blocks = 8u * max_slice;
  // Rows without in-edges reach their fixed point (1-d) v_j in one sweep and hold it, and the
  // next sweep writes the same contribution into the other buffer: from the third sweep on
  // only the row workgroups up to the last row with in-edges run (degree ordering puts
  // every row without in-edges last, half of RMAT-22's rows).
  const uint32_t rows_per_block = (kThreads / tile) * rows_per_group;
  const uint32_t live_row_blocks = (uint32_t)std::min<uint64_t>(ceil_div(st.row_end, rows_per_block), row_blocks);

  const float init = (float)(1.0 / n);
  NWG_HIP(hipEventRecord(ctx->ev0, s));
  {
    const uint64_t per_tile = (uint64_t)n << P.kt_shift;
    NWG_REQUIRE(ceil_div(per_tile, kThreads) < (1ull << 31) && P.tiles < 65536u, "batched solve too large for one launch");
    ppr_init_kernel<<<dim3((uint32_t)ceil_div(per_tile, kThreads), P.tiles), kThreads, 0, s>>>(n, k, P.kt_shift,
                                                                                              a.deg, init, c0p);
  }
  NWG_LAUNCH_CHECK();
  ppr_zero_acc_kernel<<<grid1(max_iters * k), kThreads, 0, s>>>(st.acc.p, max_iters * k);
  NWG_LAUNCH_CHECK();
  const uint64_t grid64 = (uint64_t)P.heavy_blocks + (uint64_t)row_blocks * tiles;
  NWG_REQUIRE(grid64 < (1ull << 31), "batched solve too large for one launch");
  const uint64_t fin_a = ceil_div((uint64_t)n_heavy * k, kFinItems), fin_b = ceil_div(row_blocks, kFoldRows);
  const uint32_t fin = (uint32_t)std::max<uint64_t>(std::max(fin_a, fin_b), 1);
  for (uint64_t it = 0; it < max_iters; ++it) {
    P.iter = (uint32_t)it;
    P.cin = (it & 1) ? c1p : c0p;
    P.cout = (it & 1) ? c0p : c1p;
    P.row_blocks_run = it >= 2 ? live_row_blocks : row_blocks;
    const uint32_t grid = P.heavy_blocks + P.row_blocks_run * tiles;  // 0 only for a graph without edges
    switch (tile) {
      case 8:
        if (grid) ppr_sweep_kernel<8><<<grid, kThreads, 0, s>>>(P);
        ppr_finish_kernel<8><<<fin, kThreads, 0, s>>>(P, tiles);
        break;
      case 16:
        if (grid) ppr_sweep_kernel<16><<<grid, kThreads, 0, s>>>(P);
        ppr_finish_kernel<16><<<fin, kThreads, 0, s>>>(P, tiles);
        break;
      case 32:
        if (grid) ppr_sweep_kernel<32><<<grid, kThreads, 0, s>>>(P);
        ppr_finish_kernel<32><<<fin, kThreads, 0, s>>>(P, tiles);
        break;
      default:
        if (grid) ppr_sweep_kernel<64><<<grid, kThreads, 0, s>>>(P);
        ppr_finish_kernel<64><<<fin, kThreads, 0, s>>>(P, tiles);
        break;
    }
    NWG_LAUNCH_CHECK();
  }
  NWG_HIP(hipEventRecord(ctx->ev1, s));
  st.its.ensure(k);
  ppr_its_kernel<<<grid1(k), kThreads, 0, s>>>(k, st.acc.p, max_iters, (double)threshold, st.its.p);
  ppr_scores_kernel<<<grid1(nk), kThreads, 0, s>>>(n, k, P.kt_shift, a.order, st.degflag.p, c0p, c1p, st.its.p,
                                                   scores_dev);
  NWG_LAUNCH_CHECK();
  std::vector<unsigned long long> h_acc(std::max<uint64_t>(max_iters, 1) * k);
  NWG_HIP(hipMemcpyAsync(h_acc.data(), st.acc.p, max_iters * k * 8ull, hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  NWG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  uint64_t max_it = 0;
  for (uint32_t j = 0; j < k; ++j) {
    uint64_t itj = max_iters;
    for (uint64_t t = 0; t < max_iters; ++t)
      if (fixed_to_err(h_acc[t * k + j]) < (double)threshold) {
        itj = t + 1;
        break;
      }
    if (iterations) iterations[j] = itj;
    max_it = std::max(max_it, itj);
  }
  if (info) {
    info->iterations = max_it;
    info->final_error = 0.0;
    for (uint32_t j = 0; j < k && max_it; ++j) info->final_error += fixed_to_err(h_acc[(max_it - 1) * k + j]);
    info->ms_total = ms;
    info->ms_per_iter = max_it ? ms / (double)max_it : 0.0;
    info->sweep_ms_avg = 0.0;
  }
}

}  // namespace nwg
