// relabel.hip -- degree-ordered relabelling of a CSC (locality preprocessing).
//
// The pull sweep gathers contrib[u] once per out-edge of u, so a vertex's gather
// frequency is its out-degree.  Renumbering vertices by out-degree (descending, ties by
// original id; vertices without in-edges after all others) packs the most gathered contributions into the fewest cache lines: on an
// RMAT graph the hottest million vertices then occupy one contiguous 4 MB range that the
// L2 of each XCD keeps, instead of being spread over the whole 134 MB array at scale 25.
// Rows are renumbered with the same permutation, so scores and contributions stay
// contiguous; the solve permutes its result back to the original ids.  The relabelled
// CSC keeps rows sorted by new source id.  (The reference's counterpart is
// relabel_by_degree / permute_by_degree, build.hpp:580-615, compressed.hpp:523-598,
// which the PageRank driver does not call; results are invariant under relabelling up
// to FP32 summation order.)
#include <hip/hip_runtime.h>

#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kThreads = 256;

inline uint32_t grid1(uint64_t m) { return (uint32_t)ceil_div(m ? m : 1, kThreads); }

// key = (no in-edges << (32 + b)) | ((2^32-1 - outdeg) << b) | id : ascending order = the
// vertices with in-edges first, each group by out-degree descending, id ascending.  (A vertex
// without in-edges reaches its fixed point in one sweep; grouping them last leaves the
// binned sweep whole bins with nothing to add, which it stops visiting.)
__global__ __launch_bounds__(kThreads) void degree_key_kernel(uint32_t n, const uint32_t* __restrict__ deg,
                                                              const uint32_t* __restrict__ ptr, int b,
                                                              uint64_t* __restrict__ keys) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    const uint64_t no_in = ptr[i + 1] == ptr[i] ? 1ull : 0ull;
    keys[i] = (no_in << (32 + b)) | ((uint64_t)(0xFFFFFFFFu - deg[i]) << b) | i;
  }
}

__global__ __launch_bounds__(kThreads) void order_kernel(uint32_t n, const uint64_t* __restrict__ keys, int b,
                                                         const uint32_t* __restrict__ deg,
                                                         uint32_t* __restrict__ order, uint32_t* __restrict__ rank,
                                                         uint32_t* __restrict__ deg_new) {
  const uint32_t r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= n) return;
  const uint32_t v = (uint32_t)(keys[r] & ((1ull << b) - 1));
  order[r] = v;
  rank[v] = r;
  deg_new[r] = deg[v];
}

// relabelled edge keys (rank[row] << b) | rank[src], one thread per CSC position (the
// row found by binary search, so hub rows do not serialise on one lane)
__global__ __launch_bounds__(kThreads) void edge_key_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ ptr,
                                                            const uint32_t* __restrict__ col,
                                                            const uint32_t* __restrict__ rank, int b,
                                                            uint64_t* __restrict__ keys) {
  const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= nnz) return;
  uint32_t lo = 0, hi = n;  // first row r with ptr[r + 1] > p
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ptr[mid + 1] <= p) lo = mid + 1;
    else hi = mid;
  }
  keys[p] = ((uint64_t)rank[lo] << b) | rank[col[p]];
}

}  // namespace

void relabel_build(hipStream_t s, nwg_graph* g) {
  if (g->rl) return;
  const uint32_t n = g->n;
  const uint64_t nnz = g->nnz;
  const int b = key_bits_for(n);
  NWG_REQUIRE(b <= 31, "degree ordering needs at most 2^31 vertices");
  std::unique_ptr<Relabeled> rl(new Relabeled());
  rl->order.alloc(n);
  rl->rank.alloc(n);
  rl->outdeg.alloc(n);
  {
    DevBuf<uint64_t> keys(n), tmp(n);
    degree_key_kernel<<<grid1(n), kThreads, 0, s>>>(n, g->outdeg.p, g->csc_ptr.p, b, keys.p);
    NWG_LAUNCH_CHECK();
    radix_sort_pairs(s, keys.p, tmp.p, nullptr, nullptr, n, 33 + b);
    order_kernel<<<grid1(n), kThreads, 0, s>>>(n, keys.p, b, g->outdeg.p, rl->order.p, rl->rank.p, rl->outdeg.p);
    NWG_LAUNCH_CHECK();
    NWG_HIP(hipStreamSynchronize(s));
  }
  rl->csc_ptr.alloc((uint64_t)n + 1);
  rl->csc_idx.alloc(nnz ? nnz : 1);
  {
    DevBuf<uint64_t> keys(nnz ? nnz : 1), tmp(nnz ? nnz : 1);
    if (nnz) {
      edge_key_kernel<<<grid1(nnz), kThreads, 0, s>>>(n, nnz, g->csc_ptr.p, g->csc_idx.p, rl->rank.p, b, keys.p);
      NWG_LAUNCH_CHECK();
      radix_sort_pairs(s, keys.p, tmp.p, nullptr, nullptr, nnz, 2 * b);
    }
    sorted_keys_to_csx(s, keys.p, nnz, n, b, rl->csc_ptr.p, rl->csc_idx.p);
    NWG_HIP(hipStreamSynchronize(s));
  }
  g->rl = std::move(rl);
}

}  // namespace nwg
