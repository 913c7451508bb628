// multi.hip -- many independent graphs solved by one persistent launch per sweep.
//
// The graphs are concatenated into one block-diagonal CSC (vertex ids offset per
// graph).  The union is cut into work items (graph, rows [r0, r1)): up to 256
// consecutive rows of one graph whose in-edges total at most kItemNnz, or one row
// with more in-edges (a hub item).  A persistent grid walks the items round-robin;
// each sweep:
//   normal item  the workgroup gathers contrib[col[p]] for the whole contiguous edge
//                range into LDS (coalesced id stream, every lane busy), then one lane
//                per row adds its row's values from LDS in row order
//   hub item     the whole workgroup strides over the row and reduces
// Each graph keeps the reference's scalars (init = 1/N_g, base = (1-d)/N_g,
// page_rank.hpp:92-94) and its own stopping rule: the L1 change of graph g in sweep k
// is accumulated in fixed point in acc[k][g]; items of a converged graph are skipped.
// The init launch and max_iters sweep launches are captured in one HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kItemNnz = 4096;
constexpr uint32_t kItemRows = 256;

struct MultiParams {
  const uint32_t* rp;
  const uint32_t* col;
  const uint32_t* deg;
  const uint32_t* item_graph;
  const uint32_t* item_r0;
  const uint32_t* item_r1;
  uint32_t n_items;
  const float* gbase;  // per graph (1-d)/N_g
  uint32_t ngraphs;
  const float* cin;
  float* cout;
  float* x;
  float damp;
  unsigned long long* acc;  // iters x ngraphs
  uint32_t iter;
  double threshold;
};

__device__ __forceinline__ bool graph_done(const MultiParams& P, uint32_t g) {
  return P.iter > 0 && fixed_to_err(P.acc[(uint64_t)(P.iter - 1) * P.ngraphs + g]) < P.threshold;
}

__device__ __forceinline__ double multi_update(const MultiParams& P, uint32_t i, float base, float z) {
  const float xo = P.x[i];
  const float xn = __builtin_fmaf(P.damp, z, base);
  P.x[i] = xn;
  P.cout[i] = xn / (float)P.deg[i];
  return (double)fabsf(xn - xo);
}

__global__ __launch_bounds__(kThreads) void multi_sweep_kernel(MultiParams P) {
  __shared__ float vals[kItemNnz];
  __shared__ double red_d[kThreads / 64];
  __shared__ float red_f[kThreads / 64];
  for (uint32_t w = blockIdx.x; w < P.n_items; w += gridDim.x) {
    const uint32_t g = P.item_graph[w];
    if (graph_done(P, g)) continue;
    const uint32_t r0 = P.item_r0[w], r1 = P.item_r1[w];
    const uint32_t p0 = P.rp[r0], p1 = P.rp[r1];
    const float base = P.gbase[g];
    double e = 0.0;
    if (p1 - p0 > kItemNnz) {
      // hub item: a single row
      float z = 0.f;
      for (uint32_t p = p0 + threadIdx.x; p < p1; p += kThreads) z += P.cin[P.col[p]];
      z = block_sum<kThreads>(z, red_f);
      if (threadIdx.x == 0) e = multi_update(P, r0, base, z);
    } else {
      for (uint32_t q = threadIdx.x; q < p1 - p0; q += kThreads) vals[q] = P.cin[P.col[p0 + q]];
      __syncthreads();
      const uint32_t i = r0 + threadIdx.x;
      if (i < r1) {
        const uint32_t a = P.rp[i] - p0, b = P.rp[i + 1] - p0;
        float z = 0.f;
        for (uint32_t q = a; q < b; ++q) z += vals[q];
        e = multi_update(P, i, base, z);
      }
    }
    e = block_sum<kThreads>(e, red_d);
    if (threadIdx.x == 0 && e > 0.0) atomicAdd(P.acc + (uint64_t)P.iter * P.ngraphs + g, err_to_fixed(e));
    __syncthreads();  // vals and the reduction scratch are reused by the next item
  }
}

__global__ __launch_bounds__(kThreads) void multi_init_kernel(uint64_t nv, const uint32_t* __restrict__ deg,
                                                              const float* __restrict__ vinit, float* __restrict__ x,
                                                              float* __restrict__ c0,
                                                              unsigned long long* __restrict__ acc, uint64_t nacc) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < nv) {
    const float init = vinit[t];
    x[t] = init;
    c0[t] = init / (float)deg[t];
  }
  if (t < nacc) acc[t] = 0ull;
}

// offset a graph's CSC into the union
__global__ __launch_bounds__(kThreads) void multi_copy_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ rp,
                                                              const uint32_t* __restrict__ col,
                                                              const uint32_t* __restrict__ deg, uint32_t voff,
                                                              uint32_t eoff, float init, uint32_t* __restrict__ rp_all,
                                                              uint32_t* __restrict__ col_all,
                                                              uint32_t* __restrict__ deg_all,
                                                              float* __restrict__ vinit) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < n) {
    rp_all[voff + t] = eoff + rp[t];
    deg_all[voff + t] = deg[t];
    vinit[voff + t] = init;
  }
  if (t < nnz) col_all[eoff + t] = col[t] + voff;
}

// graph of every row of the union
__global__ __launch_bounds__(kThreads) void row_graph_kernel(uint32_t n, uint32_t voff, uint32_t g,
                                                             uint32_t* __restrict__ row_graph) {
  const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
  if (t < n) row_graph[voff + t] = g;
}

// per-row base (1-d)/N_g of the binned sweep
__global__ __launch_bounds__(kThreads) void row_base_kernel(uint64_t nv, const uint32_t* __restrict__ row_graph,
                                                            const float* __restrict__ gbase, float* __restrict__ base) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < nv) base[t] = gbase[row_graph[t]];
}

inline uint32_t grid1(uint64_t m) { return (uint32_t)ceil_div(m ? m : 1, kThreads); }

}  // namespace

}  // namespace nwg

struct nwg_multigraph {
  nwg_ctx* ctx = nullptr;
  int device = 0;
  uint32_t ngraphs = 0;
  uint64_t nv = 0, ne = 0;
  std::vector<uint64_t> voff;
  nwg::DevBuf<uint32_t> rp, col, deg, item_graph, item_r0, item_r1;
  nwg::DevBuf<float> vinit, x, c0, c1, gbase;
  nwg::DevBuf<unsigned long long> acc;
  uint32_t n_items = 0;
  nwg::Captured cap;
  // binned sweep of the union (plan option binned)
  bool binned = false;
  nwg::BinnedLayout bl;
  nwg::DevBuf<uint32_t> row_graph;
  nwg::DevBuf<float> row_base, bvals;
  nwg::DevBuf<unsigned long long> bpart;
};

namespace nwg {

nwg_multigraph* multigraph_create(nwg_ctx* ctx, uint32_t ngraphs, nwg_graph* const* graphs) {
  hipStream_t s = ctx->stream;
  NWG_REQUIRE(ngraphs >= 1, "need at least one graph");
  std::unique_ptr<nwg_multigraph> mg(new nwg_multigraph());
  mg->ctx = ctx;
  mg->device = ctx->device;
  mg->ngraphs = ngraphs;
  mg->voff.resize(ngraphs + 1, 0);
  std::vector<uint64_t> eoff(ngraphs + 1, 0);
  for (uint32_t g = 0; g < ngraphs; ++g) {
    NWG_REQUIRE(graphs[g] && graphs[g]->device == ctx->device, "graph missing or on another device");
    mg->voff[g + 1] = mg->voff[g] + graphs[g]->n;
    eoff[g + 1] = eoff[g] + graphs[g]->nnz;
  }
  mg->nv = mg->voff[ngraphs];
  mg->ne = eoff[ngraphs];
  NWG_REQUIRE(mg->nv < (1ull << 32) && mg->ne < (1ull << 32), "union exceeds uint32 ids");
  mg->rp.alloc(mg->nv + 1);
  mg->col.alloc(mg->ne ? mg->ne : 1);
  mg->deg.alloc(mg->nv);
  mg->vinit.alloc(mg->nv);
  for (uint32_t g = 0; g < ngraphs; ++g) {
    const nwg_graph* gr = graphs[g];
    multi_copy_kernel<<<grid1(std::max<uint64_t>(gr->n, gr->nnz)), kThreads, 0, s>>>(
        gr->n, gr->nnz, gr->csc_ptr.p, gr->csc_idx.p, gr->outdeg.p, (uint32_t)mg->voff[g], (uint32_t)eoff[g],
        (float)(1.0 / gr->n), mg->rp.p, mg->col.p, mg->deg.p, mg->vinit.p);
    NWG_LAUNCH_CHECK();
  }
  const uint32_t ne32 = (uint32_t)mg->ne;
  NWG_HIP(hipMemcpyAsync(mg->rp.p + mg->nv, &ne32, 4, hipMemcpyHostToDevice, s));
  // work items from the union's row pointers
  std::vector<uint32_t> rp(mg->nv + 1);
  NWG_HIP(hipMemcpyAsync(rp.data(), mg->rp.p, (mg->nv + 1) * 4, hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> ig, i0, i1;
  for (uint32_t g = 0; g < ngraphs; ++g) {
    uint64_t r = mg->voff[g];
    const uint64_t rend = mg->voff[g + 1];
    while (r < rend) {
      uint64_t r1 = r + 1;
      if (rp[r1] - rp[r] <= kItemNnz) {
        while (r1 < rend && r1 - r < kItemRows && rp[r1 + 1] - rp[r] <= kItemNnz) ++r1;
      }
      ig.push_back(g);
      i0.push_back((uint32_t)r);
      i1.push_back((uint32_t)r1);
      r = r1;
    }
  }
  mg->n_items = (uint32_t)ig.size();
  mg->item_graph.alloc(ig.size());
  mg->item_r0.alloc(ig.size());
  mg->item_r1.alloc(ig.size());
  NWG_HIP(hipMemcpyAsync(mg->item_graph.p, ig.data(), ig.size() * 4, hipMemcpyHostToDevice, s));
  NWG_HIP(hipMemcpyAsync(mg->item_r0.p, i0.data(), i0.size() * 4, hipMemcpyHostToDevice, s));
  NWG_HIP(hipMemcpyAsync(mg->item_r1.p, i1.data(), i1.size() * 4, hipMemcpyHostToDevice, s));
  mg->x.alloc(mg->nv);
  mg->c0.alloc(mg->nv);
  mg->c1.alloc(mg->nv);
  mg->gbase.alloc(ngraphs);
  NWG_HIP(hipStreamSynchronize(s));
  return mg.release();
}

// Selects the schedule of the union's sweeps: the items of multi_sweep_kernel, or the binned
// sweep (binned.hip) over a tiled copy of the union with a graph id per row.
void multigraph_set_binned(nwg_ctx* ctx, nwg_multigraph* mg, bool binned) {
  hipStream_t s = ctx->stream;
  if (binned && !mg->bl.built) {
    mg->row_graph.alloc(mg->nv ? mg->nv : 1);
    for (uint32_t g = 0; g < mg->ngraphs; ++g) {
      const uint32_t n = (uint32_t)(mg->voff[g + 1] - mg->voff[g]);
      if (n) row_graph_kernel<<<grid1(n), kThreads, 0, s>>>(n, (uint32_t)mg->voff[g], g, mg->row_graph.p);
    }
    NWG_LAUNCH_CHECK();
    // no hot tiles: the union's first ids are one graph's, not the union's hubs
    binned_build(s, ActiveCsc{mg->rp.p, mg->col.p, mg->deg.p, nullptr}, (uint32_t)mg->nv, mg->ne, mg->bl, false);
    mg->row_base.alloc(mg->nv ? mg->nv : 1);
    mg->bvals.alloc(mg->bl.pairs);
    mg->bpart.alloc(std::max<uint64_t>((uint64_t)mg->bl.n_slots << mg->bl.bin_bits, 1));
    NWG_HIP(hipStreamSynchronize(s));
  }
  if (mg->binned != binned) mg->cap.reset();
  mg->binned = binned;
}

void multigraph_run(nwg_ctx* ctx, nwg_multigraph* mg, float damping, float threshold, uint64_t max_iters,
                    float* scores_dev, uint64_t* iterations, double* final_error, nwg_pr_info* info) {
  hipStream_t s = ctx->stream;
  NWG_REQUIRE(max_iters < (1ull << 31), "max_iters too large");
  const uint32_t G = mg->ngraphs;
  std::vector<float> gb(G);
  for (uint32_t g = 0; g < G; ++g)
    gb[g] = (float)((1.0 - (double)damping) / (double)(mg->voff[g + 1] - mg->voff[g]));  // page_rank.hpp:94
  NWG_HIP(hipMemcpyAsync(mg->gbase.p, gb.data(), G * 4ull, hipMemcpyHostToDevice, s));
  mg->acc.ensure(std::max<uint64_t>(max_iters, 1) * G);

  int dev = 0, cus = 0;
  NWG_HIP(hipGetDevice(&dev));
  NWG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(mg->n_items, (uint64_t)cus * 8));

  MultiParams P{};
  P.rp = mg->rp.p;
  P.col = mg->col.p;
  P.deg = mg->deg.p;
  P.item_graph = mg->item_graph.p;
  P.item_r0 = mg->item_r0.p;
  P.item_r1 = mg->item_r1.p;
  P.n_items = mg->n_items;
  P.gbase = mg->gbase.p;
  P.ngraphs = G;
  P.x = mg->x.p;
  P.damp = damping;
  P.acc = mg->acc.p;
  P.threshold = (double)threshold;

  BinParams B{};
  if (mg->binned) {
    row_base_kernel<<<grid1(mg->nv), kThreads, 0, s>>>(mg->nv, mg->row_graph.p, mg->gbase.p, mg->row_base.p);
    NWG_LAUNCH_CHECK();
    const BinnedLayout& L = mg->bl;
    B.brow = L.brow.p;
    B.bsrc = L.bsrc.p;
    B.bpid = L.bpid.p;
    B.vals = mg->bvals.p;
    B.items2 = L.items2.p;
    B.items1 = L.items1.p;
    B.items3 = L.items3.p;
    B.seg_start = L.seg_start.p;
    B.seg_flat = L.seg_flat.p;
    B.part = mg->bpart.p;
    B.deg = mg->deg.p;
    B.x = mg->x.p;
    // one buffer: the chunk pass reads it before the bin pass rewrites it (no hot tiles read it there)
    B.cin = mg->c0.p;
    B.cout = mg->c0.p;
    B.basev = mg->row_base.p;
    B.damp = damping;
    B.n = (uint32_t)mg->nv;
    B.n_items2 = L.n_items2;
    B.bin_bits = L.bin_bits;
    B.n_items1 = L.n_items1;
    B.n_items3 = L.n_items3;
    B.acc = mg->acc.p;
    B.threshold = (double)threshold;
    B.row_graph = mg->row_graph.p;
    B.ngraphs = G;
  }

  CaptureKey key;
  key.ptrs[0] = mg->acc.p;
  key.ptrs[1] = mg->binned ? mg->bvals.p : nullptr;
  key.iters = max_iters;
  key.damping = damping;
  key.threshold = threshold;
  key.extra = grid | (mg->binned ? 0x80000000u : 0u);
  if (!mg->cap.valid || !(mg->cap.key == key)) {
    mg->cap.reset();
    hipStream_t cs = ctx->capture_stream;
    NWG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      multi_init_kernel<<<grid1(std::max<uint64_t>(mg->nv, max_iters * G)), kThreads, 0, cs>>>(
          mg->nv, mg->deg.p, mg->vinit.p, mg->x.p, mg->c0.p, mg->acc.p, max_iters * G);
      for (uint64_t k = 0; k < max_iters; ++k) {
        if (mg->binned) {
          B.iter = (uint32_t)k;
          binned_enqueue_sweep(cs, B);  // throws on a failed launch
          continue;
        }
        P.iter = (uint32_t)k;
        P.cin = (k & 1) ? mg->c1.p : mg->c0.p;
        P.cout = (k & 1) ? mg->c0.p : mg->c1.p;
        multi_sweep_kernel<<<grid, kThreads, 0, cs>>>(P);
      }
    } catch (...) {
      // leave the capture stream out of capture mode, so that later solves can record again
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(cs, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      mg->cap.reset();
      throw;
    }
    hipGraph_t gr = nullptr;
    const hipError_t e1 = hipGetLastError();
    NWG_HIP(hipStreamEndCapture(cs, &gr));
    mg->cap.graph = gr;
    NWG_HIP(e1);
    NWG_HIP(hipGraphInstantiate(&mg->cap.exec, mg->cap.graph, nullptr, nullptr, 0));
    mg->cap.key = key;
    mg->cap.valid = true;
  }
  NWG_HIP(hipEventRecord(ctx->ev0, s));
  NWG_HIP(hipGraphLaunch(mg->cap.exec, s));
  NWG_HIP(hipEventRecord(ctx->ev1, s));
  std::vector<unsigned long long> acc(std::max<uint64_t>(max_iters, 1) * G);
  NWG_HIP(hipMemcpyAsync(acc.data(), mg->acc.p, max_iters * G * 8ull, hipMemcpyDeviceToHost, s));
  if (scores_dev) NWG_HIP(hipMemcpyAsync(scores_dev, mg->x.p, mg->nv * 4, hipMemcpyDeviceToDevice, s));
  NWG_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  NWG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  uint64_t max_it = 0;
  for (uint32_t g = 0; g < G; ++g) {
    uint64_t itg = max_iters;
    double last = 0.0;
    for (uint64_t k = 0; k < max_iters; ++k) {
      last = fixed_to_err(acc[k * G + g]);
      if (last < (double)threshold) {
        itg = k + 1;
        break;
      }
    }
    if (iterations) iterations[g] = itg;
    if (final_error) final_error[g] = max_iters ? last : 0.0;
    max_it = std::max(max_it, itg);
  }
  if (info) {
    info->iterations = max_it;
    info->final_error = 0.0;
    info->ms_total = ms;
    info->ms_per_iter = max_it ? ms / (double)max_it : 0.0;
    info->sweep_ms_avg = 0.0;
  }
}

}  // namespace nwg

namespace nwg {
uint64_t multigraph_vertices(const nwg_multigraph* mg) { return mg->nv; }
}  // namespace nwg

extern "C" {

void nwg_multigraph_destroy(nwg_multigraph* mg) {
  if (!mg) return;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(mg->device);
  (void)hipDeviceSynchronize();
  delete mg;
  if (prev >= 0) (void)hipSetDevice(prev);
}

int nwg_multigraph_info(const nwg_multigraph* mg, uint32_t* ngraphs, uint64_t* total_n, uint64_t* total_nnz) {
  if (!mg) return NWG_EINVAL;
  if (ngraphs) *ngraphs = mg->ngraphs;
  if (total_n) *total_n = mg->nv;
  if (total_nnz) *total_nnz = mg->ne;
  return NWG_OK;
}

}  // extern "C"
