// capi.cpp -- the extern "C" boundary of libnwgraph_hip: argument checking, status
// codes, host<->device staging.  No exception crosses this file's functions.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "capi_util.hpp"
#include "nwg_internal.hpp"
#include "philox.hpp"

using nwg::check_graph;
using nwg::check_mem_kind;
using nwg::DevBuf;
using nwg::DeviceGuard;
using nwg::Error;
using nwg::guarded;
using nwg::stage_in;

namespace {

// validate a host CSC: n+1 monotone pointers ending at nnz, entries < n
void validate_host_csc(uint32_t n, uint64_t nnz, const uint32_t* ptr, const uint32_t* idx) {
  NWG_REQUIRE(ptr[0] == 0, "csc_ptr[0] must be 0");
  NWG_REQUIRE(ptr[n] == nnz, "csc_ptr[n] must equal nnz");
  for (uint32_t i = 0; i < n; ++i) NWG_REQUIRE(ptr[i] <= ptr[i + 1], "csc_ptr must be non-decreasing");
  for (uint64_t p = 0; p < nnz; ++p) NWG_REQUIRE(idx[p] < n, "csc_idx holds a vertex id >= n");
}

nwg_graph* new_graph(nwg_ctx* ctx) {
  nwg_graph* g = new nwg_graph();
  g->ctx = ctx;
  g->device = ctx->device;
  return g;
}

}  // namespace

extern "C" {

int nwg_version(void) { return 10000; }

int nwg_create(nwg_ctx** out, int device) {
  if (!out) return NWG_EINVAL;
  *out = nullptr;
  nwg_ctx* ctx = new (std::nothrow) nwg_ctx();
  if (!ctx) return NWG_ENOMEM;
  int rc = guarded(ctx, [&] {
    int count = 0;
    NWG_HIP(hipGetDeviceCount(&count));
    NWG_REQUIRE(device >= 0 && device < count, "no such HIP device");
    ctx->device = device;
    DeviceGuard dg(device);
    // a blocking stream: ordered after work a caller queued on the legacy default stream
    NWG_HIP(hipStreamCreate(&ctx->own_stream));
    // a private stream that only records HIP graphs (never executes work)
    NWG_HIP(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
    ctx->stream = ctx->own_stream;
    NWG_HIP(hipEventCreate(&ctx->ev0));
    NWG_HIP(hipEventCreate(&ctx->ev1));
    nwg::noalloc_setup(device);
  });
  if (rc != NWG_OK) {
    nwg_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return NWG_OK;
}

void nwg_destroy(nwg_ctx* ctx) {
  if (!ctx) return;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(ctx->device);
  if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
  if (prev >= 0) (void)hipSetDevice(prev);
  delete ctx;
}

int nwg_set_stream(nwg_ctx* ctx, void* stream) {
  if (!ctx) return NWG_EINVAL;
  ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
  return NWG_OK;
}

const char* nwg_last_error(const nwg_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

// ------------------------------------------------------------------ graphs
int nwg_graph_create_csc(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr, const uint32_t* csc_idx,
                         const uint32_t* outdeg, int mem_kind, nwg_graph** out) {
  if (!ctx || !out) return NWG_EINVAL;
  *out = nullptr;
  return guarded(ctx, [&] {
    check_mem_kind(mem_kind);
    NWG_REQUIRE(n >= 1, "graph needs at least one vertex");
    NWG_REQUIRE(csc_ptr && (csc_idx || nnz == 0), "csc arrays must not be null");
    NWG_REQUIRE(nnz < (1ull << 32), "nnz must be below 2^32 (uint32 indices)");
    if (mem_kind == NWG_MEM_HOST) validate_host_csc(n, nnz, csc_ptr, csc_idx);
    DeviceGuard dg(ctx->device);
    std::unique_ptr<nwg_graph> g(new_graph(ctx));
    hipStream_t s = ctx->stream;
    g->n = n;
    g->nnz = nnz;
    stage_in(s, g->csc_ptr, csc_ptr, (size_t)n + 1, mem_kind);
    stage_in(s, g->csc_idx, csc_idx, nnz, mem_kind);
    // a device CSC is checked on the device before any kernel indexes by its entries
    if (mem_kind != NWG_MEM_HOST) nwg::validate_device_csc(s, n, nnz, g->csc_ptr.p, g->csc_idx.p);
    if (outdeg) {
      // a smaller out-degree than a vertex's out-edges would let its contributions sum past
      // its score, out of the range the sweeps' exact fixed-point sums cover
      stage_in(s, g->outdeg, outdeg, n, mem_kind);
      nwg::check_outdeg_device(s, n, nnz, g->csc_idx.p, g->outdeg.p);
    } else {
      g->outdeg.alloc(n);
      nwg::out_degrees_device(s, n, nnz, g->csc_idx.p, g->outdeg.p);
    }
    NWG_HIP(hipStreamSynchronize(s));
    *out = g.release();
  });
}

int nwg_graph_build_from_edges(nwg_ctx* ctx, uint32_t n, uint64_t ne, const uint32_t* src, const uint32_t* dst,
                               int clean, int mem_kind, nwg_graph** out) {
  if (!ctx || !out) return NWG_EINVAL;
  *out = nullptr;
  return guarded(ctx, [&] {
    check_mem_kind(mem_kind);
    NWG_REQUIRE(n >= 1, "graph needs at least one vertex");
    NWG_REQUIRE((src && dst) || ne == 0, "edge arrays must not be null");
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    DevBuf<uint32_t> ds, dd;
    stage_in(s, ds, src, ne, mem_kind);
    stage_in(s, dd, dst, ne, mem_kind);
    std::unique_ptr<nwg_graph> g(new_graph(ctx));
    nwg::build_csc_device(s, n, ne, ds.p, dd.p, clean != 0, g.get());
    *out = g.release();
  });
}

void nwg_graph_destroy(nwg_graph* g) {
  if (!g) return;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(g->device);
  (void)hipDeviceSynchronize();
  delete g;
  if (prev >= 0) (void)hipSetDevice(prev);
}

int nwg_graph_info(const nwg_graph* g, uint32_t* n, uint64_t* nnz) {
  if (!g) return NWG_EINVAL;
  if (n) *n = g->n;
  if (nnz) *nnz = g->nnz;
  return NWG_OK;
}

int nwg_graph_get_csc(nwg_ctx* ctx, const nwg_graph* g, uint32_t* csc_ptr, uint32_t* csc_idx, uint32_t* outdeg) {
  if (!ctx) return NWG_EINVAL;
  return guarded(ctx, [&] {
    check_graph(ctx, g);
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    if (csc_ptr)
      NWG_HIP(hipMemcpyAsync(csc_ptr, g->csc_ptr.p, ((size_t)g->n + 1) * 4, hipMemcpyDeviceToHost, s));
    if (csc_idx && g->nnz)
      NWG_HIP(hipMemcpyAsync(csc_idx, g->csc_idx.p, (size_t)g->nnz * 4, hipMemcpyDeviceToHost, s));
    if (outdeg) NWG_HIP(hipMemcpyAsync(outdeg, g->outdeg.p, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
  });
}

int nwg_graph_get_csr(nwg_ctx* ctx, nwg_graph* g, uint32_t* csr_ptr, uint32_t* csr_idx, uint32_t* csr_to_csc) {
  if (!ctx) return NWG_EINVAL;
  return guarded(ctx, [&] {
    check_graph(ctx, g);
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    nwg::build_csr_device(s, g);
    if (csr_ptr)
      NWG_HIP(hipMemcpyAsync(csr_ptr, g->csr_ptr.p, ((size_t)g->n + 1) * 4, hipMemcpyDeviceToHost, s));
    if (csr_idx && g->nnz)
      NWG_HIP(hipMemcpyAsync(csr_idx, g->csr_idx.p, (size_t)g->nnz * 4, hipMemcpyDeviceToHost, s));
    if (csr_to_csc && g->nnz)
      NWG_HIP(hipMemcpyAsync(csr_to_csc, g->csr_to_csc.p, (size_t)g->nnz * 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
  });
}

int nwg_graph_device_csc(const nwg_graph* g, const uint32_t** csc_ptr, const uint32_t** csc_idx,
                         const uint32_t** outdeg) {
  if (!g) return NWG_EINVAL;
  if (csc_ptr) *csc_ptr = g->csc_ptr.p;
  if (csc_idx) *csc_idx = g->csc_idx.p;
  if (outdeg) *outdeg = g->outdeg.p;
  return NWG_OK;
}

// ------------------------------------------------------------- generators
int nwg_generate_rmat(nwg_ctx* ctx, uint32_t scale, uint32_t edge_factor, uint32_t a, uint32_t b, uint32_t c,
                      uint64_t seed, nwg_graph** out) {
  if (!ctx || !out) return NWG_EINVAL;
  *out = nullptr;
  return guarded(ctx, [&] {
    NWG_REQUIRE(scale >= 1 && scale <= 31, "rmat scale must be in [1, 31]");
    const uint64_t n = 1ull << scale;
    const uint64_t ne = (uint64_t)edge_factor * n;
    NWG_REQUIRE(ne < (1ull << 32), "edge_factor * 2^scale must be below 2^32");
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    DevBuf<uint32_t> ds(ne ? ne : 1), dd(ne ? ne : 1);
    nwg::rmat_edges_device(s, scale, a, b, c, seed, 0, ne, ds.p, dd.p);
    std::unique_ptr<nwg_graph> g(new_graph(ctx));
    nwg::build_csc_device(s, (uint32_t)n, ne, ds.p, dd.p, true, g.get());
    *out = g.release();
  });
}

int nwg_generate_rmat_edges(nwg_ctx* ctx, uint32_t scale, uint32_t a, uint32_t b, uint32_t c, uint64_t seed,
                            uint64_t first, uint64_t count, uint32_t* src, uint32_t* dst, int mem_kind) {
  if (!ctx) return NWG_EINVAL;
  return guarded(ctx, [&] {
    check_mem_kind(mem_kind);
    NWG_REQUIRE((src && dst) || count == 0, "edge arrays must not be null");
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    if (mem_kind == NWG_MEM_DEVICE) {
      nwg::rmat_edges_device(s, scale, a, b, c, seed, first, count, src, dst);
    } else {
      DevBuf<uint32_t> ds(count ? count : 1), dd(count ? count : 1);
      nwg::rmat_edges_device(s, scale, a, b, c, seed, first, count, ds.p, dd.p);
      if (count) {
        NWG_HIP(hipMemcpyAsync(src, ds.p, count * 4, hipMemcpyDeviceToHost, s));
        NWG_HIP(hipMemcpyAsync(dst, dd.p, count * 4, hipMemcpyDeviceToHost, s));
      }
    }
    NWG_HIP(hipStreamSynchronize(s));
  });
}

int nwg_generate_ppr_seeds(uint32_t n, uint32_t k, uint32_t s, uint64_t seed, uint32_t* ids) {
  if (!ids || n == 0 || s > n) return NWG_EINVAL;
  for (uint32_t j = 0; j < k; ++j) {
    uint32_t* row = ids + (uint64_t)j * s;
    uint32_t got = 0;
    for (uint64_t t = 0; got < s; ++t) {
      const nwg::U32x4 r = nwg::philox4x32_10(j, (uint32_t)t, 2u, nwg::kTagPprs, (uint32_t)seed,
                                              (uint32_t)(seed >> 32));
      const uint32_t id = (uint32_t)(((uint64_t)r.v[0] * n) >> 32);
      bool dup = false;
      for (uint32_t q = 0; q < got; ++q) dup |= (row[q] == id);
      if (!dup) row[got++] = id;
    }
  }
  return NWG_OK;
}

uint64_t nwg_grid_slots(uint32_t width) {
  const uint64_t W = width;
  return W < 2 ? 0 : 4 * W * (W - 1) + W * W;
}

int nwg_generate_grid_edges(nwg_ctx* ctx, uint32_t width, uint32_t keep_per10000, uint32_t long_per10000,
                            uint64_t seed, uint32_t* src, uint32_t* dst, int mem_kind) {
  if (!ctx) return NWG_EINVAL;
  return guarded(ctx, [&] {
    check_mem_kind(mem_kind);
    NWG_REQUIRE(src && dst, "edge arrays must not be null");
    NWG_REQUIRE(keep_per10000 <= 10000 && long_per10000 <= 10000, "probabilities are in units of 1/10000");
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    const uint64_t slots = nwg_grid_slots(width);
    if (mem_kind == NWG_MEM_DEVICE) {
      nwg::grid_edges_device(s, width, keep_per10000, long_per10000, seed, src, dst);
    } else {
      DevBuf<uint32_t> ds(slots), dd(slots);
      nwg::grid_edges_device(s, width, keep_per10000, long_per10000, seed, ds.p, dd.p);
      NWG_HIP(hipMemcpyAsync(src, ds.p, slots * 4, hipMemcpyDeviceToHost, s));
      NWG_HIP(hipMemcpyAsync(dst, dd.p, slots * 4, hipMemcpyDeviceToHost, s));
    }
    NWG_HIP(hipStreamSynchronize(s));
  });
}

int nwg_generate_grid(nwg_ctx* ctx, uint32_t width, uint32_t keep_per10000, uint32_t long_per10000, uint64_t seed,
                      nwg_graph** out) {
  if (!ctx || !out) return NWG_EINVAL;
  *out = nullptr;
  return guarded(ctx, [&] {
    NWG_REQUIRE(keep_per10000 <= 10000 && long_per10000 <= 10000, "probabilities are in units of 1/10000");
    NWG_REQUIRE(width >= 2 && (uint64_t)width * width < (1ull << 32), "grid width must be in [2, 65535]");
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    const uint64_t slots = nwg_grid_slots(width);
    NWG_REQUIRE(slots < (1ull << 32), "grid too large for uint32 edge counts");
    DevBuf<uint32_t> ds(slots), dd(slots);
    nwg::grid_edges_device(s, width, keep_per10000, long_per10000, seed, ds.p, dd.p);
    std::unique_ptr<nwg_graph> g(new_graph(ctx));
    nwg::build_csc_device(s, width * width, slots, ds.p, dd.p, true, g.get());
    *out = g.release();
  });
}

// ---------------------------------------------------------------- PageRank
static int graph_pagerank(nwg_ctx* ctx, nwg_graph* g, float* scores, const float* teleport, float damping,
                          float threshold, uint64_t max_iters, int mem_kind, nwg_pr_info* info,
                          nwg::PrSchedule sched) {
  if (!ctx) return NWG_EINVAL;
  return guarded(ctx, [&] {
    check_graph(ctx, g);
    check_mem_kind(mem_kind);
    NWG_REQUIRE(scores != nullptr, "scores must not be null");
    DeviceGuard dg(ctx->device);
    hipStream_t s = ctx->stream;
    DevBuf<float> tele_dev, out_dev;
    const float* tele = teleport;
    float* sc = scores;
    if (mem_kind == NWG_MEM_HOST) {
      if (teleport) {
        stage_in(s, tele_dev, teleport, g->n, NWG_MEM_HOST);
        tele = tele_dev.p;
      }
      out_dev.alloc(g->n);
      sc = out_dev.p;
    }
    nwg::pagerank_run(ctx, g, tele, damping, threshold, max_iters, info, sc, sched);
    if (mem_kind == NWG_MEM_HOST) {
      NWG_HIP(hipMemcpyAsync(scores, out_dev.p, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
      NWG_HIP(hipStreamSynchronize(s));
    }
  });
}

int nwg_graph_pagerank(nwg_ctx* ctx, nwg_graph* g, float* scores, const float* teleport, float damping,
                       float threshold, uint64_t max_iters, int mem_kind, nwg_pr_info* info) {
  return graph_pagerank(ctx, g, scores, teleport, damping, threshold, max_iters, mem_kind, info,
                        nwg::PrSchedule::kJacobi);
}

int nwg_graph_pagerank_inplace(nwg_ctx* ctx, nwg_graph* g, float* scores, const float* teleport, float damping,
                               float threshold, uint64_t max_iters, int mem_kind, nwg_pr_info* info) {
  return graph_pagerank(ctx, g, scores, teleport, damping, threshold, max_iters, mem_kind, info,
                        nwg::PrSchedule::kInPlace);
}

int nwg_graph_pagerank_delta(nwg_ctx* ctx, nwg_graph* g, float* scores, float damping, float threshold,
                             uint64_t max_iters, int mem_kind, nwg_pr_info* info) {
  return graph_pagerank(ctx, g, scores, nullptr, damping, threshold, max_iters, mem_kind, info,
                        nwg::PrSchedule::kDelta);
}

int nwg_pagerank_f32(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr, const uint32_t* csc_idx,
                     const uint32_t* outdeg, float* scores, float damping, float threshold, uint64_t max_iters,
                     nwg_pr_info* info) {
  if (!ctx) return NWG_EINVAL;
  nwg_graph* g = nullptr;
  int rc = nwg_graph_create_csc(ctx, n, nnz, csc_ptr, csc_idx, outdeg, NWG_MEM_HOST, &g);
  if (rc != NWG_OK) return rc;
  // A graph made for one solve takes the pull sweep on its own numbering: it has nothing to
  // build, while the binned layout and the degree-ordered copy cost more than they save in
  // one solve of a few tens of sweeps (RMAT-25: ~160 ms to build, ~115 ms saved over 20).
  const nwg_plan_params pull = nwg::default_plan_params();
  rc = nwg_graph_set_plan(ctx, g, &pull);
  if (rc == NWG_OK) rc = nwg_graph_pagerank(ctx, g, scores, nullptr, damping, threshold, max_iters, NWG_MEM_HOST, info);
  nwg_graph_destroy(g);
  return rc;
}

int nwg_pagerank_f32_noalloc(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr,
                             const uint32_t* csc_idx, const uint32_t* outdeg, float* scores, float damping,
                             float threshold, uint64_t max_iters, nwg_pr_info* info) {
  if (!ctx) return NWG_EINVAL;
  return guarded(ctx, [&] {
    DeviceGuard dg(ctx->device);
    nwg::pagerank_noalloc(ctx, n, nnz, csc_ptr, csc_idx, outdeg, scores, damping, threshold, max_iters, info);
  });
}

int nwg_graph_set_plan(nwg_ctx* ctx, nwg_graph* g, const nwg_plan_params* p) {
  if (!ctx) return NWG_EINVAL;
  return guarded(ctx, [&] {
    check_graph(ctx, g);
    NWG_REQUIRE(p != nullptr, "plan params must not be null");
    DeviceGuard dg(ctx->device);
    nwg::plan_build(ctx->stream, g, *p);
  });
}

int nwg_graph_get_plan(const nwg_graph* g, nwg_plan_params* p) {
  if (!g || !p) return NWG_EINVAL;
  *p = g->plan.built ? g->plan.params : nwg::default_plan_for(g);
  return NWG_OK;
}

}  // extern "C"
