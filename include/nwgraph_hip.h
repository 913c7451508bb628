/*
 * nwgraph_hip.h -- C ABI of the MI355X (gfx950) PageRank library libnwgraph_hip.so.
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point returns 0
 * (NWG_OK) or a negative status; the message of the last failure on a context is
 * available from nwg_last_error().  Nothing throws across this boundary.
 *
 * Memory: functions whose pointer arguments are marked [host] copy in and out of
 * device memory themselves and are synchronous.  Arguments marked [device] take
 * device pointers and run on the context's stream (nwg_set_stream); they are
 * synchronous too unless stated otherwise.
 *
 * Reference interfaces (pnnl/NWGraph) replaced by each entry point are cited
 * beside it.  The reference has no FFI: its path is the header template
 *   nw::graph::page_rank(graph, degrees, page_rank, damping, threshold, max_iters, num_threads)
 *   (include/nwgraph/algorithms/page_rank.hpp:89-91)
 * which include/nwgraph_hip/page_rank.hpp re-exposes on top of this ABI.
 */
#ifndef NWGRAPH_HIP_H
#define NWGRAPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define NWG_OK 0
#define NWG_EINVAL (-22)  /* bad argument: null pointer, size mismatch, n == 0, ... */
#define NWG_ENOMEM (-12)  /* device or host allocation failed */
#define NWG_EHIP (-5)     /* a HIP runtime call failed */
#define NWG_ENOTSUP (-95) /* unsupported request (e.g. graph without weights for the weighted path) */
#define NWG_EFORMAT (-74) /* malformed input file */

/* ---- memory kinds for buffer arguments ---- */
#define NWG_MEM_HOST 0
#define NWG_MEM_DEVICE 1

typedef struct nwg_ctx nwg_ctx;             /* one device, one stream, reusable buffers */
typedef struct nwg_graph nwg_graph;         /* device-resident CSC (+ optional CSR, weights, plan) */
typedef struct nwg_multigraph nwg_multigraph; /* block-diagonal union of many graphs */

/* Result of one PageRank solve; the reference prints the same facts per sweep
 * (pagerank::trace, page_rank.hpp:42-52,142). */
typedef struct nwg_pr_info {
  uint64_t iterations;   /* sweeps executed (<= max_iters) */
  double final_error;    /* L1 change of the last sweep */
  double ms_total;       /* device time of the solve (HIP events), init included */
  double ms_per_iter;    /* ms_total / iterations */
  double* error_trace;   /* [host, optional] receives one L1 change per executed sweep */
  uint64_t error_trace_len; /* capacity of error_trace */
  uint32_t flags;        /* in: NWG_INFO_* request bits */
  double sweep_ms_avg;   /* out with NWG_INFO_KERNEL_TIMING: mean duration of one sweep launch */
  double* time_trace;    /* [host, optional] with NWG_INFO_KERNEL_TIMING: each executed sweep's
                            duration in ms (the reference traces every sweep's time, page_rank.hpp:123-142) */
  uint64_t time_trace_len; /* capacity of time_trace */
  double init_ms[2];     /* out with NWG_INFO_KERNEL_TIMING: device time of the score init and of the
                            contribution init (life_timers "init page rank", "init contrib",
                            page_rank.hpp:99-118) */
  double threshold_used; /* out (weighted solves): the absolute L1 change the stop compared with */
} nwg_pr_info;

/* nwg_pr_info.flags: launch the sweeps one by one (instead of as one captured HIP
 * graph) and time each sweep kernel with HIP events on the context's stream. */
#define NWG_INFO_KERNEL_TIMING 1u
/* nwg_pr_info.flags, weighted forward / adjoint only: the threshold is relative.  A solve
 * stops when its L1 change falls below threshold * ||b||_1 / (1 - d), b the constant term of
 * its fixed point: (1 - d) v for the forward (so the scale is ||v||_1, and 1 for the uniform
 * teleport: a unit teleport stops exactly as with an absolute threshold), grad_x for the
 * adjoint (whose L1 changes scale with ||grad_x||_1: an absolute threshold meant for a unit
 * teleport asks a large gradient for many more orders of magnitude).  A zero b keeps the
 * absolute threshold. */
#define NWG_INFO_RELATIVE_THRESHOLD 2u

/* Tunable scheduling plan of the pull sweep (see nwg_graph_autotune). */
typedef struct nwg_plan_params {
  uint32_t short_max;    /* rows with in-degree <= short_max: one lane per row */
  uint32_t group_max;    /* rows with in-degree <= group_max: a 4/8/16/32-lane group per row */
  uint32_t hub_min;      /* rows with in-degree >= hub_min: split over several workgroups */
  uint32_t hub_segment;  /* nonzeros per workgroup segment of a hub row */
  uint32_t slice_min;    /* rows with slice_min <= in-degree < hub_min split their in-edges by source
                            slice ((id / 32) mod 8) so that each XCD gathers from one slice (0 = off;
                            else above short_max); costs a second copy of the column ids */
  uint32_t stream_hints; /* non-temporal hints.  Pull sweep: 0 none, 1 the column-id stream, 2 every
                            streamed array.  Binned sweep: 0 none, 1 loads of the streams read once
                            per sweep, 2 also the stores of pair values, partial rows and scores */
  uint32_t ppr_tile;     /* k-tile of the batched solve: 8, 16, 32 or 64 vectors per lane group (0 = widest <= k) */
  uint32_t relabel;      /* 1: sweep a copy of the CSC renumbered by out-degree (results returned in original ids) */
  uint32_t binned;       /* 1: binned sweep (propagation blocking over a tiled copy of the CSC, 8 bytes per
                            entry; row sums exact in fixed point) instead of the degree-binned pull sweep */
  uint32_t bin_bits;     /* binned sweep: rows per bin, log2: 13 or 14 (0 = by size: 14 from 2^28 edges) */
} nwg_plan_params;

/* ---------------------------------------------------------------- context */
int nwg_version(void);
/* Create a context on HIP device `device`.  Creates its own stream. */
int nwg_create(nwg_ctx** out, int device);
void nwg_destroy(nwg_ctx* ctx);
/* Run subsequent work on `stream` (a hipStream_t; NULL = the context's own stream). */
int nwg_set_stream(nwg_ctx* ctx, void* stream);
const char* nwg_last_error(const nwg_ctx* ctx);

/* ------------------------------------------------------------ drop-in solve */
/* One-call PageRank over a host CSC: replaces nw::graph::page_rank
 * (page_rank.hpp:89-148) for adjacency<1> = CSC with uint32 indices_ (n+1) and
 * to_be_indexed_ column 0 (nnz source ids).  outdeg = build_degrees() output
 * (bench/gapbs/common.hpp:110-121) or NULL to compute it on the device.  scores
 * (n floats) is written in place like the reference's pre-sized page_rank vector.
 * Semantics: Jacobi sweeps (the v10 twin, experimental/algorithms/page_rank.hpp:377-432)
 * with the reference's scalars; stops after the first sweep whose L1 change is
 * below threshold, or after max_iters sweeps.  The graph lives for this call only, so it
 * runs the pull sweep, which needs no layout built; a graph solved more than once is
 * faster resident (nwg_graph_create_csc + nwg_graph_pagerank), where the binned layout is
 * built once. */
int nwg_pagerank_f32(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr /*[host]*/,
                     const uint32_t* csc_idx /*[host]*/, const uint32_t* outdeg /*[host] or NULL*/,
                     float* scores /*[host] n*/, float damping, float threshold, uint64_t max_iters,
                     nwg_pr_info* info /*nullable*/);

/* Allocation-free variant of nwg_pagerank_f32.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The CSC, the optional out-degrees and the scores
 * are [device] pointers; the call runs on the context's stream and makes no hipMalloc,
 * hipFree or other allocating call.  Its working memory (contributions, counted degrees, L1
 * partials and accumulators) is one static 48.2 MiB device array per GPU, looked up by
 * nwg_create and held until the process ends.  Limits: n <= 4194304 vertices and
 * max_iters <= 4096, else NWG_EINVAL; csc_idx 16-byte aligned; NWG_INFO_KERNEL_TIMING is
 * not supported (NWG_EINVAL).  Calls that overlap on one GPU share the array and race
 * (calls through one context are ordered by its stream).  Semantics as nwg_pagerank_f32
 * (Jacobi sweeps, the reference's scalars); its sweep needs no plan, so row sums may differ
 * from nwg_pagerank_f32's in FP32 summation order. */
int nwg_pagerank_f32_noalloc(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr /*[device]*/,
                             const uint32_t* csc_idx /*[device]*/, const uint32_t* outdeg /*[device] or NULL*/,
                             float* scores /*[device] n*/, float damping, float threshold, uint64_t max_iters,
                             nwg_pr_info* info /*nullable*/);

/* ------------------------------------------------------------ graph objects */
/* Upload a CSC (rows = targets, entries = sources; each row ascending or not) --
 * the layout of nw::graph::adjacency<1> (adjacency.hpp:86-192, compressed.hpp:64-102).
 * A given outdeg must count at least each vertex's entries in csc_idx (build_degrees
 * counts exactly those, common.hpp:110-121), else NWG_EINVAL. */
int nwg_graph_create_csc(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr,
                         const uint32_t* csc_idx, const uint32_t* outdeg /*nullable*/, int mem_kind,
                         nwg_graph** out);
/* Build a cleaned CSC on the device from an edge list (src[i] -> dst[i]):
 * lexical sort, uniq, remove_self_loops (build.hpp:57-70,392-410), then
 * fill_directed<1> (build.hpp:209-232) with rows sorted (compressed.hpp:491-502),
 * and out-degrees (common.hpp:110-121).  clean=0 keeps duplicates and self-loops
 * (the reference's adjacency<1> as built by pr.cpp) but still sorts every row. */
int nwg_graph_build_from_edges(nwg_ctx* ctx, uint32_t n, uint64_t ne, const uint32_t* src,
                               const uint32_t* dst, int clean, int mem_kind, nwg_graph** out);
void nwg_graph_destroy(nwg_graph* g);
int nwg_graph_info(const nwg_graph* g, uint32_t* n, uint64_t* nnz);
/* Copy the CSC (and out-degrees) back; any pointer may be NULL. */
int nwg_graph_get_csc(nwg_ctx* ctx, const nwg_graph* g, uint32_t* csc_ptr /*[host] n+1*/,
                      uint32_t* csc_idx /*[host] nnz*/, uint32_t* outdeg /*[host] n*/);
/* Build (once) and copy the CSR (fill_directed<0>, rows sorted) and the map from CSR
 * position to CSC position. */
int nwg_graph_get_csr(nwg_ctx* ctx, nwg_graph* g, uint32_t* csr_ptr /*[host] n+1*/,
                      uint32_t* csr_idx /*[host] nnz*/, uint32_t* csr_to_csc /*[host] nnz, nullable*/);
/* Device pointers of the resident CSC arrays (valid until nwg_graph_destroy). */
int nwg_graph_device_csc(const nwg_graph* g, const uint32_t** csc_ptr, const uint32_t** csc_idx,
                         const uint32_t** outdeg);

/* -------------------------------------------------------------- generators */
/* RMAT graph with 2^scale vertices and edge_factor * 2^scale raw edges, quadrant
 * probabilities a, b, c (and d = 1 - a - b - c) in units of 1/10000, Philox4x32-10
 * keyed by seed, no noise and no vertex permutation; generated and cleaned on the
 * device (self-loops and duplicates removed). */
int nwg_generate_rmat(nwg_ctx* ctx, uint32_t scale, uint32_t edge_factor, uint32_t a, uint32_t b,
                      uint32_t c, uint64_t seed, nwg_graph** out);
/* The raw (uncleaned) RMAT edges [first, first+count) -- for generator parity checks. */
int nwg_generate_rmat_edges(nwg_ctx* ctx, uint32_t scale, uint32_t a, uint32_t b, uint32_t c,
                            uint64_t seed, uint64_t first, uint64_t count, uint32_t* src,
                            uint32_t* dst, int mem_kind);
/* width x width 4-neighbour lattice, each undirected lattice edge kept with probability
 * keep/10000 and stored both ways, plus one random long-range edge from each vertex
 * with probability long/10000; cleaned on the device. */
int nwg_generate_grid(nwg_ctx* ctx, uint32_t width, uint32_t keep_per10000, uint32_t long_per10000,
                      uint64_t seed, nwg_graph** out);
int nwg_generate_grid_edges(nwg_ctx* ctx, uint32_t width, uint32_t keep_per10000,
                            uint32_t long_per10000, uint64_t seed, uint32_t* src, uint32_t* dst,
                            int mem_kind);
/* number of raw slots nwg_generate_grid_edges writes */
uint64_t nwg_grid_slots(uint32_t width);
/* Teleport seed sets for personalised PageRank: k vectors of s distinct ids in [0, n),
 * vector j draws mulhi(Philox({j, t, 2, 'PPRS'}, seed)[0], n) for t = 0, 1, ... keeping
 * the first s distinct ids.  ids: [host] k*s, row-major. */
int nwg_generate_ppr_seeds(uint32_t n, uint32_t k, uint32_t s, uint64_t seed, uint32_t* ids);

/* ------------------------------------------------------ plain PageRank solve */
/* PageRank on a resident graph.  scores: n floats ([host] or [device] per mem_kind).
 * teleport: NULL for the reference's uniform base (1-d)/N, else n floats (same
 * mem_kind) giving the personalised base (1-d)*teleport[i]; its entries must be finite,
 * in [0, 1] and sum to at most 1 (+1e-4 rounding slack), else NWG_EINVAL (scores then
 * stay in [0, 1], the range the sweeps' exact fixed-point sums cover).  Replaces page_rank
 * (page_rank.hpp:89-148) with the v10 Jacobi schedule; the whole max_iters loop
 * runs as one captured HIP graph with on-device convergence (no host round trip
 * per sweep). */
int nwg_graph_pagerank(nwg_ctx* ctx, nwg_graph* g, float* scores, const float* teleport,
                       float damping, float threshold, uint64_t max_iters, int mem_kind,
                       nwg_pr_info* info);
/* Same solve with the reference's own in-place schedule: nw::graph::page_rank
 * (page_rank.hpp:121-147) updates outgoing_contrib inside the sweep (:135), so rows later
 * in a sweep see earlier rows' new contributions.  Here one contribution array is updated
 * during the sweep.  With the pull kernels, rows see whatever neighbours have already
 * written.  With the binned sweep (plan option binned), each sweep's non-hot contributions
 * are staged before any row is finished, so updates reach other rows within the sweep only
 * through the hot sources.  Like the reference's multi-threaded run the result depends on
 * the schedule, so it is not reproducible bit for bit; it converges to the same fixed point. */
int nwg_graph_pagerank_inplace(nwg_ctx* ctx, nwg_graph* g, float* scores, const float* teleport,
                               float damping, float threshold, uint64_t max_iters, int mem_kind,
                               nwg_pr_info* info);
/* The reference's delta (residual) PageRank, page_rank_v13
 * (include/nwgraph/experimental/algorithms/page_rank.hpp:560-625, pr --version 13).  Every
 * vertex starts with residual 1 - d and score 0.  An iteration absorbs each residual above
 * `threshold` into the score and sends r * d / outdeg to the out-neighbours (:580-594); then
 * every residual becomes the sum of what the vertex received (:597-612), so residuals at or
 * below the threshold that were not absorbed are dropped.  The loop ends after the first
 * iteration in which no vertex with out-edges pushed (:616-620), or after max_iters.  Scores
 * are not normalised (they sum to about N, not 1).  The pull phase runs on the plan's pull
 * kernels (the binned option is ignored); the loop is one captured HIP graph that stops
 * working after the last iteration without a host round trip.  info->iterations counts the
 * iterations executed; the error trace is 0, as the reference traces it (:615). */
int nwg_graph_pagerank_delta(nwg_ctx* ctx, nwg_graph* g, float* scores, float damping,
                             float threshold, uint64_t max_iters, int mem_kind, nwg_pr_info* info);
/* Scheduling plan of the sweep: set explicitly, read back, or autotune.  Without either, the
 * first PageRank solve of a graph builds a plan by size: the binned sweep from 2^18 edges,
 * on the degree-ordered copy from 2^23 edges, else the pull sweep (get_plan reports it). */
int nwg_graph_set_plan(nwg_ctx* ctx, nwg_graph* g, const nwg_plan_params* p);
int nwg_graph_get_plan(const nwg_graph* g, nwg_plan_params* p);
/* Time candidate plans on this graph with HIP events and keep the fastest; results are
 * cached per (in-degree histogram, device) for the process and, when the environment
 * variable NWG_TUNE_CACHE names a file, across processes. */
int nwg_graph_autotune(nwg_ctx* ctx, nwg_graph* g, nwg_plan_params* chosen /*nullable*/);

/* ------------------------------------------- batched personalised PageRank */
/* k personalised PageRank vectors solved together as one SpMM per sweep.
 * Teleport j is uniform over seeds[seed_ptr[j] .. seed_ptr[j+1]) ([host] arrays).
 * scores: n x k row-major (score of vertex i for vector j at [i*k + j]).
 * Each vector stops updating after its own first sweep with L1 change < threshold.  The solve
 * stores no scores between sweeps: a row's old score is read back from its previous
 * contribution as (x / deg) * deg, within 1 ulp of x, so a vector's L1 change can differ from a
 * single-vector solve's by up to ~2^-23 * sum(x) (< 1.2e-7 for scores summing to 1), and a
 * vector whose change lands that close to the threshold may stop one sweep apart from it.
 * iterations: [host] k counts, nullable. */
int nwg_ppr_batched(nwg_ctx* ctx, nwg_graph* g, uint32_t k, const uint32_t* seed_ptr,
                    const uint32_t* seeds, float* scores, float damping, float threshold,
                    uint64_t max_iters, int mem_kind, uint64_t* iterations, nwg_pr_info* info);
/* Time the batched sweep with every k-tile (8, 16, 32, 64; at most k rounded up) on this
 * graph and teleport set, keep the fastest in the graph's plan (ppr_tile) and return it. */
int nwg_ppr_autotune(nwg_ctx* ctx, nwg_graph* g, uint32_t k, const uint32_t* seed_ptr,
                     const uint32_t* seeds, uint32_t* tile);

/* --------------------------------------------------- many graphs, one launch */
/* Concatenate graphs into one block-diagonal device graph. */
int nwg_multigraph_create(nwg_ctx* ctx, uint32_t ngraphs, nwg_graph* const* graphs,
                          nwg_multigraph** out);
void nwg_multigraph_destroy(nwg_multigraph* mg);
int nwg_multigraph_info(const nwg_multigraph* mg, uint32_t* ngraphs, uint64_t* total_n,
                        uint64_t* total_nnz);
/* All graphs solved by one persistent kernel per sweep (per-graph convergence),
 * the sweep loop captured in a HIP graph.  scores: concatenation of all graphs'
 * scores (total_n floats).  iterations / final_error: [host] ngraphs each, nullable. */
int nwg_multigraph_pagerank(nwg_ctx* ctx, nwg_multigraph* mg, float* scores, float damping,
                            float threshold, uint64_t max_iters, int mem_kind,
                            uint64_t* iterations, double* final_error, nwg_pr_info* info);
/* Schedule of the union's sweeps: only plan->binned is read (0: one persistent kernel per
 * sweep over row items; 1: the binned sweep over a tiled copy of the union, rows of a
 * converged graph frozen).  No counterpart in the reference (it has no batched solve). */
int nwg_multigraph_set_plan(nwg_ctx* ctx, nwg_multigraph* mg, const nwg_plan_params* plan);
/* Times both schedules over a few sweeps and keeps the faster; *binned_out (nullable)
 * receives the choice. */
int nwg_multigraph_autotune(nwg_ctx* ctx, nwg_multigraph* mg, uint32_t* binned_out);

/* ------------------------------------------- weighted PageRank and adjoint */
/* Edge weights in CSC order (weight of edge col[p] -> row).  With normalize=1 each
 * source's outgoing weights are divided by their sum (row-normalised transition). */
int nwg_graph_set_weights(nwg_ctx* ctx, nwg_graph* g, const float* w_csc, int normalize,
                          int mem_kind);
/* Fill raw weights w(u,v) = 0.1 + 0.9 t, t ~ U[0,1) from Philox (u, v, seed), then set them. */
int nwg_graph_random_weights(nwg_ctx* ctx, nwg_graph* g, uint64_t seed, int normalize);
int nwg_graph_get_weights(nwg_ctx* ctx, const nwg_graph* g, float* w_csc /*[host] nnz*/);
/* Forward: x = (1-d) v + d A x,  A[i][u] = w(u->i) (normalised weights).  teleport:
 * n floats or NULL (uniform 1/n). */
int nwg_wpr_forward(nwg_ctx* ctx, nwg_graph* g, const float* teleport, float* x, float damping,
                    float threshold, uint64_t max_iters, int mem_kind, nwg_pr_info* info);
/* Backward of L(x) with dL/dx = grad_x at the forward fixed point x: solves the adjoint
 * fixed point lambda = grad_x + d A^T lambda (pull over the CSR, no unrolled iterates),
 * then grad_w[p] = d * lambda[row(p)] * x[col(p)] for the normalised weights (CSC order)
 * and grad_teleport = (1-d) lambda.  With raw_weights=1 grad_w is chained through the
 * row normaliser to the raw weights.  grad_w / grad_teleport nullable. */
int nwg_wpr_backward(nwg_ctx* ctx, nwg_graph* g, const float* x, const float* grad_x,
                     float* grad_w, float* grad_teleport, float damping, float threshold,
                     uint64_t max_iters, int raw_weights, int mem_kind, nwg_pr_info* info);

/* ------------------------------------------------------------ calibration */
/* Plain device copy of `bytes` (multiple of 16) from src to dst ([device] pointers) with a
 * streaming kernel; returns the kernel time in *ms.  Used to calibrate HBM bandwidth and
 * the hardware byte counters against a transfer of known size. */
int nwg_stream_copy(nwg_ctx* ctx, void* dst, const void* src, uint64_t bytes, double* ms);
/* Streaming ceilings: kind 0 copies src to dst, 1 only reads src, 2 only writes dst, 3 reads
 * src 16 times in one launch (the rate of a buffer that stays cached);
 * flags bit 0 = nontemporal loads, bit 1 = nontemporal stores; wg_per_cu workgroups of
 * 256 lanes per CU (1..32), four 16-byte accesses in flight per lane.  Returns the kernel
 * time in *ms.  No counterpart in the reference (a measurement aid). */
#define NWG_PROBE_COPY 0u
#define NWG_PROBE_READ 1u
#define NWG_PROBE_WRITE 2u
#define NWG_PROBE_REREAD 3u
#define NWG_PROBE_NT_LOAD 1u
#define NWG_PROBE_NT_STORE 2u
int nwg_stream_probe(nwg_ctx* ctx, uint32_t kind, uint32_t flags, uint32_t wg_per_cu, void* dst,
                     const void* src, uint64_t bytes, double* ms);

/* ------------------------------------------------------------------ input */
/* Matrix Market reader (read_mm<directed>, io/mmio.hpp:171-217 with mm_fill :86-103):
 * 1-based -> 0-based, symmetric files mirror off-diagonal entries, n = max(header
 * rows, max id + 1).  *src and *dst are malloc'ed; free with nwg_free. */
int nwg_read_mm(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst);
/* NW binary edge list (edge_list::serialize/deserialize, edge_list.hpp:237-280).  A file with
 * an id at or above its own vertex count is malformed (NWG_EFORMAT). */
int nwg_read_nw(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst);
int nwg_write_nw(const char* path, uint32_t n, uint64_t ne, const uint32_t* src, const uint32_t* dst);
/* Binary adjacency (index_adjacency<idx>::serialize / deserialize, adjacency.hpp:176-191,
 * containers/compressed.hpp:377-413, graph_base.hpp:72-87, soa.hpp:333-375), the
 * attribute-free uint32 form: n + 1 row offsets ptr and nnz ids idx (a CSC for idx = 1, a
 * CSR for idx = 0).  The reader checks the layout, offsets and ids and mallocs *ptr (n + 1)
 * and *idx (nnz); free with nwg_free. */
int nwg_write_adjacency(const char* path, uint32_t n, uint64_t nnz, const uint32_t* ptr, const uint32_t* idx);
int nwg_read_adjacency(const char* path, uint32_t* n, uint64_t* nnz, uint32_t** ptr, uint32_t** idx);
void nwg_free(void* p);

#ifdef __cplusplus
}
#endif
#endif /* NWGRAPH_HIP_H */
