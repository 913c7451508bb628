// nwg_internal.hpp -- shared host/device definitions of libnwgraph_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nwgraph_hip.h"

namespace nwg {

constexpr int kWave = 64;  // CDNA wavefront width

// Error carried from the implementation to the C boundary, where it becomes a status.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define NWG_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::nwg::Error(e_ == hipErrorOutOfMemory ? NWG_ENOMEM : NWG_EHIP,                \
                         std::string(#call) + " failed: " + hipGetErrorString(e_));        \
  } while (0)

#define NWG_REQUIRE(cond, msg)                                                             \
  do {                                                                                     \
    if (!(cond)) throw ::nwg::Error(NWG_EINVAL, msg);                                      \
  } while (0)

#define NWG_LAUNCH_CHECK() NWG_HIP(hipGetLastError())

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Owning device allocation.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count) NWG_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  // grow-only reuse
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};


// Scheduling plan of the pull sweep over one CSC.
// Row classes by in-degree d (rows of a CSC are destinations):
//   d <= short_max             one lane per row, inside 256-row tiles covering all rows
//   short_max < d <= group_max a group of 4 / 8 / 16 / 32 lanes per row (by d)
//   group_max < d < hub_min    one wavefront (64 lanes) per row
//   d >= hub_min               hub: hub_segment nonzeros per workgroup, partials combined
//                              in a second kernel in segment order
constexpr int kGroupClasses = 5;  // 4, 8, 16, 32, 64 lanes
constexpr uint32_t kSlices = 8;   // source slices of the sliced gather: one per XCD of gfx950
// A source's slice is its 128-byte line of contributions (32 floats) mod kSlices, so that
// every L2 line an XCD fetches for its slice holds only that slice's sources.
constexpr int kSliceShift = 5;
__host__ __device__ inline uint32_t slice_of(uint32_t u) { return (u >> kSliceShift) & (kSlices - 1); }
// Binned sweep layout (propagation blocking) of the active CSC.  Rows form bins of
// 2^bin_bits, sources form chunks of 2^kChunkBits; every CSC entry (row v <- source u)
// belongs to tile (bin v >> bin_bits, chunk u >> kChunkBits).  Entries are stored tile
// by tile, bin-major, each tile ordered by source, then row.  The entries of one bin that
// share a source form a pair (bin, source): the value of a pair is written and read once,
// however many rows of the bin the source feeds.  Per entry the layout keeps the row's
// low bits and a bit that marks the first entry of a pair; per pair, the source's low
// bits.  The chunk pass writes contrib[u] into every pair of a chunk (from the chunk
// staged in LDS), the bin pass streams each bin's entries, finds each entry's pair by
// counting pair starts, and adds the pair's value into LDS row accumulators.
constexpr int kChunkBits = 15;  // sources per chunk: 32768 floats = 128 KiB of LDS
constexpr uint32_t kBinMaxSegs = 2048;  // tiles of one chunk-pass work item
constexpr uint32_t kBinBlock = 512;     // entries of one bin-pass wave block (8 per lane); pair index kept per block
constexpr uint32_t kPairStart = 0x8000u;  // entry flag: first entry of its pair
// Hot tiles: a bin's entries whose source is below kHotSrc (the highest out-degrees in a
// degree-ordered numbering) skip the value array: the bin pass reads them from the
// contributions staged in its LDS next to the row accumulators (15.5 KiB; more hot sources
// measured slower at RMAT-25: staging them costs more than their pairs).
constexpr uint32_t kHotSrc = 3968;
// One graph's per-sweep L1 change is added by thousands of workgroups; they spread their
// atomics over kBinAccLanes counters kBinAccPad words apart (one address would serialise
// them), which every workgroup sums when it checks convergence.
constexpr uint32_t kBinAccLanes = 64;
constexpr uint32_t kBinAccPad = 8;
constexpr uint32_t kBinAccStride = kBinAccLanes * kBinAccPad;  // words per sweep
struct BinnedLayout {
  bool built = false;
  uint32_t n_bins = 0, n_chunks = 0, n_items2 = 0, n_items1 = 0, n_items3 = 0, n_segs = 0, n_slots = 0;
  uint32_t n_items2_live = 0;  // bin-pass items of bins with entries (the others come after them)
  int bin_bits = 13;           // rows per bin: 2^13 (64 KiB of row sums, two bin-pass workgroups per CU) or 2^14
  uint32_t n_hot = 0;            // sources [0, n_hot) form each bin's hot tile (0: no hot tiles)
  uint64_t nnz = 0, padded = 0;  // entries, and the length of the per-entry array (pads included)
  uint64_t pairs = 0;            // length of the per-pair arrays (tiles padded to 8 pairs)
  DevBuf<uint16_t> brow;         // per entry: row & (2^bin_bits - 1), | kPairStart on a pair's first entry
  DevBuf<uint16_t> bsrc;         // per pair: source & (2^kChunkBits - 1), or the source in a hot tile
  DevBuf<uint32_t> bpid;         // per kBinBlock entries: pair of the block's first entry
  DevBuf<uint32_t> item_he;      // per bin-pass item: end of its part of the bin's hot tile (its first tile)
  DevBuf<uint32_t> item_hd;      // per bin-pass item: its bin's hot pairs start at pair item_hd + their index in hsrc
  DevBuf<uint32_t> hsrc;         // the hot tiles' pair sources (as in bsrc), 32-bit, bin after bin
  DevBuf<uint4> items2;          // bin pass: {bin, first entry, end entry, partial slot (or none)}
  DevBuf<uint4> items3;          // combine pass units: {split bin, first slot, slots, first row | log2 lanes << 16}
  DevBuf<uint4> items1;          // chunk pass: {chunk, first tile, end tile, groups of 8 pairs}
  DevBuf<uint32_t> seg_start;    // non-empty tiles of each chunk: first pair
  DevBuf<uint32_t> seg_flat;     // and offset inside their chunk-pass item
  DevBuf<float> bw;              // weighted layouts: per entry its edge weight (pads 0)
};

struct Plan {
  nwg_plan_params params{};
  bool built = false;
  uint32_t n_tiles = 0;
  uint32_t cls_count[kGroupClasses] = {};
  uint32_t cls_off[kGroupClasses + 1] = {};  // offsets of each class in group_rows
  DevBuf<uint32_t> group_rows;               // row ids, ascending inside each class
  uint32_t n_hub_rows = 0, n_hub_segs = 0;
  DevBuf<uint32_t> hub_rows;     // ascending row ids
  DevBuf<uint32_t> hub_seg_ptr;  // n_hub_rows + 1: first segment of each hub row
  DevBuf<uint32_t> hub_seg_row;  // segment -> index into hub_rows
  // Source-sliced rows (slice_min <= in-degree < hub_min): their in-edges are regrouped
  // by source slice (slice_of) inside the row, in `scol` (a full copy of the column
  // ids, other rows unchanged); sseg[k*(kSlices+1) + x] is where slice x of sliced row k
  // starts.
  uint32_t n_srows = 0;
  uint32_t scls_off[kGroupClasses + 1] = {};  // sliced rows per lane class (by in-degree / kSlices)
  DevBuf<uint32_t> srows;  // sliced row ids, ascending
  DevBuf<uint32_t> sseg;   // n_srows x (kSlices + 1) segment starts
  DevBuf<uint32_t> slist;  // indices into srows grouped by lane class
  DevBuf<uint32_t> scol;
  BinnedLayout bl;  // built when params.binned
};

// Device parameters of one binned sweep.
struct BinParams {
  const uint16_t* brow;
  const uint16_t* bsrc;
  const uint32_t* bpid;
  float* vals;  // one value per pair
  const uint4* items2;
  const uint4* items1;
  const uint4* items3;
  const uint32_t* seg_start;
  const uint32_t* seg_flat;
  unsigned long long* part;  // n_slots << bin_bits partial row sums (fixed point) of the split bins' items
  const uint32_t* deg;
  float* x;
  const float* cin;   // contributions of the previous sweep (read by the chunk pass and the hot tiles)
  float* cout;        // contributions written by the row updates (may equal cin without hot tiles)
  const uint32_t* item_he;  // per bin-pass item: end of its hot entries
  const uint32_t* item_hd;  // per bin-pass item: pair - item_hd indexes hsrc in its hot tile
  const uint32_t* hsrc;     // hot tiles' pair sources, 32-bit
  uint32_t n_hot;
  const float* basev;
  float base;
  float damp;
  uint32_t n;
  uint32_t n_items2, n_items1, n_items3;
  uint32_t n_items2_live;  // one graph, sweeps from the third: only these items run
  int bin_bits;            // the layout's rows per bin (log2)
  unsigned long long* acc;    // one graph: iters x kBinAccStride (lanes); several graphs: iters x ngraphs
  uint32_t iter;
  double threshold;
  const uint32_t* row_graph;  // several graphs in one union: graph of every row (null: one graph)
  uint32_t ngraphs;
  uint32_t hints;  // non-temporal hints (the plan's stream_hints): 0 none, 1 loads, 2 loads and stores
  // Weighted sweep (bw set): entry value = weight * pair value, scaled by 2^(*qshift) into
  // [-1/2, 1/2] and summed in signed fixed point; cin holds the previous iterate and x the
  // new one (double-buffered), basev the per-row constant term; the L1 change of a workgroup
  // goes to wpart[slot] (bin items, then combine units) and a fold stores werr[iter]; a sweep
  // exits when werr[iter - 1] < threshold * (*thr_scale, or 1).
  const float* bw;
  const int* qshift;  // device word: set per solve, read by the captured sweeps
  double* wpart;
  double* werr;
  const double* thr_scale;
};

// A solve's sweeps run as captured HIP graphs of kSegFirst sweeps, then kSegNext at a time,
// with the host checking the last sweep's L1 change between segments: sweeps after
// convergence still launch (and exit at once), and at a few microseconds per launch, 80 of
// them cost more than the check.
constexpr uint64_t kSegFirst = 20;
constexpr uint64_t kSegNext = 8;
inline uint64_t segment_end(uint64_t start, uint64_t max_iters) {
  const uint64_t len = start == 0 ? kSegFirst : kSegNext;
  return start + len < max_iters ? start + len : max_iters;
}

// Identity of a captured sweep loop: re-capture when any of it changes.
struct CaptureKey {
  static constexpr int kPtrs = 16;
  const void* ptrs[kPtrs] = {};
  uint64_t plan_version = 0;
  uint64_t iters = 0;
  float damping = 0.f, threshold = 0.f;
  uint32_t extra = 0;
  bool operator==(const CaptureKey& o) const {
    for (int i = 0; i < kPtrs; ++i)
      if (ptrs[i] != o.ptrs[i]) return false;
    return plan_version == o.plan_version && iters == o.iters && damping == o.damping &&
           threshold == o.threshold && extra == o.extra;
  }
};

// A captured and instantiated HIP graph.
struct Captured {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  CaptureKey key;
  bool valid = false;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
  ~Captured() { reset(); }
};

// Cached buffers of the weighted forward / adjoint (wpr.hip).
struct WState {
  DevBuf<float> y0, y1, base, lambda, t;
  DevBuf<double> part, err;  // per-workgroup L1 partials; per-sweep L1 change
  DevBuf<double> scale;      // relative threshold: ||b||_1 / (1 - d)
  DevBuf<uint32_t> scale_done;
  DevBuf<uint32_t> done;     // per sweep: workgroups finished
  DevBuf<uint32_t> row_of;  // CSC position -> row
  // binned sweep: weighted layouts of the CSC (forward) and the CSR (adjoint), built on first
  // use and dropped when the weights change; pair values, split-bin partial rows, L1 partials
  std::unique_ptr<BinnedLayout> bl_csc, bl_csr;
  DevBuf<float> bvals;
  DevBuf<unsigned long long> bpart;
  DevBuf<double> wpart, stats;
  // captured solve loops (init + max_iters sweeps) of the forward and the adjoint; a call's
  // varying inputs go to fixed buffers first: the adjoint's grad_x to gin, the binned sweep's
  // scale exponent to qs
  std::vector<std::unique_ptr<Captured>> cap_fwd, cap_adj;  // one per segment of sweeps
  DevBuf<float> gin;
  DevBuf<int> qs;
  int qs_host = 0;
};

// Degree-ordered copy of a CSC (relabel.hip): new id r is original vertex order[r].
struct Relabeled {
  DevBuf<uint32_t> order, rank;  // new -> original, original -> new
  DevBuf<uint32_t> csc_ptr, csc_idx, outdeg;
};

// The CSC arrays a sweep runs on (original or relabelled).
struct ActiveCsc {
  const uint32_t* ptr;
  const uint32_t* idx;
  const uint32_t* deg;
  const uint32_t* order;  // null when not relabelled
};

// Cached state of PageRank solves on one graph (buffers + captured HIP graph).
struct PrState {
  DevBuf<float> x, contrib0, contrib1, base;
  DevBuf<unsigned long long> acc;  // per-sweep L1 change, fixed point (see device_utils.hpp)
  DevBuf<float> hub_part;          // per hub segment partial sums
  DevBuf<double> blk_err;          // per-workgroup L1 partials of one sweep
  DevBuf<float> spart;             // kSlices x n_srows partial sums of the sliced rows
  DevBuf<float> bvals;             // binned sweep: one value per (bin, source) pair
  DevBuf<unsigned long long> bpart;  // binned sweep: partial rows of split bins
  DevBuf<unsigned long long> bacc;   // binned sweep: per-sweep L1 change, kBinAccLanes lanes
  DevBuf<double> tcheck;             // teleport check: sum, invalid-entry flag
  std::vector<std::unique_ptr<Captured>> caps;  // the solve's segments of sweeps
};

// Reusable state of the batched personalised solve: the heavy-row segments of the active
// CSC (rebuilt when the plan or the k-tile changes) and grow-only buffers.
struct PprState {
  uint64_t seg_version = 0;  // plan_version the segments were built for (0 = none)
  uint32_t seg_tile = 0, seg_slices = 0;
  uint32_t n_heavy = 0, n_segs = 0;
  uint32_t row_end = 0;  // one past the last row with in-edges (active numbering)
  DevBuf<uint32_t> heavy_rows, heavy_seg_ptr, heavy_seg_row;
  // heavy rows' in-edges regrouped by (row, source slice); segments [start, end) of it,
  // numbered row by row; the segment ids of each slice, slice after slice
  DevBuf<uint32_t> hcol, seg_start, seg_end, slice_segs;
  std::vector<uint32_t> slice_off;
  DevBuf<uint32_t> lrp, lcol;  // light rows' in-edges, compacted (heavy rows empty)
  DevBuf<uint32_t> degflag, seed_row, seed_vec, ids, ids_new;
  DevBuf<uint32_t> its;  // sweeps each vector executed (read by the final-scores kernel)
  DevBuf<float> seed_base, c0, c1, hub_part;
  DevBuf<double> blk_err;
  DevBuf<unsigned long long> acc;
};

}  // namespace nwg

struct nwg_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t capture_stream = nullptr;  // used only to record HIP graphs
  std::string last_error;
  std::mutex mu;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct nwg_graph {
  nwg_ctx* ctx = nullptr;
  int device = 0;
  uint32_t n = 0;
  uint64_t nnz = 0;
  nwg::DevBuf<uint32_t> csc_ptr, csc_idx, outdeg;
  // optional transposed structure for the adjoint
  bool has_csr = false;
  nwg::DevBuf<uint32_t> csr_ptr, csr_idx, csr_to_csc;
  // optional weights (CSC order, normalised) and raw weights
  bool has_weights = false;
  bool w_normalized = false;
  bool w_nonneg = false;  // every raw weight finite and >= 0
  nwg::DevBuf<float> w_csc, w_raw, w_csr, w_rowsum;
  // work items of the weighted sweeps over the CSC (forward) and the CSR (adjoint)
  nwg::DevBuf<uint32_t> w_items_csc_r0, w_items_csc_r1, w_items_csr_r0, w_items_csr_r1;
  uint32_t w_items_csc_n = 0, w_items_csr_n = 0;
  nwg::Plan plan;
  uint64_t plan_version = 1;
  std::unique_ptr<nwg::Relabeled> rl;  // built when a plan asks for relabel=1
  nwg::ActiveCsc active(bool relabel) const {
    if (relabel && rl) return {rl->csc_ptr.p, rl->csc_idx.p, rl->outdeg.p, rl->order.p};
    return {csc_ptr.p, csc_idx.p, outdeg.p, nullptr};
  }
  std::unique_ptr<nwg::PrState> pr;
  std::unique_ptr<nwg::WState> wst;
  std::unique_ptr<nwg::PprState> ppr;
};

namespace nwg {

// ---- builder (build.hip) ----
// Sort-based CSC/CSR construction on the device.  src/dst are device arrays of ne
// raw edges (they are consumed: contents are undefined afterwards).
void build_csc_device(hipStream_t s, uint32_t n, uint64_t ne, uint32_t* src, uint32_t* dst,
                      bool clean, nwg_graph* g);
void build_csr_device(hipStream_t s, nwg_graph* g);
int key_bits_for(uint32_t n);  // bits of a vertex id field in packed sort keys
// row pointers + entries from keys (row << b | entry) sorted ascending
void sorted_keys_to_csx(hipStream_t s, const uint64_t* keys, uint64_t m, uint32_t n, int b, uint32_t* ptr,
                        uint32_t* ent);
// throws NWG_EINVAL unless ptr is monotone from 0 to nnz and every entry of idx is below n;
// flag: one device word of scratch, or null to allocate it
void validate_device_csc(hipStream_t s, uint32_t n, uint64_t nnz, const uint32_t* ptr, const uint32_t* idx,
                         uint32_t* flag = nullptr);
// out-degree histogram of col (ids < n); may_allocate = false keeps to the atomic histogram,
// which needs no device memory beyond deg
void out_degrees_device(hipStream_t s, uint32_t n, uint64_t nnz, const uint32_t* col, uint32_t* deg,
                        bool may_allocate = true);
// throws NWG_EINVAL when some outdeg[u] is below the number of entries u has in col;
// counted (n words) and flag (one word): device scratch, or null to allocate them
void check_outdeg_device(hipStream_t s, uint32_t n, uint64_t nnz, const uint32_t* col, const uint32_t* outdeg,
                         uint32_t* counted = nullptr, uint32_t* flag = nullptr);

// ---- primitives (scan.hip / sort.hip) ----
// exclusive scan of n values; out has n+1 entries (out[n] = total).
void exclusive_scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t n);
// stable LSD radix sort of 64-bit keys over bits [0, key_bits), optional 32-bit payload.
// Sorted result ends in keys/vals (tmp buffers must have n entries).
void radix_sort_pairs(hipStream_t s, uint64_t* keys, uint64_t* keys_tmp, uint32_t* vals,
                      uint32_t* vals_tmp, uint64_t n, int key_bits);
// stable LSD radix sort of 32-bit keys by their bits [first_bit, end_bit) (the others are
// carried along unsorted); result in keys (keys_tmp: n entries of scratch).
void radix_sort_keys_u32(hipStream_t s, uint32_t* keys, uint32_t* keys_tmp, uint64_t n, int first_bit, int end_bit);

// ---- generators (gen.hip) ----
void rmat_edges_device(hipStream_t s, uint32_t scale, uint32_t a, uint32_t b, uint32_t c, uint64_t seed,
                       uint64_t first, uint64_t count, uint32_t* src, uint32_t* dst);
void grid_edges_device(hipStream_t s, uint32_t width, uint32_t keep, uint32_t lng, uint64_t seed,
                       uint32_t* src, uint32_t* dst);
void edge_weights_device(hipStream_t s, uint32_t n, const uint32_t* csc_ptr, const uint32_t* csc_idx,
                         uint64_t seed, float* w);
// streaming copy for bandwidth / counter calibration; returns the kernel time in ms
float stream_copy_device(hipStream_t s, hipEvent_t e0, hipEvent_t e1, void* dst, const void* src, uint64_t bytes);
// bandwidth probe (kind 0 copy, 1 read only, 2 write only; flags bit 0 nontemporal loads,
// bit 1 nontemporal stores; wg_per_cu workgroups of 256 lanes per CU); kernel time in ms
float stream_probe_device(hipStream_t s, hipEvent_t e0, hipEvent_t e1, uint32_t kind, uint32_t flags,
                          uint32_t wg_per_cu, void* dst, const void* src, uint64_t bytes);

// ---- PageRank (pagerank.hip) ----
// (re)build the sweep plan of g with the given parameters; g is unchanged on failure
void plan_build(hipStream_t s, nwg_graph* g, const nwg_plan_params& pp);
void relabel_build(hipStream_t s, nwg_graph* g);
// rows with in-degree >= min_deg (ascending) and their split into `seg`-nonzero segments
void build_segments(hipStream_t s, const uint32_t* rp, uint32_t n, uint32_t min_deg, uint32_t seg,
                    DevBuf<uint32_t>& rows, DevBuf<uint32_t>& seg_ptr, DevBuf<uint32_t>& seg_row, uint32_t* nrows,
                    uint32_t* nsegs);
nwg_plan_params default_plan_params();
// the plan a PageRank solve of g builds on first use (binned sweep and degree ordering by size)
nwg_plan_params default_plan_for(const nwg_graph* g);
// binned sweep (binned.hip)
// w (device, CSC order of a, nullable): also lay out the entries' weights (L.bw)
// bin_bits: rows per bin, log2 (13 or 14; 0 = by size)
void binned_build(hipStream_t s, const ActiveCsc& a, uint32_t n, uint64_t nnz, BinnedLayout& L, bool hot,
                  const float* w = nullptr, int bin_bits = 0);
void binned_enqueue_sweep(hipStream_t s, const BinParams& P);
// one weighted sweep (P.bw set): chunk, bin and combine passes, then the fold of the L1 partials
void binned_enqueue_weighted_sweep(hipStream_t s, const BinParams& P);
// lane counters of iters sweeps cleared / summed into acc[0 .. iters)
void binned_clear_residuals(hipStream_t s, unsigned long long* lanes, uint64_t iters);
void binned_fold_residuals(hipStream_t s, const unsigned long long* lanes, unsigned long long* acc, uint64_t iters);
// Schedules of pagerank_run:
//   kJacobi   page_rank_v10: double-buffered contributions (pull sweep or binned sweep)
//   kInPlace  page_rank (v11): one contribution array updated during the sweep
//             (page_rank.hpp:135), pull kernels only
//   kDelta    page_rank_v13 (experimental/algorithms/page_rank.hpp:560-625): residuals
//             pushed as deltas, pulled by the pull kernels; un-normalised scores, no teleport
enum class PrSchedule : uint32_t { kJacobi = 0, kInPlace = 1, kDelta = 2 };
void pagerank_run(nwg_ctx* ctx, nwg_graph* g, const float* teleport_dev, float damping, float threshold,
                  uint64_t max_iters, nwg_pr_info* info, float* scores_dev,
                  PrSchedule sched = PrSchedule::kJacobi);

// ---- allocation-free one-call solve (noalloc.hip) ----
// looks up the device's static scratch array once (called by nwg_create, device current)
void noalloc_setup(int device);
uint64_t noalloc_max_vertices();
uint64_t noalloc_max_iters();
void pagerank_noalloc(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr, const uint32_t* csc_idx,
                      const uint32_t* outdeg, float* scores, float damping, float threshold, uint64_t max_iters,
                      nwg_pr_info* info);

// ---- batched personalised PageRank (ppr.hip) ----
void ppr_run(nwg_ctx* ctx, nwg_graph* g, uint32_t k, const std::vector<uint32_t>& seed_ptr,
             const std::vector<uint32_t>& seeds, float damping, float threshold, uint64_t max_iters, uint32_t tile,
             float* scores_dev, uint64_t* iterations, nwg_pr_info* info);

// ---- many graphs (multi.hip) ----
nwg_multigraph* multigraph_create(nwg_ctx* ctx, uint32_t ngraphs, nwg_graph* const* graphs);
void multigraph_run(nwg_ctx* ctx, nwg_multigraph* mg, float damping, float threshold, uint64_t max_iters,
                    float* scores_dev, uint64_t* iterations, double* final_error, nwg_pr_info* info);
uint64_t multigraph_vertices(const nwg_multigraph* mg);
void multigraph_set_binned(nwg_ctx* ctx, nwg_multigraph* mg, bool binned);

// ---- weighted PageRank and adjoint (wpr.hip) ----
void build_items(hipStream_t s, const uint32_t* rp_dev, uint32_t n, DevBuf<uint32_t>& r0, DevBuf<uint32_t>& r1,
                 uint32_t* count);
void weights_set(hipStream_t s, nwg_graph* g, const float* w_dev, bool normalize);
void wpr_forward(nwg_ctx* ctx, nwg_graph* g, const float* teleport_dev, float* x_dev, float damping, float threshold,
                 uint64_t max_iters, nwg_pr_info* info);
void wpr_backward(nwg_ctx* ctx, nwg_graph* g, const float* x_dev, const float* gx_dev, float* gw_dev, float* gv_dev,
                  float damping, float threshold, uint64_t max_iters, bool raw, nwg_pr_info* info);

// ---- autotuner (autotune.cpp) ----
nwg_plan_params autotune_plan(nwg_ctx* ctx, nwg_graph* g);
// time `iters` sweeps of the plain kernel under the graph's current plan (ms per sweep)
double time_plan_sweeps(nwg_ctx* ctx, nwg_graph* g, uint64_t iters);

}  // namespace nwg
