// nwgraph_hip/page_rank.hpp -- C++ drop-in for nw::graph::page_rank on MI355X.
//
// Replaces (pnnl/NWGraph) include/nwgraph/algorithms/page_rank.hpp:89-148:
//
//   template <adjacency_list_graph Graph, typename Real>
//   void page_rank(const Graph& graph, const std::vector<vertex_id_type>& degrees,
//                  std::vector<Real>& page_rank, Real damping_factor, Real threshold,
//                  size_t max_iters, size_t num_threads);
//
// with the same parameters, the same in-place output, the same per-sweep trace on
// stdout (pagerank::trace, page_rank.hpp:42-52) and the same init timers
// (life_timer, util/timer.hpp:120-133).  The work runs in libnwgraph_hip.so through the
// C ABI of <nwgraph_hip.h>: the CSC is uploaded, the sweep loop runs on the GPU (Jacobi
// schedule of page_rank_v10, convergence checked on the device), and the scores are copied
// back.  With the trace on (trace_enabled(), the default, as the reference always prints
// it) the inits and sweeps are launched one by one and each is timed with HIP events;
// with it off the loop is one captured HIP graph.  Errors become exceptions:
// std::invalid_argument, std::bad_alloc or nw::graph::hip::error.
//
// Accepted graphs: anything laid out like adjacency<1> -- `indices_` (N+1 row offsets)
// and `std::get<0>(to_be_indexed_)` (source id of every in-edge), both 32-bit -- such as
// the reference's adjacency<1> and nw::graph::hip::csc_graph below.
//
// Including this header after the reference's <nwgraph/adjacency.hpp> adds a
// non-template overload nw::graph::page_rank(const adjacency<1>&, ...) for float scores,
// which overload resolution prefers to the reference's template, so the reference's
// drivers (bench/gapbs/pr.cpp:137-139) bind to the GPU path unchanged.
#pragma once

#include <nwgraph_hip.h>

#include <cstdint>
#include <iomanip>
#include <iostream>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace nw {
namespace graph {
namespace hip {

class error : public std::runtime_error {
public:
  error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const noexcept { return code_; }

private:
  int code_;
};

inline void check(nwg_ctx* ctx, int rc, const char* what) {
  if (rc == NWG_OK) return;
  std::string msg = std::string(what) + ": " + (ctx ? nwg_last_error(ctx) : "no context");
  if (rc == NWG_ENOMEM) throw std::bad_alloc();
  if (rc == NWG_EINVAL) throw std::invalid_argument(msg);
  throw error(rc, msg);
}

// One device context (stream + reusable device buffers).  Not shareable between
// threads; thread_default() gives each host thread its own on device 0.
class context {
public:
  explicit context(int device = 0) {
    nwg_ctx* c = nullptr;
    check(nullptr, nwg_create(&c, device), "nwg_create");
    ctx_ = c;
  }
  ~context() { nwg_destroy(ctx_); }
  context(const context&) = delete;
  context& operator=(const context&) = delete;
  nwg_ctx* get() const noexcept { return ctx_; }
  static context& thread_default() {
    thread_local context c(0);
    return c;
  }

private:
  nwg_ctx* ctx_ = nullptr;
};

// The CSC container with adjacency<1>'s layout (compressed.hpp:64-102).
struct csc_graph {
  using vertex_id_type = uint32_t;
  std::vector<uint32_t> indices_;
  std::tuple<std::vector<uint32_t>> to_be_indexed_;
  size_t size() const { return indices_.empty() ? 0 : indices_.size() - 1; }
  size_t num_edges() const { return std::get<0>(to_be_indexed_).size(); }
};

// Print the reference's per-sweep trace (page_rank.hpp:42-52) and init timers?
inline bool& trace_enabled() {
  static bool on = true;
  return on;
}

namespace detail {
inline void trace_row(const std::string& a, const std::string& b, const std::string& c, const std::string& d) {
  std::cout << std::setw(5) << std::left << a << std::setw(20) << std::left << b << std::setw(20) << std::left << c
            << std::setw(20) << std::left << d << "\n";
}
inline void trace_iter(uint64_t iter, double error, double seconds) {
  auto p = std::cout.precision(10);
  auto f = std::cout.setf(std::ios::fixed, std::ios::floatfield);
  std::cout << std::setw(5) << std::left << iter << std::setw(20) << std::left << error << std::setw(20) << std::left
            << seconds << std::setw(20) << std::left << 0 << "\n";
  std::cout.precision(p);
  std::cout.setf(f, std::ios::floatfield);
}
}  // namespace detail

// The reference's output around a solve (page_rank.hpp:99-142): the init timers, the trace
// header and one line per sweep with its L1 change and its time in seconds.  With
// NWG_INFO_KERNEL_TIMING the times are the device's: each init launch and each sweep.
inline void print_trace(const nwg_pr_info& info, const std::vector<double>& errors,
                        const std::vector<double>& sweep_ms) {
  std::cout << "# [ init page rank ]: " << info.init_ms[0] << " ms" << std::endl;
  detail::trace_row("iter", "error", "time", "outgoing");
  std::cout << "# [ init contrib ]: " << info.init_ms[1] << " ms" << std::endl;
  for (uint64_t k = 0; k < info.iterations && k < errors.size(); ++k)
    detail::trace_iter(k, errors[k], k < sweep_ms.size() ? sweep_ms[k] * 1e-3 : 0.0);
}

// Core entry on raw CSC arrays (host memory).  Returns the sweeps executed.
inline uint64_t page_rank_csc(uint32_t n, uint64_t nnz, const uint32_t* indices, const uint32_t* sources,
                              const uint32_t* degrees, float* scores, float damping_factor, float threshold,
                              size_t max_iters, nwg_ctx* ctx = nullptr) {
  if (!ctx) ctx = context::thread_default().get();
  std::vector<double> trace(max_iters ? max_iters : 1), times(max_iters ? max_iters : 1);
  nwg_pr_info info{};
  info.error_trace = trace.data();
  info.error_trace_len = trace.size();
  if (trace_enabled()) {
    // sweeps launched one by one, each timed, as the reference times each sweep
    info.flags = NWG_INFO_KERNEL_TIMING;
    info.time_trace = times.data();
    info.time_trace_len = times.size();
  }
  check(ctx, nwg_pagerank_f32(ctx, n, nnz, indices, sources, degrees, scores, damping_factor, threshold, max_iters,
                              &info),
        "nwg_pagerank_f32");
  if (trace_enabled()) print_trace(info, trace, times);
  return info.iterations;
}

template <class Graph>
concept csc_layout = requires(const Graph& g) {
  { g.indices_.data() };
  { std::get<0>(g.to_be_indexed_).data() };
};

// Same signature as the reference template (page_rank.hpp:89-91).  Real must be float
// (the device path computes in FP32); ids and offsets must be 32-bit.
template <csc_layout Graph, typename Real>
void page_rank(const Graph& graph, const std::vector<typename Graph::vertex_id_type>& degrees,
               std::vector<Real>& page_rank, Real damping_factor, Real threshold, size_t max_iters,
               size_t /*num_threads: the reference ignores it too*/) {
  static_assert(std::is_same_v<Real, float>, "the MI355X path computes float scores");
  static_assert(sizeof(typename Graph::vertex_id_type) == 4, "vertex ids must be 32-bit");
  static_assert(sizeof(graph.indices_[0]) == 4, "row offsets must be 32-bit");
  const auto& src = std::get<0>(graph.to_be_indexed_);
  const size_t n = graph.indices_.size() - 1;
  if (page_rank.size() != n) throw std::invalid_argument("page_rank must be pre-sized to graph.size()");
  if (!degrees.empty() && degrees.size() != n) throw std::invalid_argument("degrees must have graph.size() entries");
  page_rank_csc((uint32_t)n, src.size(), reinterpret_cast<const uint32_t*>(graph.indices_.data()),
                reinterpret_cast<const uint32_t*>(src.data()),
                degrees.empty() ? nullptr : reinterpret_cast<const uint32_t*>(degrees.data()), page_rank.data(),
                damping_factor, threshold, max_iters);
}

// Same signature as the reference's Jacobi twin page_rank_v10
// (experimental/algorithms/page_rank.hpp:377-432): the device path of page_rank above is
// this schedule already.
template <csc_layout Graph, typename Real>
void page_rank_v10(const Graph& graph, const std::vector<typename Graph::vertex_id_type>& degrees,
                   std::vector<Real>& page_rank, Real damping_factor, Real threshold, size_t max_iters,
                   size_t num_threads) {
  hip::page_rank(graph, degrees, page_rank, damping_factor, threshold, max_iters, num_threads);
}

// Same signature as the reference's delta PageRank page_rank_v13
// (experimental/algorithms/page_rank.hpp:560-625): un-normalised scores (about N in total).
// The graph lives for this call: it runs on the pull kernels (nwg_graph_pagerank_delta).
template <csc_layout Graph, typename Real>
void page_rank_v13(const Graph& graph, const std::vector<typename Graph::vertex_id_type>& degrees,
                   std::vector<Real>& page_rank, Real damping_factor, Real threshold, size_t max_iters,
                   size_t /*num_threads*/) {
  static_assert(std::is_same_v<Real, float>, "the MI355X path computes float scores");
  static_assert(sizeof(typename Graph::vertex_id_type) == 4, "vertex ids must be 32-bit");
  static_assert(sizeof(graph.indices_[0]) == 4, "row offsets must be 32-bit");
  const auto& src = std::get<0>(graph.to_be_indexed_);
  const size_t n = graph.indices_.size() - 1;
  if (page_rank.size() != n) throw std::invalid_argument("page_rank must be pre-sized to graph.size()");
  if (!degrees.empty() && degrees.size() != n) throw std::invalid_argument("degrees must have graph.size() entries");
  nwg_ctx* ctx = context::thread_default().get();
  nwg_graph* g = nullptr;
  check(ctx,
        nwg_graph_create_csc(ctx, (uint32_t)n, src.size(), reinterpret_cast<const uint32_t*>(graph.indices_.data()),
                             reinterpret_cast<const uint32_t*>(src.data()),
                             degrees.empty() ? nullptr : reinterpret_cast<const uint32_t*>(degrees.data()),
                             NWG_MEM_HOST, &g),
        "nwg_graph_create_csc");
  std::vector<double> times(max_iters ? max_iters : 1);
  nwg_pr_info info{};
  if (trace_enabled()) {
    info.flags = NWG_INFO_KERNEL_TIMING;
    info.time_trace = times.data();
    info.time_trace_len = times.size();
  }
  const int rc = nwg_graph_pagerank_delta(ctx, g, page_rank.data(), damping_factor, threshold, max_iters,
                                          NWG_MEM_HOST, &info);
  nwg_graph_destroy(g);
  check(ctx, rc, "nwg_graph_pagerank_delta");
  if (trace_enabled()) {
    detail::trace_row("iter", "error", "time", "outgoing");
    for (uint64_t k = 0; k < info.iterations; ++k) detail::trace_iter(k, 0.0, times[k] * 1e-3);
  }
}

}  // namespace hip

#ifdef NW_GRAPH_ADJACENCY_HPP
// Exact-match overload for the reference's CSC with float scores: preferred over the
// reference's template by overload resolution.
inline void page_rank(const adjacency<1>& graph, const std::vector<adjacency<1>::vertex_id_type>& degrees,
                      std::vector<float>& page_rank, float damping_factor, float threshold, size_t max_iters,
                      size_t num_threads) {
  hip::page_rank(graph, degrees, page_rank, damping_factor, threshold, max_iters, num_threads);
}
#endif

#ifndef NW_GRAPH_PAGE_RANK_HPP
// Without the reference's algorithm header, nw::graph::page_rank is the GPU path.
using hip::page_rank;
#endif
#ifndef NW_GRAPH_PAGE_RANK_EXPERIMENTAL_HPP
// Without the reference's experimental header, its v10 and v13 are the GPU paths too.
using hip::page_rank_v10;
using hip::page_rank_v13;
#endif

}  // namespace graph
}  // namespace nw
