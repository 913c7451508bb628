// dropin_run.cpp -- a caller of the C++ drop-in, written the way the reference's own drivers
// call nw::graph::page_rank (bench/gapbs/pr.cpp:137-139): a CSC laid out like adjacency<1>,
// degrees counted as build_degrees counts them (bench/gapbs/common.hpp:110-121), a pre-sized
// score vector.  It also declares a stand-in for the reference's class template adjacency<idx>
// (only the members the drop-in reads) and a generic page_rank template in the same namespace,
// to check that the drop-in's exact-match overload for adjacency<1> is what the call binds to.
//
//   dropin_run ADJACENCY_FILE OUT_PREFIX
// reads a CSC written by nwg_write_adjacency and writes OUT_PREFIX.v11 / .v10 (float32 scores
// from page_rank and page_rank_v10 on csc_graph) and OUT_PREFIX.adj (page_rank on the
// adjacency<1> stand-in); prints the per-sweep trace and "template calls: N".
#include <nwgraph_hip.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

namespace nw::graph {
template <int idx>
struct adjacency {
  using vertex_id_type = uint32_t;
  std::vector<uint32_t> indices_;
  std::tuple<std::vector<uint32_t>> to_be_indexed_;
  size_t size() const { return indices_.size() - 1; }
};
inline int template_calls = 0;
// a generic template with the reference's signature: the overload for adjacency<1> must win
template <class Graph, typename Real>
void page_rank(const Graph&, const std::vector<typename Graph::vertex_id_type>&, std::vector<Real>&, Real, Real,
               size_t, size_t) {
  ++template_calls;
}
}  // namespace nw::graph

#define NW_GRAPH_ADJACENCY_HPP
#define NW_GRAPH_PAGE_RANK_HPP
#include <nwgraph_hip/page_rank.hpp>

static void write_floats(const std::string& path, const std::vector<float>& v) {
  std::ofstream out(path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(float)));
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: dropin_run ADJACENCY_FILE OUT_PREFIX\n");
    return 2;
  }
  uint32_t n = 0;
  uint64_t nnz = 0;
  uint32_t *ptr = nullptr, *idx = nullptr;
  if (nwg_read_adjacency(argv[1], &n, &nnz, &ptr, &idx) != NWG_OK) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 1;
  }
  nw::graph::hip::csc_graph g;
  g.indices_.assign(ptr, ptr + n + 1);
  std::get<0>(g.to_be_indexed_).assign(idx, idx + nnz);
  nwg_free(ptr);
  nwg_free(idx);
  std::vector<uint32_t> degrees(n, 0);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t p = g.indices_[i]; p < g.indices_[i + 1]; ++p) ++degrees[std::get<0>(g.to_be_indexed_)[p]];

  std::vector<float> rankings(n, 0.0f);
  nw::graph::hip::page_rank(g, degrees, rankings, 0.85f, 1e-4f, 20, 1);
  write_floats(std::string(argv[2]) + ".v11", rankings);
  std::vector<float> r10(n, 0.0f);
  nw::graph::hip::page_rank_v10(g, degrees, r10, 0.85f, 1e-4f, 20, 1);
  write_floats(std::string(argv[2]) + ".v10", r10);

  nw::graph::adjacency<1> a;
  a.indices_ = g.indices_;
  a.to_be_indexed_ = g.to_be_indexed_;
  std::vector<float> ra(n, 0.0f);
  nw::graph::page_rank(a, degrees, ra, 0.85f, 1e-4f, 20, 1);  // binds to the drop-in's overload
  write_floats(std::string(argv[2]) + ".adj", ra);

  // misuse reaches the caller as an exception, not a crash
  std::vector<float> wrong(n + 1);
  try {
    nw::graph::hip::page_rank(g, degrees, wrong, 0.85f, 1e-4f, 20, 1);
    return 3;
  } catch (const std::invalid_argument&) {
  }
  std::printf("template calls: %d\n", nw::graph::template_calls);
  return 0;
}
