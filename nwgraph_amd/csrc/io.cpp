// io.cpp -- graph input formats of the reference's PageRank driver (load_graph,
// bench/gapbs/common.hpp:82-101): Matrix Market text and the NW binary edge list; and
// the binary form of an adjacency (index_adjacency::serialize / deserialize).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nwgraph_hip.h"

namespace {

// Runs a reader or writer body so that nothing throws across the C boundary: allocation
// failures become NWG_ENOMEM, any other exception (length_error from a count read out of a
// file, stream failures) NWG_EFORMAT.
template <class F>
int io_guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return NWG_ENOMEM;
  } catch (...) {
    return NWG_EFORMAT;
  }
}

// Size of an open stream in bytes (0 when it cannot be told); counts read from a file
// are checked against it before anything is allocated for them.
uint64_t stream_bytes(std::istream& in) {
  const std::streampos here = in.tellg();
  if (here < 0) return 0;
  in.seekg(0, std::ios::end);
  const std::streampos end = in.tellg();
  in.seekg(here);
  return end < 0 ? 0 : (uint64_t)end;
}

struct Edges {
  std::vector<uint32_t> s, d;
};

// hand the edges to the caller in malloc'ed arrays (freed with nwg_free)
int export_edges(const Edges& e, uint64_t* ne, uint32_t** src, uint32_t** dst) {
  const size_t m = e.s.size();
  uint32_t* a = static_cast<uint32_t*>(std::malloc((m ? m : 1) * sizeof(uint32_t)));
  uint32_t* b = static_cast<uint32_t*>(std::malloc((m ? m : 1) * sizeof(uint32_t)));
  if (!a || !b) {
    std::free(a);
    std::free(b);
    return NWG_ENOMEM;
  }
  if (m) {
    std::memcpy(a, e.s.data(), m * sizeof(uint32_t));
    std::memcpy(b, e.d.data(), m * sizeof(uint32_t));
  }
  *ne = m;
  *src = a;
  *dst = b;
  return NWG_OK;
}

}  // namespace

extern "C" {

// read_mm<directed> (io/mmio.hpp:171-208) + mm_fill (mmio.hpp:86-103): header
// "%%MatrixMarket matrix coordinate <field> <symmetric|general>", '%' comment lines,
// "n0 n1 nnz", then nnz lines "i j [value]" (1-based).  A symmetric file also adds
// (j, i) for i != j.  n = max(n0, max id + 1) (edge_list.hpp:188-199).
static int read_mm(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst) {
  if (!path || !n || !ne || !src || !dst) return NWG_EINVAL;
  std::ifstream in(path);
  if (!in) return NWG_EINVAL;
  std::string line;
  if (!std::getline(in, line)) return NWG_EFORMAT;
  std::stringstream h(line);
  std::string hdr[5];
  for (auto& t : hdr) h >> t;
  if (hdr[0] != "%%MatrixMarket") return NWG_EFORMAT;
  bool sym;
  if (hdr[4] == "symmetric") sym = true;
  else if (hdr[4] == "general") sym = false;
  else return NWG_EFORMAT;
  while (std::getline(in, line))
    if (line.empty() || line[0] != '%') break;
  unsigned long long n0 = 0, n1 = 0, nnz = 0;
  if (std::sscanf(line.c_str(), "%llu %llu %llu", &n0, &n1, &nnz) != 3) return NWG_EFORMAT;
  // every entry line takes at least 4 bytes ("i j\n"): a header count beyond that is malformed
  const uint64_t bytes = stream_bytes(in);
  if (nnz > bytes / 4) return NWG_EFORMAT;
  Edges e;
  e.s.reserve((sym ? 2 : 1) * nnz);
  e.d.reserve((sym ? 2 : 1) * nnz);
  uint64_t N = n0;
  for (unsigned long long k = 0; k < nnz; ++k) {
    if (!std::getline(in, line)) return NWG_EFORMAT;
    unsigned long long a = 0, b = 0;
    if (std::sscanf(line.c_str(), "%llu %llu", &a, &b) != 2 || a == 0 || b == 0) return NWG_EFORMAT;
    if (a > 0xFFFFFFFFull || b > 0xFFFFFFFFull) return NWG_EFORMAT;
    e.s.push_back((uint32_t)(a - 1));
    e.d.push_back((uint32_t)(b - 1));
    if (sym && a != b) {
      e.s.push_back((uint32_t)(b - 1));
      e.d.push_back((uint32_t)(a - 1));
    }
    const uint64_t m = a > b ? a : b;
    if (m > N) N = m;
  }
  if (N > 0xFFFFFFFFull) return NWG_EFORMAT;
  *n = (uint32_t)N;
  return export_edges(e, ne, src, dst);
}

int nwg_read_mm(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst) {
  return io_guarded([&] { return read_mm(path, n, ne, src, dst); });
}

// NW binary edge list: char magic[27] "NW Graph index_edge_list", int directedness,
// size_t N, bool is_open, then struct_of_arrays: size_t nedges, size_t narrays and per
// array size_t count, size_t elsize, raw data (edge_list.hpp:237-280,
// graph_base.hpp:72-87, containers/soa.hpp:332-375).
static const char kMagic[27] = "NW Graph index_edge_list";

static int read_nw(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst) {
  if (!path || !n || !ne || !src || !dst) return NWG_EINVAL;
  std::ifstream in(path, std::ios::binary);
  if (!in) return NWG_EINVAL;
  const uint64_t bytes = stream_bytes(in);
  char magic[27];
  int32_t dirn = 0;
  uint64_t N = 0, st = 0, na = 0;
  char open_flag = 0;
  in.read(magic, 27);
  if (!in || std::memcmp(magic, kMagic, 27) != 0) return NWG_EFORMAT;
  in.read(reinterpret_cast<char*>(&dirn), 4);
  in.read(reinterpret_cast<char*>(&N), 8);
  in.read(&open_flag, 1);
  in.read(reinterpret_cast<char*>(&st), 8);
  in.read(reinterpret_cast<char*>(&na), 8);
  if (!in || na < 2 || N > 0xFFFFFFFFull) return NWG_EFORMAT;
  Edges e;
  std::vector<uint32_t>* cols[2] = {&e.s, &e.d};
  for (int a = 0; a < 2; ++a) {
    uint64_t cnt = 0, es = 0;
    in.read(reinterpret_cast<char*>(&cnt), 8);
    in.read(reinterpret_cast<char*>(&es), 8);
    if (!in || cnt != st || es != 4 || cnt > bytes / 4) return NWG_EFORMAT;
    cols[a]->resize(cnt);
    in.read(reinterpret_cast<char*>(cols[a]->data()), (std::streamsize)(cnt * 4));
    if (!in) return NWG_EFORMAT;
    // a list the reference writes has N >= max id + 1 (edge_list.hpp:188-199)
    for (uint32_t v : *cols[a])
      if (v >= N) return NWG_EFORMAT;
  }
  *n = (uint32_t)N;
  return export_edges(e, ne, src, dst);
}

int nwg_read_nw(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst) {
  return io_guarded([&] { return read_nw(path, n, ne, src, dst); });
}

static int write_nw(const char* path, uint32_t n, uint64_t ne, const uint32_t* src, const uint32_t* dst) {
  if (!path || ((!src || !dst) && ne)) return NWG_EINVAL;
  std::ofstream out(path, std::ios::binary);
  if (!out) return NWG_EINVAL;
  const int32_t dirn = 1;  // directedness::directed
  const uint64_t N = n, narr = 2, es = 4;
  const char open_flag = 0;
  out.write(kMagic, 27);
  out.write(reinterpret_cast<const char*>(&dirn), 4);
  out.write(reinterpret_cast<const char*>(&N), 8);
  out.write(&open_flag, 1);
  out.write(reinterpret_cast<const char*>(&ne), 8);
  out.write(reinterpret_cast<const char*>(&narr), 8);
  const uint32_t* cols[2] = {src, dst};
  for (int a = 0; a < 2; ++a) {
    out.write(reinterpret_cast<const char*>(&ne), 8);
    out.write(reinterpret_cast<const char*>(&es), 8);
    if (ne) out.write(reinterpret_cast<const char*>(cols[a]), (std::streamsize)(ne * 4));
  }
  return out ? NWG_OK : NWG_EFORMAT;
}

int nwg_write_nw(const char* path, uint32_t n, uint64_t ne, const uint32_t* src, const uint32_t* dst) {
  return io_guarded([&] { return write_nw(path, n, ne, src, dst); });
}

// Adjacency (index_adjacency<idx>::serialize, adjacency.hpp:176-191): the graph base
// (graph_base.hpp:72-87: size_t N, bool is_open), then the indexed struct of arrays
// (containers/compressed.hpp:377-413): char magic[34] "NW GRAPH indexed_struct_of_arrays",
// 8 bytes written from its 4-byte index_t N_ (the upper 4 are whatever follows N_ in
// memory: written as 0 here and ignored when read; an adjacency built from an edge list,
// common.hpp:103-108, holds N_ = n + 1), size_t n + 1, size_t sizeof(index_t), the n + 1
// row offsets, then the struct of arrays (soa.hpp:350-356: size_t entries, size_t arrays,
// and per array size_t count, size_t element size, raw data).  Only the attribute-free
// uint32 form (adjacency<idx>, defaults.hpp:23-24) is read and written.
static const char kAdjMagic[34] = "NW GRAPH indexed_struct_of_arrays";

static int write_adjacency(const char* path, uint32_t n, uint64_t nnz, const uint32_t* ptr, const uint32_t* idx) {
  if (!path || !ptr || (!idx && nnz)) return NWG_EINVAL;
  if ((uint64_t)ptr[n] != nnz || ptr[0] != 0) return NWG_EINVAL;
  std::ofstream out(path, std::ios::binary);
  if (!out) return NWG_EINVAL;
  const uint64_t N = n, st = (uint64_t)n + 1, es = 4, narr = 1;
  const char open_flag = 0;
  const uint32_t n_field[2] = {n + 1u, 0u};
  out.write(reinterpret_cast<const char*>(&N), 8);
  out.write(&open_flag, 1);
  out.write(kAdjMagic, 34);
  out.write(reinterpret_cast<const char*>(n_field), 8);
  out.write(reinterpret_cast<const char*>(&st), 8);
  out.write(reinterpret_cast<const char*>(&es), 8);
  out.write(reinterpret_cast<const char*>(ptr), (std::streamsize)(st * 4));
  out.write(reinterpret_cast<const char*>(&nnz), 8);
  out.write(reinterpret_cast<const char*>(&narr), 8);
  out.write(reinterpret_cast<const char*>(&nnz), 8);
  out.write(reinterpret_cast<const char*>(&es), 8);
  if (nnz) out.write(reinterpret_cast<const char*>(idx), (std::streamsize)(nnz * 4));
  return out ? NWG_OK : NWG_EFORMAT;
}

int nwg_write_adjacency(const char* path, uint32_t n, uint64_t nnz, const uint32_t* ptr, const uint32_t* idx) {
  return io_guarded([&] { return write_adjacency(path, n, nnz, ptr, idx); });
}

static int read_adjacency(const char* path, uint32_t* n, uint64_t* nnz, uint32_t** ptr, uint32_t** idx) {
  if (!path || !n || !nnz || !ptr || !idx) return NWG_EINVAL;
  std::ifstream in(path, std::ios::binary);
  if (!in) return NWG_EINVAL;
  const uint64_t bytes = stream_bytes(in);
  uint64_t N = 0, nfield = 0, st = 0, es = 0, ne = 0, narr = 0, cnt = 0, es2 = 0;
  char open_flag = 0, magic[34];
  in.read(reinterpret_cast<char*>(&N), 8);
  in.read(&open_flag, 1);
  in.read(magic, 34);
  if (!in || std::memcmp(magic, kAdjMagic, 34) != 0) return NWG_EFORMAT;
  in.read(reinterpret_cast<char*>(&nfield), 8);
  in.read(reinterpret_cast<char*>(&st), 8);
  in.read(reinterpret_cast<char*>(&es), 8);
  if (!in || es != 4 || N > 0xFFFFFFFFull || st != N + 1 || st > bytes / 4) return NWG_EFORMAT;
  std::vector<uint32_t> rp(st);
  in.read(reinterpret_cast<char*>(rp.data()), (std::streamsize)(st * 4));
  in.read(reinterpret_cast<char*>(&ne), 8);
  in.read(reinterpret_cast<char*>(&narr), 8);
  in.read(reinterpret_cast<char*>(&cnt), 8);
  in.read(reinterpret_cast<char*>(&es2), 8);
  if (!in || narr != 1 || cnt != ne || es2 != 4 || rp[0] != 0 || rp[N] != ne || ne > bytes / 4)
    return NWG_EFORMAT;
  for (uint64_t i = 0; i < N; ++i)
    if (rp[i + 1] < rp[i]) return NWG_EFORMAT;
  std::vector<uint32_t> col(ne);
  if (ne) in.read(reinterpret_cast<char*>(col.data()), (std::streamsize)(ne * 4));
  if (!in) return NWG_EFORMAT;
  for (uint64_t q = 0; q < ne; ++q)
    if (col[q] >= N) return NWG_EFORMAT;
  uint32_t* a = static_cast<uint32_t*>(std::malloc(st * sizeof(uint32_t)));
  uint32_t* b = static_cast<uint32_t*>(std::malloc((ne ? ne : 1) * sizeof(uint32_t)));
  if (!a || !b) {
    std::free(a);
    std::free(b);
    return NWG_ENOMEM;
  }
  std::memcpy(a, rp.data(), st * 4);
  if (ne) std::memcpy(b, col.data(), ne * 4);
  *n = (uint32_t)N;
  *nnz = ne;
  *ptr = a;
  *idx = b;
  return NWG_OK;
}

int nwg_read_adjacency(const char* path, uint32_t* n, uint64_t* nnz, uint32_t** ptr, uint32_t** idx) {
  return io_guarded([&] { return read_adjacency(path, n, nnz, ptr, idx); });
}

void nwg_free(void* p) { std::free(p); }

}  // extern "C"
