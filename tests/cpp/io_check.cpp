// io_check.cpp -- drives the file readers and writers of nwgraph_amd/csrc/io.cpp, built
// together with it under AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitize.py: valid files and round trips, malformed and oversized headers,
// every truncation of valid files and seeded random corruptions.  Every input must give a
// status (never an exception or a crash), the malformed ones a failure; the sanitizers abort
// on any invalid access or undefined behaviour.
//
//   io_check GOLDEN_DIR TMP_DIR
#include <nwgraph_hip.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what.c_str());
    ++failures;
  }
}

bool known_status(int rc) {
  return rc == NWG_OK || rc == NWG_EINVAL || rc == NWG_ENOMEM || rc == NWG_EFORMAT;
}

std::vector<char> slurp(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  return std::vector<char>(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

void spit(const std::string& path, const char* data, size_t len) {
  std::ofstream out(path, std::ios::binary | std::ios::trunc);
  out.write(data, (std::streamsize)len);
}

void spit(const std::string& path, const std::string& text) { spit(path, text.data(), text.size()); }

using EdgeReader = int (*)(const char*, uint32_t*, uint64_t*, uint32_t**, uint32_t**);

// runs an edge-list reader; frees what it returns
int run_edges(EdgeReader fn, const std::string& path, uint32_t* n_out = nullptr, std::vector<uint32_t>* s = nullptr,
              std::vector<uint32_t>* d = nullptr) {
  uint32_t n = 0;
  uint64_t ne = 0;
  uint32_t* src = nullptr;
  uint32_t* dst = nullptr;
  const int rc = fn(path.c_str(), &n, &ne, &src, &dst);
  if (rc == NWG_OK) {
    for (uint64_t k = 0; k < ne; ++k) expect(src[k] < n && dst[k] < n, "edge id below n: " + path);
    if (n_out) *n_out = n;
    if (s) s->assign(src, src + ne);
    if (d) d->assign(dst, dst + ne);
    nwg_free(src);
    nwg_free(dst);
  }
  return rc;
}

int run_adjacency(const std::string& path, uint32_t* n_out = nullptr, std::vector<uint32_t>* ptr = nullptr,
                  std::vector<uint32_t>* idx = nullptr) {
  uint32_t n = 0;
  uint64_t nnz = 0;
  uint32_t* p = nullptr;
  uint32_t* q = nullptr;
  const int rc = nwg_read_adjacency(path.c_str(), &n, &nnz, &p, &q);
  if (rc == NWG_OK) {
    expect(p[0] == 0 && p[n] == nnz, "adjacency offsets span the ids: " + path);
    for (uint32_t i = 0; i < n; ++i) expect(p[i] <= p[i + 1], "adjacency offsets monotone: " + path);
    for (uint64_t k = 0; k < nnz; ++k) expect(q[k] < n, "adjacency id below n: " + path);
    if (n_out) *n_out = n;
    if (ptr) ptr->assign(p, p + (size_t)n + 1);
    if (idx) idx->assign(q, q + nnz);
    nwg_free(p);
    nwg_free(q);
  }
  return rc;
}

// every prefix of a valid file (and, for long files, a sample of them), then seeded byte
// corruptions: each must give a known status
template <class Run>
void mutate(const std::vector<char>& good, const std::string& tmp, const std::string& tag, Run run) {
  const std::string path = tmp + "/mut_" + tag;
  const size_t step = std::max<size_t>(1, good.size() / 512);
  for (size_t len = 0; len < good.size(); len += (len < 256 ? 1 : step)) {
    spit(path, good.data(), len);
    expect(known_status(run(path)), tag + " truncated to " + std::to_string(len));
  }
  std::mt19937_64 rng(0x5EED0000u + good.size());
  for (int t = 0; t < 300; ++t) {
    std::vector<char> b = good;
    const int flips = 1 + (int)(rng() % 8);
    for (int f = 0; f < flips && !b.empty(); ++f) {
      const size_t at = rng() % b.size();
      switch (rng() % 3) {
        case 0: b[at] = (char)(rng() & 0xFF); break;
        case 1: b[at] = (char)0xFF; break;
        default: b[at] ^= (char)(1u << (rng() % 8)); break;
      }
    }
    spit(path, b.data(), b.size());
    expect(known_status(run(path)), tag + " corrupted, trial " + std::to_string(t));
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: io_check GOLDEN_DIR TMP_DIR\n");
    return 2;
  }
  const std::string golden = argv[1], tmp = argv[2];

  // the reference's karate.mtx: 34 vertices, 78 symmetric entries -> 156 directed edges
  uint32_t n = 0;
  std::vector<uint32_t> s, d;
  int rc = run_edges(nwg_read_mm, golden + "/karate.mtx", &n, &s, &d);
  expect(rc == NWG_OK && n == 34 && s.size() == 156, "karate.mtx reads as 34 vertices, 156 edges");

  // NW edge list round trip
  const std::string nw = tmp + "/karate.nw";
  expect(nwg_write_nw(nw.c_str(), n, s.size(), s.data(), d.data()) == NWG_OK, "write_nw");
  uint32_t n2 = 0;
  std::vector<uint32_t> s2, d2;
  expect(run_edges(nwg_read_nw, nw, &n2, &s2, &d2) == NWG_OK && n2 == n && s2 == s && d2 == d, "NW round trip");

  // adjacency round trip: a CSC of the karate edges (rows = targets, sources ascending)
  std::vector<std::pair<uint32_t, uint32_t>> e;
  for (size_t k = 0; k < s.size(); ++k) e.emplace_back(d[k], s[k]);
  std::sort(e.begin(), e.end());
  std::vector<uint32_t> ptr(n + 1, 0), idx;
  for (auto& [v, u] : e) {
    ++ptr[v + 1];
    idx.push_back(u);
  }
  for (uint32_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  const std::string adj = tmp + "/karate.adj";
  expect(nwg_write_adjacency(adj.c_str(), n, idx.size(), ptr.data(), idx.data()) == NWG_OK, "write_adjacency");
  uint32_t n3 = 0;
  std::vector<uint32_t> p3, i3;
  expect(run_adjacency(adj, &n3, &p3, &i3) == NWG_OK && n3 == n && p3 == ptr && i3 == idx, "adjacency round trip");

  // malformed Matrix Market files: each is a failure, not a crash
  const std::string banner = "%%MatrixMarket matrix coordinate pattern general\n";
  const std::vector<std::pair<std::string, std::string>> bad_mm = {
      {"empty", ""},
      {"banner only", banner},
      {"bad banner", "%%MatrixMarkt matrix coordinate pattern general\n2 2 1\n1 2\n"},
      {"skew-symmetric", "%%MatrixMarket matrix coordinate pattern skew-symmetric\n2 2 1\n1 2\n"},
      {"huge nnz", banner + "2 2 4000000000000000000\n1 2\n"},
      {"nnz beyond the file", banner + "2 2 1000\n1 2\n2 1\n"},
      {"zero id", banner + "2 2 1\n0 2\n"},
      {"id beyond 32 bits", banner + "2 2 1\n1 4294967297\n"},
      {"letters", banner + "2 2 1\nx y\n"},
      {"short size line", banner + "2 2\n1 2\n"},
  };
  for (auto& [tag, text] : bad_mm) {
    const std::string p = tmp + "/bad.mtx";
    spit(p, text);
    rc = run_edges(nwg_read_mm, p);
    expect(rc != NWG_OK && known_status(rc), "malformed Matrix Market (" + tag + ") fails cleanly");
  }
  expect(run_edges(nwg_read_mm, tmp + "/no_such_file.mtx") == NWG_EINVAL, "missing file is EINVAL");

  // oversized counts in the binary formats: overwrite every 8-byte window of the header with
  // a huge count; each read must fail without allocating for it (or read valid data)
  for (const auto& [path, tag] : {std::pair<std::string, std::string>{nw, "nw"}, {adj, "adjacency"}}) {
    const std::vector<char> good = slurp(path);
    for (size_t at = 0; at + 8 <= std::min<size_t>(good.size(), 160); ++at) {
      std::vector<char> b = good;
      const uint64_t huge = 0x4000000000000000ull;
      std::memcpy(b.data() + at, &huge, 8);
      const std::string p = tmp + "/huge_" + tag;
      spit(p, b.data(), b.size());
      rc = tag == "nw" ? run_edges(nwg_read_nw, p) : run_adjacency(p);
      expect(known_status(rc), tag + " with a huge count at byte " + std::to_string(at));
    }
  }

  // every truncation and random corruptions of each valid file
  mutate(slurp(golden + "/karate.mtx"), tmp, "mtx", [](const std::string& p) { return run_edges(nwg_read_mm, p); });
  mutate(slurp(nw), tmp, "nw", [](const std::string& p) { return run_edges(nwg_read_nw, p); });
  mutate(slurp(adj), tmp, "adj", [](const std::string& p) { return run_adjacency(p); });

  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("io_check: all reader checks passed\n");
  return 0;
}
