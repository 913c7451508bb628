// pr.cpp -- PageRank benchmark driver over libnwgraph_hip, the counterpart of the
// reference's bench/gapbs/pr.cpp (usage pr.cpp:16-34, main pr.cpp:54-177).
//
//   pr [--version ID...] -f FILE... [-i NUM] [-t NUM] [-n NUM] [-dvV] [--log FILE] [--log-header] [--autotune] [THREADS]...
//
// FILE is a Matrix Market file, an NW binary edge list (load_graph,
// bench/gapbs/common.hpp:82-101), or a synthetic spec generated on the device:
//   rmat:SCALE[:EDGE_FACTOR[:SEED]]   RMAT, A/B/C/D = .57/.19/.19/.05, cleaned
//   grid:WIDTH[:SEED]                 4-neighbour lattice (p=0.85) + 0.5% long-range edges
// Version 10 runs the Jacobi schedule of page_rank_v10 (nwg_graph_pagerank, deterministic);
// version 11 runs the reference's in-place schedule (nwg_graph_pagerank_inplace: one
// contribution array updated during the sweep, on the pull kernels or the binned sweep,
// schedule-dependent like the reference's TBB run); version 13 the delta (residual) PageRank of page_rank_v13
// (nwg_graph_pagerank_delta, un-normalised scores).  --autotune picks the sweep plan first
// (nwg_graph_autotune).  Defaults follow
// the reference: version 11, -i 20 (0 means 1), -t 1e-4, -n 1 (0 means 1).  THREADS are
// accepted and reported like the reference's (which only uses them for TBB); the GPU
// ignores them.
#include <nwgraph_hip.h>

#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include <nwgraph_hip/page_rank.hpp>

namespace {

struct Args {
  std::vector<long> versions;
  std::vector<std::string> files;
  long max_iters = 20;
  float tolerance = 1e-4f;
  long trials = 1;
  bool debug = false, verify = false, verbose = false, log_header = false, autotune = false;
  std::string log_file;
  std::vector<long> threads;
};

[[noreturn]] void usage(int rc) {
  std::cout << "pr: MI355X page rank benchmark driver.\n"
               "  Usage:\n"
               "      pr (-h | --help)\n"
               "      pr [--version ID...] -f FILE... [-i NUM] [-t NUM] [-n NUM] [-dvV] [--log FILE] [--log-header] "
               "[--autotune] [THREADS]...\n\n"
               "  Options:\n"
               "      -h, --help          show this screen\n"
               "      --version ID        algorithm version to run [default: 11]\n"
               "      -f FILE             input file path (.mtx, NW binary, rmat:SCALE[:EF[:SEED]], grid:W[:SEED])\n"
               "      -i NUM              maximum iteration [default: 20]\n"
               "      -t NUM              tolerance [default: 1e-4]\n"
               "      -n NUM              number of trials [default: 1]\n"
               "      --log FILE          log times to a file\n"
               "      --log-header        add a header to the log file\n"
               "      -d, --debug         run in debug mode\n"
               "      -v, --verify        verify results\n"
               "      -V, --verbose       run in verbose mode\n"
               "      --autotune          choose the sweep plan on the device first (not in the reference)\n"
               "  Versions: 10 = Jacobi schedule (reproducible), 11 = in-place schedule (the reference's default),\n"
               "            13 = delta PageRank (un-normalised scores)\n";
  std::exit(rc);
}

Args parse(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto need = [&](const char* what) -> std::string {
      if (i + 1 >= argc) {
        std::cerr << "missing value for " << what << "\n";
        usage(1);
      }
      return argv[++i];
    };
    if (s == "-h" || s == "--help") usage(0);
    else if (s == "--version") a.versions.push_back(std::stol(need("--version")));
    else if (s == "-f") a.files.push_back(need("-f"));
    else if (s == "-i") a.max_iters = std::stol(need("-i"));
    else if (s == "-t") a.tolerance = std::stof(need("-t"));
    else if (s == "-n") a.trials = std::stol(need("-n"));
    else if (s == "--log") a.log_file = need("--log");
    else if (s == "--log-header") a.log_header = true;
    else if (s == "--autotune") a.autotune = true;
    else if (s.size() > 1 && s[0] == '-' && s[1] != '-' && !std::isdigit((unsigned char)s[1])) {
      for (size_t k = 1; k < s.size(); ++k) {
        if (s[k] == 'd') a.debug = true;
        else if (s[k] == 'v') a.verify = true;
        else if (s[k] == 'V') a.verbose = true;
        else usage(1);
      }
    } else if (s == "--debug") a.debug = true;
    else if (s == "--verify") a.verify = true;
    else if (s == "--verbose") a.verbose = true;
    else a.threads.push_back(std::stol(s));
  }
  if (a.files.empty()) usage(1);
  if (a.versions.empty()) a.versions.push_back(11);
  if (a.max_iters == 0) a.max_iters = 1;  // `?: 1` in the reference (pr.cpp:63)
  if (a.trials == 0) a.trials = 1;        // pr.cpp:62
  if (a.threads.empty()) a.threads.push_back(1);
  return a;
}

void check(nwg_ctx* ctx, int rc, const char* what) { nw::graph::hip::check(ctx, rc, what); }

struct LifeTimer {  // util/timer.hpp:120-133
  std::string msg;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit LifeTimer(std::string m) : msg(std::move(m)) {}
  ~LifeTimer() {
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "# [ " + msg + " ]: " << ms << " ms" << std::endl;
  }
};

nwg_graph* load(nwg_ctx* ctx, const std::string& file, bool verbose) {
  nwg_graph* g = nullptr;
  auto field = [&](size_t k, long dflt) {
    size_t pos = 0;
    for (size_t q = 0; q < k; ++q) {
      pos = file.find(':', pos);
      if (pos == std::string::npos) return dflt;
      ++pos;
    }
    return std::stol(file.substr(pos));
  };
  if (file.rfind("rmat:", 0) == 0) {
    LifeTimer t("generate rmat + build adjacency");
    check(ctx, nwg_generate_rmat(ctx, (uint32_t)field(1, 20), (uint32_t)field(2, 16), 5700, 1900, 1900,
                                 (uint64_t)field(3, 1), &g),
          "nwg_generate_rmat");
    return g;
  }
  if (file.rfind("grid:", 0) == 0) {
    LifeTimer t("generate grid + build adjacency");
    check(ctx, nwg_generate_grid(ctx, (uint32_t)field(1, 1024), 8500, 50, (uint64_t)field(2, 1), &g),
          "nwg_generate_grid");
    return g;
  }
  std::ifstream in(file);
  std::string type;
  in >> type;
  uint32_t n = 0;
  uint64_t ne = 0;
  uint32_t *src = nullptr, *dst = nullptr;
  if (type == "NW") {
    LifeTimer t("deserialize");
    check(ctx, nwg_read_nw(file.c_str(), &n, &ne, &src, &dst), "nwg_read_nw");
  } else if (type == "%%MatrixMarket") {
    std::cout << "Reading matrix market input " << file << " (slow)\n";
    LifeTimer t("read mm");
    check(ctx, nwg_read_mm(file.c_str(), &n, &ne, &src, &dst), "nwg_read_mm");
  } else {
    std::cerr << "Did not recognize graph input file " << file << "\n";
    std::exit(1);
  }
  if (verbose) std::cout << "% edge list: " << n << " vertices, " << ne << " edges\n";
  {
    LifeTimer t("build adjacency");  // adjacency<1> keeps self-loops and duplicates
    const int rc = nwg_graph_build_from_edges(ctx, n, ne, src, dst, 0, NWG_MEM_HOST, &g);
    nwg_free(src);
    nwg_free(dst);
    check(ctx, rc, "nwg_graph_build_from_edges");
  }
  return g;
}

// Compiler facts of the Log's CXX / CXX_ID / CXX_VER / Build columns (the reference takes
// them from its CMake configuration, config.h).
#ifndef NWG_CXX_PATH
#define NWG_CXX_PATH "c++"
#endif
#if defined(__clang__)
constexpr const char* kCxxId = "Clang";
#elif defined(__GNUC__)
constexpr const char* kCxxId = "GNU";
#else
constexpr const char* kCxxId = "unknown";
#endif
constexpr const char* kCxxVersion = __VERSION__;
constexpr const char* kBuildType = "Release";
constexpr const char* kTbbMalloc = "OFF";  // no TBB here

// The reference's result log (bench/gapbs/Log.hpp:39-179): one row per sample, with a run
// uuid, library, git branch and revision, compiler, build, date, host, benchmark, version
// ("v" + id), threads, the graph's file stem cut at its first '_', and the sample's time.
// "-" writes to stdout.  Column widths and number formats follow the reference.
class Log {
public:
  explicit Log(const std::string& path) : os_(path == "-" ? std::cout : out_) {
    std::random_device rd;
    std::mt19937 gen(rd());
    std::uniform_int_distribution<int> dis(97, 122);
    uuid_.resize(kUuidSize);
    for (char& c : uuid_) c = (char)dis(gen);
    std::memset(host_, 0, sizeof host_);
    if (gethostname(host_, sizeof host_) != 0) std::cerr << "truncated host name\n";
    host_[sizeof host_ - 1] = 0;
    if (path != "-") {
      std::filesystem::path full = path;
      if (full.has_parent_path()) std::filesystem::create_directory(full.parent_path());
      out_.open(path, std::ofstream::out | std::ofstream::app);
    }
    const std::time_t now = std::time(nullptr);
    char day[16];
    std::strftime(day, sizeof day, "%Y-%m-%d", std::gmtime(&now));
    date_ = day;
    git_branch_ = run("git rev-parse --abbrev-ref HEAD 2>/dev/null");
    git_version_ = run("git log --pretty=format:%h -n 1 2>/dev/null");
    if (!git_version_.empty() && std::system("git diff --quiet --exit-code 2>/dev/null") != 0) git_version_ += "+";
    cxx_w_ = std::max(std::strlen(NWG_CXX_PATH), std::strlen("CXX")) + 2;
    cxx_id_w_ = std::max(std::strlen(kCxxId), std::strlen("CXX_ID")) + 2;
    cxx_ver_w_ = std::max(std::strlen(kCxxVersion), std::strlen("CXX_VER")) + 2;
    build_w_ = std::max(std::strlen(kBuildType), std::strlen("Build")) + 2;
    tbb_w_ = std::max(std::strlen(kTbbMalloc), std::strlen("TBBMALLOC")) + 2;
  }

  void header(const std::vector<std::string>& times) {
    auto p = os_.precision(15);
    auto f = os_.setf(std::ios::fixed, std::ios::floatfield);
    os_ << std::setw(kUuidSize + 2) << std::left << "UUID" << std::setw(10) << std::left << "Library"
        << std::setw(10) << std::left << "Branch" << std::setw(10) << std::left << "Revision" << std::setw(cxx_w_)
        << std::left << "CXX" << std::setw(cxx_id_w_) << std::left << "CXX_ID" << std::setw(cxx_ver_w_) << std::left
        << "CXX_VER" << std::setw(build_w_) << std::left << "Build" << std::setw(tbb_w_) << std::left << "TBBMALLOC"
        << std::setw(15) << std::left << "Date" << std::setw(sizeof host_) << std::left << "Host" << std::setw(10)
        << std::left << "Benchmark" << std::setw(10) << std::left << "Version" << std::setw(10) << std::left
        << "Threads" << std::setw(32) << std::left << "Graph";
    for (const auto& t : times) os_ << std::setw(20) << std::left << t;
    os_ << "\n";
    os_.precision(p);
    os_.setf(f, std::ios::floatfield);
  }

  void row(const std::string& benchmark, long version, long threads, const std::string& graph, double seconds) {
    auto p = os_.precision(8);
    auto f = os_.setf(std::ios::scientific, std::ios::floatfield);
    std::string stem = std::filesystem::path(graph).stem().string();
    const size_t cut = stem.find('_');
    if (cut < stem.size()) stem.resize(cut);
    os_ << std::setw(kUuidSize + 2) << std::left << uuid_ << std::setw(10) << std::left << "NWG_HIP" << std::setw(10)
        << std::left << git_branch_ << std::setw(10) << std::left << git_version_ << std::setw(cxx_w_) << std::left
        << NWG_CXX_PATH << std::setw(cxx_id_w_) << std::left << kCxxId << std::setw(cxx_ver_w_) << std::left
        << kCxxVersion << std::setw(build_w_) << std::left << kBuildType << std::setw(tbb_w_) << std::left
        << kTbbMalloc << std::setw(15) << std::left << date_ << std::setw(sizeof host_) << std::left << host_
        << std::setw(10) << std::left << benchmark << std::setw(10) << std::left << ("v" + std::to_string(version))
        << std::setw(10) << std::left << threads << std::setw(32) << std::left << stem << std::setw(20) << std::left
        << seconds << "\n";
    os_.precision(p);
    os_.setf(f, std::ios::floatfield);
  }

private:
  static std::string run(const char* cmd) {
    std::string out;
    if (FILE* f = popen(cmd, "r")) {
      char buf[256];
      while (std::fgets(buf, sizeof buf, f)) out += buf;
      pclose(f);
    }
    std::istringstream in(out);
    std::string word;
    in >> word;
    return word;
  }
  static constexpr size_t kUuidSize = 24;
  std::string uuid_, date_, git_branch_, git_version_;
  char host_[20];
  size_t cxx_w_ = 0, cxx_id_w_ = 0, cxx_ver_w_ = 0, build_w_ = 0, tbb_w_ = 0;
  std::ofstream out_;
  std::ostream& os_;
};

}  // namespace

int main(int argc, char** argv) {
  Args args = parse(argc, argv);
  nw::graph::hip::context context(0);
  nwg_ctx* ctx = context.get();
  // Times (common.hpp:176-255): (file, version, threads) -> samples
  std::map<std::tuple<std::string, long, long>, std::vector<double>> times;

  for (const auto& file : args.files) {
    nwg_graph* g = load(ctx, file, args.verbose);
    uint32_t n = 0;
    uint64_t nnz = 0;
    nwg_graph_info(g, &n, &nnz);
    if (args.verbose) std::cout << "% adjacency<1>: " << n << " vertices, " << nnz << " edges\n";
    std::cout << "# [ degrees ]: built on the device with the adjacency" << std::endl;
    if (args.autotune) {
      nwg_plan_params plan{};
      check(ctx, nwg_graph_autotune(ctx, g, &plan), "nwg_graph_autotune");
      if (args.verbose)
        std::cout << "% plan: binned " << plan.binned << " relabel " << plan.relabel << " short_max " << plan.short_max
                  << " group_max " << plan.group_max << " hub_min " << plan.hub_min << "\n";
    }
    if (args.debug) {
      std::vector<uint32_t> ptr(n + 1);
      check(ctx, nwg_graph_get_csc(ctx, g, ptr.data(), nullptr, nullptr), "nwg_graph_get_csc");
      for (uint32_t i = 0; i <= std::min<uint32_t>(n, 32); ++i) std::cout << ptr[i] << " ";
      std::cout << "\n";
    }
    std::vector<float> rankings(n);
    std::vector<double> trace(args.max_iters), sweep_ms(args.max_iters);
    for (long thread : args.threads) {
      for (long id : args.versions) {
        for (long trial = 0; trial < args.trials; ++trial) {
          if (id != 10 && id != 11 && id != 13) {
            std::cerr << "Unknown version id " << id << std::endl;
            continue;
          }
          nwg_pr_info info{};
          info.error_trace = trace.data();
          info.error_trace_len = trace.size();
          // every sweep timed on the device, as the reference times each sweep (page_rank.hpp:123-142)
          info.flags = NWG_INFO_KERNEL_TIMING;
          info.time_trace = sweep_ms.data();
          info.time_trace_len = sweep_ms.size();
          const auto t0 = std::chrono::steady_clock::now();
          // version 11 is the reference's in-place schedule, version 10 its Jacobi twin,
          // version 13 the delta (residual) PageRank with un-normalised scores
          if (id == 13) {
            check(ctx,
                  nwg_graph_pagerank_delta(ctx, g, rankings.data(), 0.85f, args.tolerance, (uint64_t)args.max_iters,
                                           NWG_MEM_HOST, &info),
                  "nwg_graph_pagerank_delta");
          } else {
            const auto solve = id == 11 ? nwg_graph_pagerank_inplace : nwg_graph_pagerank;
            check(ctx,
                  solve(ctx, g, rankings.data(), nullptr, 0.85f, args.tolerance, (uint64_t)args.max_iters,
                        NWG_MEM_HOST, &info),
                  id == 11 ? "nwg_graph_pagerank_inplace" : "nwg_graph_pagerank");
          }
          const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
          nw::graph::hip::print_trace(info, trace, sweep_ms);
          times[std::make_tuple(file, id, thread)].push_back(sec);
        }
        if (args.verify) {  // print_n_ranks (pr.cpp:46-52): top 10 by score
          std::cout << "Verifying\n";
          std::vector<size_t> perm(n);
          std::iota(perm.begin(), perm.end(), 0);
          std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return rankings[a] > rankings[b]; });
          for (size_t i = 0; i < 10 && i < n; ++i)
            std::cout << std::to_string(perm[i]) + ": " << std::to_string(rankings[perm[i]]) << std::endl;
        }
      }
    }
    nwg_graph_destroy(g);
  }

  // Times::print (common.hpp:210-236)
  size_t w = 4;
  for (auto& [cfg, s] : times) w = std::max(w, std::get<0>(cfg).size());
  std::cout << std::setw(w + 2) << std::left << "File" << std::setw(10) << std::left << "Version" << std::setw(10)
            << std::left << "Threads" << std::setw(20) << std::left << "Min" << std::setw(20) << std::left << "Avg"
            << std::setw(20) << std::left << "Max" << "\n";
  for (auto& [cfg, s] : times) {
    const auto [mn, mx] = std::minmax_element(s.begin(), s.end());
    const double avg = std::accumulate(s.begin(), s.end(), 0.0) / s.size();
    std::cout << std::setw(w + 2) << std::left << std::get<0>(cfg) << std::setw(10) << std::left << std::get<1>(cfg)
              << std::setw(10) << std::left << std::get<2>(cfg) << std::setw(20) << std::left << std::setprecision(6)
              << std::fixed << *mn << std::setw(20) << std::left << std::setprecision(6) << std::fixed << avg
              << std::setw(20) << std::left << std::setprecision(6) << std::fixed << *mx << "\n";
  }
  if (!args.log_file.empty()) {  // log("pr", file, times, header, "Time(s)", "Tolerance") (pr.cpp:171-175)
    Log log(args.log_file);
    if (args.log_header) log.header({"Time(s)", "Tolerance"});
    for (auto& [cfg, s] : times)
      for (double t : s) log.row("pr", std::get<1>(cfg), std::get<2>(cfg), std::get<0>(cfg), t);
  }
  return 0;
}
