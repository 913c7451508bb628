// noalloc.hip -- one-call PageRank over a CSC already in device memory that allocates
// nothing: nwg_pagerank_f32_noalloc.
//
// The solve is the Jacobi (page_rank_v10) schedule of nwg_pagerank_f32, with the
// reference's scalars (page_rank.hpp:93-94,116,133-135,144): x = (float)(1.0 / N),
// contrib = x / (float)deg, x' = fma(d, z, base) with base = (float)((1 - (double)d) / N),
// the L1 change |x' - x| summed in double and compared with (double)threshold.
//
// Everything beyond the caller's inputs (CSC, optional out-degrees) and output (scores)
// lives in `scratch`, one file-scope device array per GPU: two contribution arrays, the
// counted out-degrees, per-workgroup L1 partials, the per-sweep L1 accumulators and a
// validation flag.  Its size caps the vertices (kNaMaxN) and sweeps (kNaMaxIters) of one
// call.  The sweep needs no plan (the pull sweep's row lists would need allocations):
// one workgroup per 256 consecutive rows; a row of at most kNaLaneMax in-edges is summed
// by its own lane in row order, a longer row by its whole wavefront (lane-strided partial
// sums, then a butterfly).  Every call on a GPU uses that GPU's array, so calls that
// overlap on one GPU race; each call clears what it reads before use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "capi_util.hpp"
#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr uint32_t kNaThreads = 256;
constexpr uint32_t kNaLaneMax = 32;  // rows with more in-edges are summed by a whole wave
constexpr uint32_t kNaUnroll = 4;    // gathers in flight per lane
constexpr uint32_t kNaFoldChunk = 2048;  // workgroup partials folded by one workgroup
constexpr uint32_t kNaMaxN = 1u << 22;
constexpr uint64_t kNaMaxIters = 4096;
constexpr uint32_t kNaMaxBlocks = kNaMaxN / kNaThreads;
constexpr int kNaMaxDevices = 64;

// byte offsets of the regions of `scratch` (each 256-byte aligned)
constexpr size_t kOffC0 = 0;
constexpr size_t kOffC1 = kOffC0 + 4ull * kNaMaxN;
constexpr size_t kOffDeg = kOffC1 + 4ull * kNaMaxN;
constexpr size_t kOffBlk = kOffDeg + 4ull * kNaMaxN;
constexpr size_t kOffAcc = kOffBlk + 8ull * kNaMaxBlocks;
constexpr size_t kOffFlag = kOffAcc + 8ull * kNaMaxIters;
constexpr size_t kScratchBytes = kOffFlag + 256;
static_assert(kOffC1 % 256 == 0 && kOffDeg % 256 == 0 && kOffBlk % 256 == 0 && kOffAcc % 256 == 0 &&
                  kOffFlag % 256 == 0,
              "scratch regions must stay 256-byte aligned");

}  // namespace

// The library's only statically sized device buffer: 48.2 MiB per GPU, held until the
// process ends.  Host code never passes its name to a kernel or a copy (on the host it is
// a stand-in symbol, not a device address); it uses the address hipGetSymbolAddress
// returns on each device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};
hipError_t g_na_status[kNaMaxDevices] = {};
bool g_na_done[kNaMaxDevices] = {};

struct NaParams {
  const uint32_t* rp;
  const uint32_t* col;
  const uint32_t* deg;
  const float* cin;
  float* cout;
  float* x;
  double* blk_err;
  unsigned long long* acc;
  float base;
  float damp;
  uint32_t n;
  uint32_t nblk;
  uint32_t iter;
  double threshold;
};

__device__ __forceinline__ bool na_converged_before(const NaParams& P) {
  return P.iter > 0 && fixed_to_err(P.acc[P.iter - 1]) < P.threshold;
}

__device__ __forceinline__ double na_update(const NaParams& P, uint32_t i, float z) {
  const float xo = P.x[i];
  const float xn = __builtin_fmaf(P.damp, z, P.base);
  P.x[i] = xn;
  P.cout[i] = xn / (float)P.deg[i];
  return (double)fabsf(xn - xo);
}

// z + cin[col[q]] for q = p0, p0 + S, ... < p1, added in that order, kNaUnroll gathers in flight
template <uint32_t S>
__device__ __forceinline__ float na_gather(const NaParams& P, uint64_t p0, uint64_t p1, float z) {
  for (uint64_t p = p0; p < p1; p += (uint64_t)S * kNaUnroll) {
    float v[kNaUnroll];
#pragma unroll
    for (uint32_t r = 0; r < kNaUnroll; ++r) {
      const uint64_t q = p + (uint64_t)r * S;
      v[r] = q < p1 ? P.cin[P.col[q]] : 0.f;
    }
#pragma unroll
    for (uint32_t r = 0; r < kNaUnroll; ++r)
      if (p + (uint64_t)r * S < p1) z += v[r];
  }
  return z;
}

__global__ __launch_bounds__(kNaThreads) void na_init_kernel(uint32_t n, const uint32_t* __restrict__ deg, float init,
                                                             float* __restrict__ x, float* __restrict__ c0,
                                                             unsigned long long* __restrict__ acc, uint64_t iters) {
  const uint64_t i = (uint64_t)blockIdx.x * kNaThreads + threadIdx.x;
  if (i < n) {
    x[i] = init;
    c0[i] = init / (float)deg[i];
  }
  if (i < iters) acc[i] = 0ull;
}

// One sweep over 256 consecutive rows per workgroup; the workgroup's L1 change goes to
// blk_err[blockIdx.x].
__global__ __launch_bounds__(kNaThreads) void na_sweep_kernel(NaParams P) {
  if (na_converged_before(P)) return;
  __shared__ double red[kNaThreads / 64];
  const uint32_t i = blockIdx.x * kNaThreads + threadIdx.x;
  uint32_t p0 = 0, p1 = 0;
  if (i < P.n) {
    p0 = P.rp[i];
    p1 = P.rp[i + 1];
  }
  const bool wide = p1 - p0 > kNaLaneMax;
  double e = 0.0;
  if (i < P.n && !wide) e = na_update(P, i, na_gather<1>(P, p0, p1, 0.f));
  // rows with more than kNaLaneMax in-edges: one at a time, summed by the whole wave
  uint64_t m = __ballot(wide);
  const uint32_t lane = lane_id();
  while (m) {
    const uint32_t l = (uint32_t)__builtin_ctzll(m);
    m &= m - 1;
    const uint32_t q0 = __shfl(p0, (int)l), q1 = __shfl(p1, (int)l);
    float z = na_gather<64>(P, (uint64_t)q0 + lane, q1, 0.f);
    z = wave_sum(z);
    if (lane == l) e += na_update(P, i, z);
  }
  e = block_sum<kNaThreads>(e, red);
  if (threadIdx.x == 0) P.blk_err[blockIdx.x] = e;
}

// Folds kNaFoldChunk workgroup partials per workgroup into acc[iter] (integer atomics:
// the total does not depend on the order the workgroups finish in).
__global__ __launch_bounds__(kNaThreads) void na_fold_kernel(NaParams P) {
  if (na_converged_before(P)) return;
  __shared__ double red[kNaThreads / 64];
  const uint32_t c0 = blockIdx.x * kNaFoldChunk;
  double e = 0.0;
  for (uint32_t q = c0 + threadIdx.x; q < c0 + kNaFoldChunk && q < P.nblk; q += kNaThreads) e += P.blk_err[q];
  e = block_sum<kNaThreads>(e, red);
  if (threadIdx.x == 0 && e > 0.0) atomicAdd(P.acc + P.iter, err_to_fixed(e));
}

void* scratch_on(int device) {
  NWG_REQUIRE(device >= 0 && device < kNaMaxDevices, "device index beyond the allocation-free path's table");
  std::lock_guard<std::mutex> lk(g_na_mu);
  NWG_REQUIRE(g_na_done[device], "allocation-free scratch was not set up on this device (nwg_create does it)");
  if (g_na_status[device] != hipSuccess)
    throw Error(NWG_EHIP, std::string("allocation-free scratch unavailable: hipGetSymbolAddress failed: ") +
                              hipGetErrorString(g_na_status[device]));
  return g_na_ptr[device];
}

}  // namespace

// Called by nwg_create with `device` current: looks up this device's copy of `scratch`
// once.  A failure is recorded (and reported by nwg_pagerank_f32_noalloc), so a context
// still serves the other entry points.
void noalloc_setup(int device) {
  if (device < 0 || device >= kNaMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_done[device]) return;
  void* p = nullptr;
  g_na_status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_na_ptr[device] = g_na_status[device] == hipSuccess ? p : nullptr;
  g_na_done[device] = true;
}

uint64_t noalloc_max_vertices() { return kNaMaxN; }
uint64_t noalloc_max_iters() { return kNaMaxIters; }

void pagerank_noalloc(nwg_ctx* ctx, uint32_t n, uint64_t nnz, const uint32_t* csc_ptr, const uint32_t* csc_idx,
                      const uint32_t* outdeg, float* scores, float damping, float threshold, uint64_t max_iters,
                      nwg_pr_info* info) {
  NWG_REQUIRE(n > 0, "n must be positive");
  NWG_REQUIRE(n <= kNaMaxN, "nwg_pagerank_f32_noalloc: n exceeds the scratch capacity of 4194304 vertices");
  NWG_REQUIRE(max_iters <= kNaMaxIters, "nwg_pagerank_f32_noalloc: max_iters exceeds 4096");
  NWG_REQUIRE(nnz <= 0xFFFFFFFFull, "nnz must fit the 32-bit row offsets");
  NWG_REQUIRE(csc_ptr && scores && (nnz == 0 || csc_idx), "CSC and scores must not be null");
  NWG_REQUIRE(!info || !(info->flags & NWG_INFO_KERNEL_TIMING),
              "nwg_pagerank_f32_noalloc does not time sweeps one by one (NWG_INFO_KERNEL_TIMING)");
  unsigned char* base = static_cast<unsigned char*>(scratch_on(ctx->device));
  float* c0 = reinterpret_cast<float*>(base + kOffC0);
  float* c1 = reinterpret_cast<float*>(base + kOffC1);
  uint32_t* cdeg = reinterpret_cast<uint32_t*>(base + kOffDeg);
  double* blk = reinterpret_cast<double*>(base + kOffBlk);
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(base + kOffAcc);
  uint32_t* flag = reinterpret_cast<uint32_t*>(base + kOffFlag);
  hipStream_t s = ctx->stream;

  validate_device_csc(s, n, nnz, csc_ptr, csc_idx, flag);
  const uint32_t* deg = outdeg;
  if (outdeg) {
    check_outdeg_device(s, n, nnz, csc_idx, outdeg, cdeg, flag);
  } else {
    out_degrees_device(s, n, nnz, csc_idx, cdeg, false);  // build_degrees, common.hpp:110-121
    deg = cdeg;
  }

  NaParams P{};
  P.rp = csc_ptr;
  P.col = csc_idx;
  P.deg = deg;
  P.x = scores;
  P.blk_err = blk;
  P.acc = acc;
  P.base = (float)((1.0 - (double)damping) / n);  // page_rank.hpp:94
  P.damp = damping;
  P.n = n;
  P.nblk = (uint32_t)ceil_div(n, kNaThreads);
  P.threshold = (double)threshold;  // page_rank.hpp:144
  const uint32_t fold_blocks = (uint32_t)ceil_div(P.nblk, kNaFoldChunk);

  NWG_HIP(hipEventRecord(ctx->ev0, s));
  const float init = (float)(1.0 / n);  // page_rank.hpp:93
  na_init_kernel<<<(uint32_t)ceil_div(std::max<uint64_t>(n, max_iters), kNaThreads), kNaThreads, 0, s>>>(
      n, deg, init, scores, c0, acc, max_iters);
  NWG_LAUNCH_CHECK();
  for (uint64_t k = 0; k < max_iters; ++k) {
    P.iter = (uint32_t)k;
    P.cin = (k & 1) ? c1 : c0;
    P.cout = (k & 1) ? c0 : c1;
    na_sweep_kernel<<<P.nblk, kNaThreads, 0, s>>>(P);
    na_fold_kernel<<<fold_blocks, kNaThreads, 0, s>>>(P);
    NWG_LAUNCH_CHECK();
  }
  NWG_HIP(hipEventRecord(ctx->ev1, s));

  // the per-sweep L1 changes, read back in pieces through a fixed host array
  constexpr uint64_t kPiece = 512;
  unsigned long long h[kPiece];
  uint64_t iters = max_iters;
  double last = 0.0;
  bool stopped = false;
  for (uint64_t k0 = 0; k0 < max_iters && !stopped; k0 += kPiece) {
    const uint64_t cnt = std::min<uint64_t>(kPiece, max_iters - k0);
    NWG_HIP(hipMemcpyAsync(h, acc + k0, cnt * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    for (uint64_t q = 0; q < cnt; ++q) {
      const double e = fixed_to_err(h[q]);
      last = e;
      if (info && info->error_trace && k0 + q < info->error_trace_len) info->error_trace[k0 + q] = e;
      if (e < (double)threshold) {
        iters = k0 + q + 1;
        stopped = true;
        break;
      }
    }
  }
  NWG_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  NWG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (info) {
    info->iterations = iters;
    info->final_error = max_iters ? last : 0.0;
    info->ms_total = ms;
    info->ms_per_iter = iters ? ms / (double)iters : 0.0;
    info->sweep_ms_avg = 0.0;
    info->init_ms[0] = info->init_ms[1] = 0.0;
  }
}

}  // namespace nwg
