// binned.hip -- PageRank sweep by propagation blocking (binned schedule of the pull sweep).
//
// The pull sweep of nw::graph::page_rank (include/nwgraph/algorithms/page_rank.hpp:127-136)
// gathers contrib[u] for every in-edge; at 2^25 vertices every gather is its own L2
// request and most miss.  The binned sweep turns the gathers into two streaming passes
// over a static, tiled copy of the CSC (layout in nwg_internal.hpp, BinnedLayout):
//
//   chunk pass  work item = one source chunk (2^15 sources) and a run of its tiles:
//               stage contrib[chunk] in LDS, then vals[e] = contrib[src(e)] for every
//               entry e of those tiles (2-byte source offsets in, 4-byte values out)
//   bin pass    work item = a contiguous entry range of one bin (2^14 rows): add
//               vals[e] into LDS row accumulators (4 + 2 bytes in per entry); a bin
//               split over several items adds its partial rows into gacc with global
//               atomics and the last item of the bin finishes it.  Finishing a row is
//               the reference's update (page_rank.hpp:133-135): x = fma(d, z, base),
//               |x - x_old| into the sweep's L1 change, contrib = x / deg.
//
// Row sums are exact fixed-point sums (2^-60 units, integer adds), so they do not depend
// on how the entries are split or ordered: results are deterministic, and z is the float
// rounding of the (nearly) exact sum of the float contributions.  The L1 change is
// accumulated in fixed point as in the pull sweep (device_utils.hpp), and every
// workgroup first checks the previous sweep's change against the threshold
// (page_rank.hpp:144-146), so a captured loop stops working after convergence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kBT = 1024;  // threads of a bin / chunk pass workgroup (16 waves, one workgroup per CU)
constexpr uint32_t kBinRows = 1u << kBinBits;
constexpr uint32_t kChunkSrc = 1u << kChunkBits;
constexpr uint32_t kItem2Entries = 1u << 18;  // entries per bin-pass item
constexpr uint32_t kItem1Entries = 1u << 19;  // entries per chunk-pass item
constexpr float kSumScale = 0x1p60f;          // fixed point of row sums: z < 8 fits in 63 bits
constexpr double kSumUnit = 0x1p-60;

__device__ __forceinline__ unsigned long long sum_to_fixed(float c) {
  return (unsigned long long)__builtin_rintf(c * kSumScale);  // c >= 0; exact for c >= 2^-37
}
__device__ __forceinline__ float fixed_to_sum(unsigned long long v) { return (float)((double)v * kSumUnit); }

__device__ __forceinline__ bool converged_before(const BinParams& P) {
  return P.iter > 0 && fixed_to_err(P.acc[P.iter - 1]) < P.threshold;
}

// the reference's row update (page_rank.hpp:133-135) -> |x_new - x_old|
__device__ __forceinline__ double finish_row(const BinParams& P, uint32_t i, float z) {
  const float b = P.basev ? P.basev[i] : P.base;
  const float xo = P.x[i];
  const float xn = __builtin_fmaf(P.damp, z, b);
  P.x[i] = xn;
  P.contrib[i] = xn / (float)P.deg[i];
  return (double)fabsf(xn - xo);
}

__device__ __forceinline__ double block_sum_d(double v, double* red /* kBT/64 */) {
  v = wave_sum(v);
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (uint32_t q = 0; q < kBT / 64; ++q) t += red[q];
  return t;  // valid in thread 0
}

// chunk pass: vals[e] = contrib[src(e)] over the tiles of one item
__global__ __launch_bounds__(kBT) void pr_chunk_kernel(BinParams P) {
  if (converged_before(P)) return;
  __shared__ float cst[kChunkSrc];
  __shared__ uint32_t sstart[kBinMaxSegs];
  __shared__ uint32_t sflat[kBinMaxSegs + 1];
  const uint4 it = P.items1[blockIdx.x];  // {chunk, first tile, end tile, entries}
  const uint32_t ns = it.z - it.y, m = it.w;
  for (uint32_t k = threadIdx.x; k < ns; k += kBT) {
    sstart[k] = P.seg_start[it.y + k];
    sflat[k] = P.seg_flat[it.y + k];
  }
  if (threadIdx.x == 0) sflat[ns] = m;
  const uint32_t u0 = it.x << kChunkBits;
  for (uint32_t q = threadIdx.x; q < kChunkSrc; q += kBT) cst[q] = (u0 + q < P.n) ? P.contrib[u0 + q] : 0.f;
  __syncthreads();
  // each wave takes blocks of 512 flat entries; lane l handles entries k0 + l + 64 r
  const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (uint32_t k0 = w * 512; k0 < m; k0 += (kBT / 64) * 512) {
    const uint32_t k = k0 + l;
    if (k >= m) break;
    uint32_t lo = 0, hi = ns - 1;  // last tile j with sflat[j] <= k
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (sflat[mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    uint32_t j = lo;
    uint32_t p[8];
    uint16_t u[8];
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r) {
      const uint32_t kk = k + 64 * r;
      if (kk < m) {
        while (sflat[j + 1] <= kk) ++j;
        p[r] = sstart[j] + (kk - sflat[j]);
        u[r] = P.bsrc[p[r]];
      }
    }
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r)
      if (k + 64 * r < m) P.vals[p[r]] = cst[u[r]];
  }
}

// bin pass: row sums of one entry range of one bin in LDS, then finish or combine
__global__ __launch_bounds__(kBT) void pr_bin_kernel(BinParams P) {
  if (converged_before(P)) return;
  __shared__ unsigned long long lacc[kBinRows];
  __shared__ double red[kBT / 64];
  __shared__ uint32_t is_last;
  const uint4 it = P.items2[blockIdx.x];  // {bin, first entry, end entry, items of the bin}
  const uint32_t b = it.x;
  for (uint32_t q = threadIdx.x; q < kBinRows; q += kBT) lacc[q] = 0ull;
  __syncthreads();
  // 8 consecutive entries per lane from an 8-aligned start (the arrays are padded), runs of
  // one row added in registers before one LDS atomic
  for (uint64_t q = (uint64_t)(it.y & ~7u) + 8u * threadIdx.x; q < it.z; q += 8u * kBT) {
    const uint4 rw = *reinterpret_cast<const uint4*>(P.brow + q);
    const float4 va = *reinterpret_cast<const float4*>(P.vals + q);
    const float4 vb = *reinterpret_cast<const float4*>(P.vals + q + 4);
    const uint32_t rows[8] = {rw.x & 0xFFFFu, rw.x >> 16, rw.y & 0xFFFFu, rw.y >> 16,
                              rw.z & 0xFFFFu, rw.z >> 16, rw.w & 0xFFFFu, rw.w >> 16};
    const float vs[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    uint32_t cur = 0xFFFFFFFFu;
    unsigned long long run = 0ull;
#pragma unroll
    for (uint32_t r = 0; r < 8; ++r) {
      const uint64_t e = q + r;
      if (e < it.y || e >= it.z) continue;
      const unsigned long long f = sum_to_fixed(vs[r]);
      if (rows[r] == cur) {
        run += f;
      } else {
        if (cur != 0xFFFFFFFFu) atomicAdd(&lacc[cur], run);
        cur = rows[r];
        run = f;
      }
    }
    if (cur != 0xFFFFFFFFu) atomicAdd(&lacc[cur], run);
  }
  __syncthreads();
  const uint32_t r0 = b << kBinBits;
  double e = 0.0;
  if (it.w == 1) {  // the bin is one item: finish its rows from LDS
    for (uint32_t v = threadIdx.x; v < kBinRows && r0 + v < P.n; v += kBT) e += finish_row(P, r0 + v, fixed_to_sum(lacc[v]));
  } else {
    for (uint32_t v = threadIdx.x; v < kBinRows; v += kBT)
      if (lacc[v]) atomicAdd(P.gacc + r0 + v, lacc[v]);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) is_last = (atomicAdd(P.bin_done + b, 1u) == it.w - 1) ? 1u : 0u;
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    for (uint32_t v = threadIdx.x; v < kBinRows && r0 + v < P.n; v += kBT) {
      const unsigned long long g = __hip_atomic_load(P.gacc + r0 + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      P.gacc[r0 + v] = 0ull;  // zero again for the next sweep
      e += finish_row(P, r0 + v, fixed_to_sum(g));
    }
    if (threadIdx.x == 0) P.bin_done[b] = 0u;
  }
  e = block_sum_d(e, red);
  if (threadIdx.x == 0 && e > 0.0) atomicAdd(P.acc + P.iter, err_to_fixed(e));
}

// ---- layout construction ----
// tile key of CSC position p (row v, source u) and its packed low bits
__global__ __launch_bounds__(256) void bin_key_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ rp,
                                                      const uint32_t* __restrict__ col, uint32_t n_chunks,
                                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ low) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= nnz) return;
  uint32_t lo = 0, hi = n;  // row of p: first r with rp[r + 1] > p
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rp[mid + 1] <= p) lo = mid + 1;
    else hi = mid;
  }
  const uint32_t u = col[p];
  keys[p] = (uint64_t)(lo >> kBinBits) * n_chunks + (u >> kChunkBits);
  low[p] = ((lo & (kBinRows - 1)) << 16) | (u & (kChunkSrc - 1));
}

__global__ __launch_bounds__(256) void bin_split_kernel(uint64_t nnz, const uint32_t* __restrict__ low,
                                                        uint16_t* __restrict__ brow, uint16_t* __restrict__ bsrc) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= nnz) return;
  brow[p] = (uint16_t)(low[p] >> 16);
  bsrc[p] = (uint16_t)(low[p] & 0xFFFFu);
}

// tile_ptr[t] = first sorted position with key >= t, for t in [0, n_tiles]
__global__ __launch_bounds__(256) void tile_ptr_kernel(uint64_t nnz, const uint64_t* __restrict__ keys,
                                                       uint64_t n_tiles, uint32_t* __restrict__ tile_ptr) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t > n_tiles) return;
  uint64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (keys[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  tile_ptr[t] = (uint32_t)lo;
}

inline uint32_t grid256(uint64_t m) { return (uint32_t)ceil_div(m ? m : 1, 256); }

}  // namespace

void binned_build(hipStream_t s, const ActiveCsc& a, uint32_t n, uint64_t nnz, BinnedLayout& L) {
  NWG_REQUIRE(nnz < (1ull << 32) - 16, "binned layout needs fewer than 2^32 - 16 entries");
  L = BinnedLayout();
  L.n_bins = (uint32_t)ceil_div(n, kBinRows);
  L.n_chunks = (uint32_t)ceil_div(n, kChunkSrc);
  L.nnz = nnz;
  L.padded = ceil_div(nnz, 8) * 8 + 8;
  const uint64_t n_tiles = (uint64_t)L.n_bins * L.n_chunks;
  L.brow.alloc(L.padded);
  L.bsrc.alloc(L.padded);
  NWG_HIP(hipMemsetAsync(L.brow.p, 0, L.padded * 2, s));
  NWG_HIP(hipMemsetAsync(L.bsrc.p, 0, L.padded * 2, s));
  std::vector<uint32_t> tp(n_tiles + 1, 0);
  if (nnz) {
    DevBuf<uint64_t> keys(nnz), keys_tmp(nnz);
    DevBuf<uint32_t> low(nnz), low_tmp(nnz);
    bin_key_kernel<<<grid256(nnz), 256, 0, s>>>(n, nnz, a.ptr, a.idx, L.n_chunks, keys.p, low.p);
    NWG_LAUNCH_CHECK();
    int kb = 1;
    while (kb < 64 && (1ull << kb) < n_tiles) ++kb;
    // stable: entries of one tile keep their CSC order (row, then source)
    radix_sort_pairs(s, keys.p, keys_tmp.p, low.p, low_tmp.p, nnz, kb);
    bin_split_kernel<<<grid256(nnz), 256, 0, s>>>(nnz, low.p, L.brow.p, L.bsrc.p);
    NWG_LAUNCH_CHECK();
    DevBuf<uint32_t> d_tp(n_tiles + 1);
    tile_ptr_kernel<<<grid256(n_tiles + 1), 256, 0, s>>>(nnz, keys.p, n_tiles, d_tp.p);
    NWG_LAUNCH_CHECK();
    NWG_HIP(hipMemcpyAsync(tp.data(), d_tp.p, (n_tiles + 1) * 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
  }
  const uint32_t C = L.n_chunks;
  // bin pass items: every bin, cut into ranges of at most kItem2Entries entries
  std::vector<uint4> it2;
  for (uint32_t b = 0; b < L.n_bins; ++b) {
    const uint32_t p0 = tp[(uint64_t)b * C], p1 = tp[(uint64_t)(b + 1) * C];
    const uint32_t cnt = (uint32_t)std::max<uint64_t>(ceil_div(p1 - p0, kItem2Entries), 1);
    for (uint32_t q = 0; q < cnt; ++q) {
      const uint32_t a0 = p0 + (uint32_t)((uint64_t)(p1 - p0) * q / cnt);
      const uint32_t a1 = p0 + (uint32_t)((uint64_t)(p1 - p0) * (q + 1) / cnt);
      it2.push_back(make_uint4(b, a0, a1, cnt));
    }
  }
  // chunk pass items: the non-empty tiles of each chunk, cut at kItem1Entries entries or
  // kBinMaxSegs tiles (a tile longer than kItem1Entries is an item of its own)
  std::vector<uint4> it1;
  std::vector<uint32_t> sst, sfl;
  for (uint32_t c = 0; c < C; ++c) {
    uint32_t first = (uint32_t)sst.size(), flat = 0;
    for (uint32_t b = 0; b < L.n_bins; ++b) {
      const uint32_t t0 = tp[(uint64_t)b * C + c], t1 = tp[(uint64_t)b * C + c + 1];
      if (t1 == t0) continue;
      const uint32_t len = t1 - t0;
      if ((uint32_t)sst.size() > first &&
          (flat + len > kItem1Entries || (uint32_t)sst.size() - first >= kBinMaxSegs)) {
        it1.push_back(make_uint4(c, first, (uint32_t)sst.size(), flat));
        first = (uint32_t)sst.size();
        flat = 0;
      }
      sst.push_back(t0);
      sfl.push_back(flat);
      flat += len;
    }
    if ((uint32_t)sst.size() > first) it1.push_back(make_uint4(c, first, (uint32_t)sst.size(), flat));
  }
  L.n_items2 = (uint32_t)it2.size();
  L.n_items1 = (uint32_t)it1.size();
  L.n_segs = (uint32_t)sst.size();
  L.items2.alloc(std::max<size_t>(it2.size(), 1));
  L.items1.alloc(std::max<size_t>(it1.size(), 1));
  L.seg_start.alloc(std::max<size_t>(sst.size(), 1));
  L.seg_flat.alloc(std::max<size_t>(sfl.size(), 1));
  if (!it2.empty()) NWG_HIP(hipMemcpyAsync(L.items2.p, it2.data(), it2.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
  if (!it1.empty()) NWG_HIP(hipMemcpyAsync(L.items1.p, it1.data(), it1.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
  if (!sst.empty()) {
    NWG_HIP(hipMemcpyAsync(L.seg_start.p, sst.data(), sst.size() * 4, hipMemcpyHostToDevice, s));
    NWG_HIP(hipMemcpyAsync(L.seg_flat.p, sfl.data(), sfl.size() * 4, hipMemcpyHostToDevice, s));
  }
  NWG_HIP(hipStreamSynchronize(s));
  L.built = true;
}

void binned_enqueue_sweep(hipStream_t s, const BinParams& P) {
  if (P.n_items1) pr_chunk_kernel<<<P.n_items1, kBT, 0, s>>>(P);
  pr_bin_kernel<<<P.n_items2, kBT, 0, s>>>(P);
  NWG_LAUNCH_CHECK();
}

}  // namespace nwg
