// prims.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Scan: reduce-then-scan over 2048-element tiles (256 threads x 8 items), recursing
// on the tile sums.  Radix sort: 8-bit digits; per pass an upsweep writes each
// tile's 256-bin histogram, a scan of the digit-major histogram gives every
// (digit, tile) its output base, and the scatter pass ranks keys stably inside the
// tile with wave64 ballot matching, stages the tile in LDS in digit order and
// writes it out as contiguous per-digit runs.
#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, uint64_t n,
                                                                   uint32_t* __restrict__ sums) {
  __shared__ uint32_t lds[kScanThreads / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    uint64_t i = base + (uint64_t)j * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = block_sum<kScanThreads>(s, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

// out[i] = offsets[block] + exclusive prefix inside the tile; the last tile writes out[n].
__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(const uint32_t* __restrict__ in,
                                                                  uint32_t* __restrict__ out, uint64_t n,
                                                                  const uint32_t* __restrict__ offsets) {
  __shared__ uint32_t lds[kScanThreads / 64 + 1];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    uint64_t i = base + j;
    v[j] = i < n ? in[i] : 0u;
    s += v[j];
  }
  uint32_t total;
  uint32_t pre = block_exclusive_scan<kScanThreads>(s, lds, &total);
  const uint32_t off = offsets ? offsets[blockIdx.x] : 0u;
  pre += off;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    uint64_t i = base + j;
    if (i < n) out[i] = pre;
    pre += v[j];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + total;
}

}  // namespace

void exclusive_scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, uint64_t n) {
  const uint64_t nb = ceil_div(n ? n : 1, kScanTile);
  if (nb == 1) {
    scan_tiles_kernel<<<1, kScanThreads, 0, s>>>(in, out, n, nullptr);
    NWG_LAUNCH_CHECK();
    return;
  }
  DevBuf<uint32_t> sums(nb), offs(nb + 1);
  scan_reduce_kernel<<<(uint32_t)nb, kScanThreads, 0, s>>>(in, n, sums.p);
  NWG_LAUNCH_CHECK();
  exclusive_scan_u32(s, sums.p, offs.p, nb);
  scan_tiles_kernel<<<(uint32_t)nb, kScanThreads, 0, s>>>(in, out, n, offs.p);
  NWG_LAUNCH_CHECK();
  // the temporaries are freed on return: wait for the kernels that read them
  NWG_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
namespace {

constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;
constexpr int kRsWaves = kRsThreads / 64;

template <class K>
__global__ __launch_bounds__(kRsThreads) void rs_upsweep_kernel(const K* __restrict__ keys, uint64_t n,
                                                                int shift, uint32_t nb,
                                                                uint32_t* __restrict__ tile_counts,
                                                                uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
#pragma unroll 4
  for (int j = 0; j < kRsItems; ++j) {
    uint64_t i = base + (uint64_t)j * kRsThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  const uint32_t c = cnt[threadIdx.x];
  tile_counts[(uint64_t)blockIdx.x * 256 + threadIdx.x] = c;
  hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = c;
}

template <class K, bool HasVals>
__global__ __launch_bounds__(kRsThreads) void rs_scatter_kernel(const K* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in,
                                                                K* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out, uint64_t n,
                                                                int shift, uint32_t nb,
                                                                const uint32_t* __restrict__ tile_counts,
                                                                const uint32_t* __restrict__ hist_scan) {
  __shared__ K sk[kRsTile];
  __shared__ uint32_t sv[HasVals ? kRsTile : 1];
  __shared__ uint32_t wave_cnt[kRsWaves][256];
  __shared__ uint32_t lofs[256];
  __shared__ uint32_t gofs[256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t scan_lds[kRsThreads / 64 + 1];

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t w = tid / 64;
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  const uint32_t tile_n = (uint32_t)((n - base) < (uint64_t)kRsTile ? (n - base) : (uint64_t)kRsTile);

  {
    const uint32_t c = tile_counts[(uint64_t)blockIdx.x * 256 + tid];
    uint32_t total;
    const uint32_t pre = block_exclusive_scan<kRsThreads>(c, scan_lds, &total);
    lofs[tid] = pre;
    gofs[tid] = hist_scan[(uint64_t)tid * nb + blockIdx.x];
    running[tid] = 0;
#pragma unroll
    for (int q = 0; q < kRsWaves; ++q) wave_cnt[q][tid] = 0;
  }
  __syncthreads();

  const uint64_t lt_mask = __lanemask_lt();
  for (int j = 0; j < kRsItems; ++j) {
    const uint32_t li = (uint32_t)j * kRsThreads + tid;
    const bool valid = li < tile_n;
    K key = 0;
    uint32_t val = 0;
    if (valid) {
      key = keys_in[base + li];
      if (HasVals) val = vals_in[base + li];
    }
    const uint32_t dg = (uint32_t)(key >> shift) & 0xFFu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (dg >> bit) & 1u;
      const uint64_t bb = __ballot(set);
      peers &= set ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank == 0) wave_cnt[w][dg] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pre = running[dg] + rank;
      for (uint32_t q = 0; q < w; ++q) pre += wave_cnt[q][dg];
      const uint32_t lp = lofs[dg] + pre;
      sk[lp] = key;
      if (HasVals) sv[lp] = val;
    }
    __syncthreads();
    {
      uint32_t add = 0;
#pragma unroll
      for (int q = 0; q < kRsWaves; ++q) {
        add += wave_cnt[q][tid];
        wave_cnt[q][tid] = 0;
      }
      running[tid] += add;
    }
    __syncthreads();
  }
  (void)lane;

  for (int j = 0; j < kRsItems; ++j) {
    const uint32_t li = (uint32_t)j * kRsThreads + tid;
    if (li < tile_n) {
      const K key = sk[li];
      const uint32_t dg = (uint32_t)(key >> shift) & 0xFFu;
      const uint64_t o = (uint64_t)gofs[dg] + (li - lofs[dg]);
      keys_out[o] = key;
      if (HasVals) vals_out[o] = sv[li];
    }
  }
}

}  // namespace

void radix_sort_pairs(hipStream_t s, uint64_t* keys, uint64_t* keys_tmp, uint32_t* vals, uint32_t* vals_tmp,
                      uint64_t n, int key_bits) {
  if (n < 2 || key_bits <= 0) return;
  NWG_REQUIRE(n < (1ull << 32), "radix sort supports fewer than 2^32 keys");
  const uint64_t nb64 = ceil_div(n, kRsTile);
  const uint32_t nb = (uint32_t)nb64;
  DevBuf<uint32_t> tile_counts(nb64 * 256), hist(nb64 * 256), hist_scan(nb64 * 256 + 1);
  uint64_t *ki = keys, *ko = keys_tmp;
  uint32_t *vi = vals, *vo = vals_tmp;
  int passes = 0;
  for (int shift = 0; shift < key_bits; shift += 8) {
    rs_upsweep_kernel<uint64_t><<<nb, kRsThreads, 0, s>>>(ki, n, shift, nb, tile_counts.p, hist.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, hist.p, hist_scan.p, nb64 * 256);
    if (vals)
      rs_scatter_kernel<uint64_t, true><<<nb, kRsThreads, 0, s>>>(ki, vi, ko, vo, n, shift, nb, tile_counts.p,
                                                                  hist_scan.p);
    else
      rs_scatter_kernel<uint64_t, false><<<nb, kRsThreads, 0, s>>>(ki, vi, ko, vo, n, shift, nb, tile_counts.p,
                                                                   hist_scan.p);
    NWG_LAUNCH_CHECK();
    std::swap(ki, ko);
    std::swap(vi, vo);
    ++passes;
  }
  if (passes & 1) {
    NWG_HIP(hipMemcpyAsync(keys, ki, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    if (vals) NWG_HIP(hipMemcpyAsync(vals, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
  NWG_HIP(hipStreamSynchronize(s));
}

void radix_sort_keys_u32(hipStream_t s, uint32_t* keys, uint32_t* keys_tmp, uint64_t n, int first_bit, int end_bit) {
  if (n < 2 || end_bit <= first_bit) return;
  NWG_REQUIRE(n < (1ull << 32), "radix sort supports fewer than 2^32 keys");
  NWG_REQUIRE(first_bit >= 0 && end_bit <= 32, "key bits must lie in [0, 32)");
  const uint64_t nb64 = ceil_div(n, kRsTile);
  const uint32_t nb = (uint32_t)nb64;
  DevBuf<uint32_t> tile_counts(nb64 * 256), hist(nb64 * 256), hist_scan(nb64 * 256 + 1);
  uint32_t *ki = keys, *ko = keys_tmp;
  int passes = 0;
  for (int shift = first_bit; shift < end_bit; shift += 8) {
    rs_upsweep_kernel<uint32_t><<<nb, kRsThreads, 0, s>>>(ki, n, shift, nb, tile_counts.p, hist.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, hist.p, hist_scan.p, nb64 * 256);
    rs_scatter_kernel<uint32_t, false><<<nb, kRsThreads, 0, s>>>(ki, nullptr, ko, nullptr, n, shift, nb,
                                                                 tile_counts.p, hist_scan.p);
    NWG_LAUNCH_CHECK();
    std::swap(ki, ko);
    ++passes;
  }
  if (passes & 1) NWG_HIP(hipMemcpyAsync(keys, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  NWG_HIP(hipStreamSynchronize(s));
}

}  // namespace nwg
