// gen.hip -- synthetic graph generators on the device (RMAT, 2D grid) and edge weights.
//
// Every output element is a pure function of (index, seed) through Philox4x32-10,
// so any subrange can be regenerated and the host restatement produces the same
// edges bit for bit.
//
// RMAT edge e: level l (most significant bit first) takes word l%4 of
// Philox({e.lo, e.hi, l/4, 'RMAT'}, seed) and picks quadrant A=(0,0) if r < tA,
// B=(0,1) if r < tAB, C=(1,0) if r < tABC, else D=(1,1), as (source bit, target bit).
// No noise and no vertex permutation.
#include <algorithm>

#include "device_utils.hpp"
#include "nwg_internal.hpp"
#include "philox.hpp"

namespace nwg {

namespace {

__global__ __launch_bounds__(256) void rmat_kernel(uint32_t scale, uint64_t tA, uint64_t tAB, uint64_t tABC,
                                                   uint64_t seed, uint64_t first, uint64_t count,
                                                   uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint64_t e = first + k;
  uint32_t s = 0, d = 0;
  for (uint32_t l0 = 0; l0 < scale; l0 += 4) {
    const U32x4 r = philox_idx(e, l0 >> 2, kTagRmat, seed);
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      const uint32_t l = l0 + q;
      if (l < scale) {
        const uint64_t x = r.v[q];
        const uint32_t sb = x >= tAB ? 1u : 0u;
        const uint32_t db = x < tA ? 0u : (x < tAB ? 1u : (x < tABC ? 0u : 1u));
        s |= sb << (scale - 1 - l);
        d |= db << (scale - 1 - l);
      }
    }
  }
  src[k] = s;
  dst[k] = d;
}

// Lattice edge k -> slots 2k, 2k+1; long-range edge of vertex v -> slot 2S + v.
__global__ __launch_bounds__(256) void grid_kernel(uint64_t W, uint64_t tk, uint64_t tl, uint64_t seed,
                                                   uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  const uint64_t H = W * (W - 1), S = 2 * H, N = W * W;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < S) {
    const uint64_t k = t;
    uint64_t a, b;
    if (k < H) {
      const uint64_t row = k / (W - 1), col = k % (W - 1);
      a = row * W + col;
      b = a + 1;
    } else {
      const uint64_t kk = k - H, row = kk / W, col = kk % W;
      a = row * W + col;
      b = a + W;
    }
    const U32x4 r = philox_idx(k, 0, kTagGrid, seed);
    if (r.v[0] < tk) {
      src[2 * k] = (uint32_t)a;
      dst[2 * k] = (uint32_t)b;
      src[2 * k + 1] = (uint32_t)b;
      dst[2 * k + 1] = (uint32_t)a;
    } else {
      src[2 * k] = dst[2 * k] = src[2 * k + 1] = dst[2 * k + 1] = 0xFFFFFFFFu;
    }
  } else if (t < S + N) {
    const uint64_t v = t - S;
    const U32x4 r = philox_idx(v, 1, kTagGrid, seed);
    if (r.v[0] < tl) {
      const uint64_t off = ((uint64_t)r.v[1] * (N - 1)) >> 32;
      src[2 * S + v] = (uint32_t)v;
      dst[2 * S + v] = (uint32_t)((v + 1 + off) % N);
    } else {
      src[2 * S + v] = dst[2 * S + v] = 0xFFFFFFFFu;
    }
  }
}

// w(u, v) = fma(0.9, t, 0.1), t = (Philox({u, v, 3, 'WGHT'}, seed)[0] >> 8) * 2^-24, one thread per row.
__global__ __launch_bounds__(256) void weight_kernel(uint32_t n, const uint32_t* __restrict__ ptr,
                                                     const uint32_t* __restrict__ idx, uint64_t seed,
                                                     float* __restrict__ w) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  for (uint32_t p = ptr[v]; p < ptr[v + 1]; ++p) {
    const uint32_t u = idx[p];
    const U32x4 r = philox4x32_10(u, v, 3u, kTagWght, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float t = (float)(r.v[0] >> 8) * 0x1p-24f;
    w[p] = __builtin_fmaf(0.9f, t, 0.1f);
  }
}

}  // namespace

void rmat_edges_device(hipStream_t s, uint32_t scale, uint32_t a, uint32_t b, uint32_t c, uint64_t seed,
                       uint64_t first, uint64_t count, uint32_t* src, uint32_t* dst) {
  NWG_REQUIRE(scale >= 1 && scale <= 31, "rmat scale must be in [1, 31]");
  NWG_REQUIRE((uint64_t)a + b + c <= 10000, "rmat a + b + c must not exceed 10000");
  if (!count) return;
  const uint64_t tA = per10000_threshold(a), tAB = per10000_threshold(a + b), tABC = per10000_threshold(a + b + c);
  rmat_kernel<<<(uint32_t)ceil_div(count, 256), 256, 0, s>>>(scale, tA, tAB, tABC, seed, first, count, src, dst);
  NWG_LAUNCH_CHECK();
}

void grid_edges_device(hipStream_t s, uint32_t width, uint32_t keep, uint32_t lng, uint64_t seed, uint32_t* src,
                       uint32_t* dst) {
  NWG_REQUIRE(width >= 2 && (uint64_t)width * width < (1ull << 32), "grid width must be in [2, 65535]");
  const uint64_t W = width, total = 2 * W * (W - 1) + W * W;
  grid_kernel<<<(uint32_t)ceil_div(total, 256), 256, 0, s>>>(W, per10000_threshold(keep), per10000_threshold(lng),
                                                              seed, src, dst);
  NWG_LAUNCH_CHECK();
}

void edge_weights_device(hipStream_t s, uint32_t n, const uint32_t* csc_ptr, const uint32_t* csc_idx, uint64_t seed,
                         float* w) {
  if (!n) return;
  weight_kernel<<<(uint32_t)ceil_div(n, 256), 256, 0, s>>>(n, csc_ptr, csc_idx, seed, w);
  NWG_LAUNCH_CHECK();
}

}  // namespace nwg

// ---------------------------------------------------------------------------
// calibration copy: 16-byte vector loads and stores, grid-stride over the buffer
// ---------------------------------------------------------------------------
namespace nwg {
namespace {
// four independent 16-byte loads in flight per lane, then four stores
__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                          uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}
}  // namespace

// Bandwidth probes: the device's streaming ceilings for a copy, a read-only pass and a
// write-only pass, each with or without nontemporal loads / stores.  Four 16-byte accesses
// in flight per lane, grid-stride.  The read-only pass folds what it reads into one value and
// stores it only when it equals a value it never takes, so the loads cannot be dropped.
template <bool kNtLoad>
__device__ __forceinline__ u32x4 probe_ld(const u32x4* p) {
  if constexpr (kNtLoad) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool kNtStore>
__device__ __forceinline__ void probe_st(u32x4* p, u32x4 v) {
  if constexpr (kNtStore) __builtin_nontemporal_store(v, p);
  else *p = v;
}
// kKind 3 reads the buffer kRereads times in one launch (a cache-resident buffer then streams
// from the cache, not from HBM).
constexpr uint32_t kRereads = 16;
template <int kKind, bool kNtLoad, bool kNtStore>
__global__ __launch_bounds__(256) void stream_probe_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                           uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  u32x4 acc = {0u, 0u, 0u, 0u};
  if constexpr (kKind == 3) {
    for (uint32_t r = 0; r < kRereads; ++r)
      for (uint64_t q = i; q < n16; q += stride) acc ^= probe_ld<kNtLoad>(src + q);
    if (acc.x == 0x9E3779B9u && acc.y == 0x7F4A7C15u && acc.z == 0x85EBCA6Bu && acc.w == 0xC2B2AE35u) dst[0] = acc;
    return;
  }
  const u32x4 fill = {(uint32_t)i, 1u, 2u, 3u};
  for (; i + 3 * stride < n16; i += 4 * stride) {
    if constexpr (kKind == 2) {  // write only
      probe_st<kNtStore>(dst + i, fill);
      probe_st<kNtStore>(dst + i + stride, fill);
      probe_st<kNtStore>(dst + i + 2 * stride, fill);
      probe_st<kNtStore>(dst + i + 3 * stride, fill);
    } else {
      const u32x4 a = probe_ld<kNtLoad>(src + i), b = probe_ld<kNtLoad>(src + i + stride),
                  c = probe_ld<kNtLoad>(src + i + 2 * stride), d = probe_ld<kNtLoad>(src + i + 3 * stride);
      if constexpr (kKind == 0) {  // copy
        probe_st<kNtStore>(dst + i, a);
        probe_st<kNtStore>(dst + i + stride, b);
        probe_st<kNtStore>(dst + i + 2 * stride, c);
        probe_st<kNtStore>(dst + i + 3 * stride, d);
      } else {  // read only
        acc ^= a ^ b ^ c ^ d;
      }
    }
  }
  for (; i < n16; i += stride) {
    if constexpr (kKind == 2) probe_st<kNtStore>(dst + i, fill);
    else if constexpr (kKind == 0) probe_st<kNtStore>(dst + i, probe_ld<kNtLoad>(src + i));
    else acc ^= probe_ld<kNtLoad>(src + i);
  }
  if constexpr (kKind == 1)
    if (acc.x == 0x9E3779B9u && acc.y == 0x7F4A7C15u && acc.z == 0x85EBCA6Bu && acc.w == 0xC2B2AE35u)
      dst[0] = acc;
}

template <int kKind>
static void launch_probe(hipStream_t s, uint32_t grid, uint32_t flags, const u32x4* src, u32x4* dst, uint64_t n16) {
  switch (flags & 3u) {
    case 0: stream_probe_kernel<kKind, false, false><<<grid, 256, 0, s>>>(src, dst, n16); break;
    case 1: stream_probe_kernel<kKind, true, false><<<grid, 256, 0, s>>>(src, dst, n16); break;
    case 2: stream_probe_kernel<kKind, false, true><<<grid, 256, 0, s>>>(src, dst, n16); break;
    default: stream_probe_kernel<kKind, true, true><<<grid, 256, 0, s>>>(src, dst, n16); break;
  }
}

float stream_probe_device(hipStream_t s, hipEvent_t e0, hipEvent_t e1, uint32_t kind, uint32_t flags,
                          uint32_t wg_per_cu, void* dst, const void* src, uint64_t bytes) {
  NWG_REQUIRE(kind <= 3, "probe kind must be 0 (copy), 1 (read), 2 (write) or 3 (read 16 times)");
  NWG_REQUIRE(bytes % 16 == 0 && bytes > 0, "probe size must be a positive multiple of 16 bytes");
  NWG_REQUIRE(((uintptr_t)dst | (uintptr_t)src) % 16 == 0, "probe buffers must be 16-byte aligned");
  NWG_REQUIRE(dst != nullptr && (kind == 2 || src != nullptr), "probe buffers must not be null");
  NWG_REQUIRE(wg_per_cu >= 1 && wg_per_cu <= 32, "workgroups per CU must be in [1, 32]");
  const uint64_t n16 = bytes / 16;
  int dev = 0, cus = 0;
  NWG_HIP(hipGetDevice(&dev));
  NWG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n16, 256), (uint64_t)cus * wg_per_cu);
  const u32x4* sp = static_cast<const u32x4*>(src);
  u32x4* dp = static_cast<u32x4*>(dst);
  NWG_HIP(hipEventRecord(e0, s));
  if (kind == 0) launch_probe<0>(s, grid, flags, sp, dp, n16);
  else if (kind == 1) launch_probe<1>(s, grid, flags, sp, dp, n16);
  else if (kind == 2) launch_probe<2>(s, grid, flags, sp, dp, n16);
  else launch_probe<3>(s, grid, flags, sp, dp, n16);
  NWG_LAUNCH_CHECK();
  NWG_HIP(hipEventRecord(e1, s));
  NWG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NWG_HIP(hipEventElapsedTime(&ms, e0, e1));
  return ms;
}

float stream_copy_device(hipStream_t s, hipEvent_t e0, hipEvent_t e1, void* dst, const void* src, uint64_t bytes) {
  NWG_REQUIRE(bytes % 16 == 0, "copy size must be a multiple of 16 bytes");
  NWG_REQUIRE(((uintptr_t)dst | (uintptr_t)src) % 16 == 0, "copy buffers must be 16-byte aligned");
  const uint64_t n16 = bytes / 16;
  int dev = 0, cus = 0;
  NWG_HIP(hipGetDevice(&dev));
  NWG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n16 ? n16 : 1, 256), (uint64_t)cus * 8);
  NWG_HIP(hipEventRecord(e0, s));
  stream_copy_kernel<<<grid, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
  NWG_LAUNCH_CHECK();
  NWG_HIP(hipEventRecord(e1, s));
  NWG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  NWG_HIP(hipEventElapsedTime(&ms, e0, e1));
  return ms;
}
}  // namespace nwg
