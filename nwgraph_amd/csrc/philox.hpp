// philox.hpp -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), host and device.
// Used so that the device generators and any host code draw identical streams
// keyed by (index, seed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nwg {

struct U32x4 {
  uint32_t v[4];
};

__host__ __device__ __forceinline__ U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                         uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return U32x4{{c0, c1, c2, c3}};
}

// Stream tags (ASCII) separating the generators' counter spaces.
constexpr uint32_t kTagRmat = 0x524D4154u;  // 'RMAT'
constexpr uint32_t kTagGrid = 0x47524944u;  // 'GRID'
constexpr uint32_t kTagPprs = 0x50505253u;  // 'PPRS'
constexpr uint32_t kTagWght = 0x57474854u;  // 'WGHT'

__host__ __device__ __forceinline__ U32x4 philox_idx(uint64_t a, uint32_t b, uint32_t tag, uint64_t seed) {
  return philox4x32_10((uint32_t)a, (uint32_t)(a >> 32), b, tag, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// probability p/10000 as a 33-bit threshold on a uniform 32-bit word
__host__ __device__ __forceinline__ uint64_t per10000_threshold(uint32_t p) { return ((uint64_t)p << 32) / 10000u; }

}  // namespace nwg
