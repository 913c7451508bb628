// device_utils.hpp -- wave64 / workgroup collectives used by the kernels.
// All reductions use a fixed combination tree, so results are run-to-run
// deterministic for a fixed launch shape.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nwg {

// Native 4 x 32-bit vector (one 16-byte load or store); the nontemporal builtins take it.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Butterfly sum over the low `width` lanes of each width-aligned group (width <= 64).
template <int Width, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = Width / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, Width);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  return group_sum<64>(v);
}

// Sum over a workgroup of NT threads (NT multiple of 64); result valid in thread 0.
template <int NT, class T>
__device__ __forceinline__ T block_sum(T v, T* lds /* NT/64 entries */) {
  v = wave_sum(v);
  const int w = threadIdx.x / 64;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  T r = T(0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r += lds[i];
  }
  __syncthreads();
  return r;
}

// Inclusive prefix sum across the wave.
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(v, o, 64);
    if (l >= (uint32_t)o) v += y;
  }
  return v;
}

// Exclusive prefix sum across a workgroup of NT threads; *total gets the block sum.
template <int NT, class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds /* NT/64 + 1 entries */, T* total) {
  constexpr int NW = NT / 64;
  const int w = threadIdx.x / 64;
  T inc = wave_inclusive_scan(v);
  if (lane_id() == 63) lds[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = T(0);
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      T t = lds[i];
      lds[i] = run;
      run += t;
    }
    lds[NW] = run;
  }
  __syncthreads();
  T r = inc - v + lds[w];
  *total = lds[NW];
  __syncthreads();
  return r;
}

// Fixed-point accumulation of the L1 change: integer atomics make the sum
// independent of block completion order (resolution 2^-58, range [0, 32)).
constexpr double kErrScale = 288230376151711744.0;  // 2^58
__device__ __forceinline__ unsigned long long err_to_fixed(double e) {
  if (!(e >= 0.0)) e = 31.0;  // NaN or negative cannot occur; saturate defensively
  if (e > 31.0) e = 31.0;
  return (unsigned long long)__double2ull_rn(e * kErrScale);
}
__host__ __device__ inline double fixed_to_err(unsigned long long q) { return (double)q / kErrScale; }

}  // namespace nwg
