// wpr.hip -- weighted PageRank forward, implicit adjoint and edge gradient.
//
// Forward (same fixed-point form as the reference loop, page_rank.hpp:121-147, with a
// per-edge transition weight instead of 1/outdeg and a teleport vector v):
//   x_i <- fma(d, z_i, (1-d) v_i),   z_i = sum_{p in CSC row i} w[p] * x[col[p]]
// Jacobi, x double-buffered, stop at the first sweep with L1 change < threshold.
// Adjoint of L(x) with g = dL/dx at the fixed point x = (1-d) v + d A x, A[i][u] = w(u->i):
//   lambda = g + d A^T lambda   -- solved by the same sweep in pull form over the CSR
// (row u gathers its out-edges u -> i: lambda_u <- fma(d, sum w(u->i) lambda_i, g_u)), so
// no unrolled forward iterate is stored and no atomics are needed.  Then
//   dL/dw(u->i) = d * lambda_i * x_u,   dL/dv = (1-d) lambda,
// and, for raw weights normalised per source (w = w~ / s_u, s_u = sum_k w~(u->k)),
//   dL/dw~(u->k) = (dL/dw(u->k) - sum_j dL/dw(u->j) w(u->j)) / s_u.
//
// Sweep schedule: the rows are cut into work items of up to 256 consecutive rows with at
// most kItemNnz nonzeros (values staged in LDS; one lane per short row, one wave per row of
// more than kLaneRow nonzeros) or one longer row
// (whole-workgroup reduction); a persistent grid walks the items round-robin.
//
// The L1 change of a sweep is summed in double with a fixed order: every workgroup stores
// its partial, and the last workgroup to finish (a counter per sweep) adds all partials
// in workgroup order and stores err[k].  The adjoint's changes have no a-priori bound (a
// loss such as -sum log x gives |grad_x| ~ n), so a fixed-point accumulator of bounded
// range could wrap; this sum cannot, and it does not depend on completion order, so the
// sweep count is reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kItemNnz = 4096;
constexpr uint32_t kItemRows = 256;
constexpr uint32_t kGather = 8;    // gathers a lane issues before it uses them
constexpr uint32_t kLaneRow = 16;  // longest row summed by one lane; longer rows take a wave

struct WParams {
  const uint32_t* rp;
  const uint32_t* col;
  const float* w;
  const uint32_t* item_r0;
  const uint32_t* item_r1;
  uint32_t n_items;
  const float* base;  // per row: (1-d) v_i forward, g_u adjoint
  const float* xin;
  float* xout;
  float damp;
  double* part;    // per workgroup: L1 partial of the current sweep
  uint32_t* done;  // per sweep: workgroups that stored their partial
  double* err;     // per sweep: L1 change, stored by the sweep's last workgroup
  uint32_t iter;
  double threshold;
  const double* thr_scale;  // relative threshold: the stop compares with threshold * *thr_scale
};

__device__ __forceinline__ double stop_threshold(const WParams& P) {
  return P.thr_scale ? P.threshold * *P.thr_scale : P.threshold;
}

// Workgroup-order sum of the per-workgroup partials `e` (valid in thread 0) of sweep
// P.iter into P.err[P.iter], by whichever workgroup finishes last.
__device__ __forceinline__ void fold_sweep_error(const WParams& P, double e, double* red_d) {
  __shared__ uint32_t last;
  if (threadIdx.x == 0) {
    P.part[blockIdx.x] = e;
    __threadfence();  // release the partial before counting it
    last = atomicAdd(P.done + P.iter, 1u) == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // acquire every other workgroup's partial
  double s = 0.0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += kThreads)
    s += __hip_atomic_load(P.part + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s = block_sum<kThreads>(s, red_d);
  if (threadIdx.x == 0) P.err[P.iter] = s;
}

__global__ __launch_bounds__(kThreads) void wsweep_kernel(WParams P) {
  if (P.iter > 0 && P.err[P.iter - 1] < stop_threshold(P)) return;
  __shared__ float vals[kItemNnz];
  __shared__ uint32_t lrows[kItemRows];   // an item's long rows (offsets from its first row)
  __shared__ uint32_t wlong[kThreads / 64];
  __shared__ float red_f[kThreads / 64];
  __shared__ double red_d[kThreads / 64];
  double e = 0.0;
  for (uint32_t it = blockIdx.x; it < P.n_items; it += gridDim.x) {
    const uint32_t r0 = P.item_r0[it], r1 = P.item_r1[it];
    const uint32_t p0 = P.rp[r0], p1 = P.rp[r1];
    if (p1 - p0 > kItemNnz) {
      // one long row: a lane's edges p0 + lane + t * kThreads, kGather gathers in flight,
      // added in that order
      float z = 0.f;
      for (uint32_t p = p0 + threadIdx.x; p < p1; p += kThreads * kGather) {
        uint32_t c[kGather];
        float wv[kGather], xv[kGather];
#pragma unroll
        for (uint32_t t = 0; t < kGather; ++t) {
          const uint32_t q = p + t * kThreads;
          c[t] = q < p1 ? P.col[q] : 0u;
          wv[t] = q < p1 ? P.w[q] : 0.f;
        }
#pragma unroll
        for (uint32_t t = 0; t < kGather; ++t) xv[t] = p + t * kThreads < p1 ? P.xin[c[t]] : 0.f;
#pragma unroll
        for (uint32_t t = 0; t < kGather; ++t)
          if (p + t * kThreads < p1) z = __builtin_fmaf(wv[t], xv[t], z);
      }
      z = block_sum<kThreads>(z, red_f);
      if (threadIdx.x == 0) {
        const float xn = __builtin_fmaf(P.damp, z, P.base[r0]);
        e += (double)fabsf(xn - P.xin[r0]);
        P.xout[r0] = xn;
      }
    } else {
      // stage the item's products w * x[col], kGather gathers of a lane in flight
      const uint32_t m = p1 - p0;
      for (uint32_t q0 = threadIdx.x; q0 < m; q0 += kThreads * kGather) {
        uint32_t c[kGather];
        float wv[kGather], xv[kGather];
#pragma unroll
        for (uint32_t t = 0; t < kGather; ++t) {
          const uint32_t q = q0 + t * kThreads;
          c[t] = q < m ? P.col[p0 + q] : 0u;
          wv[t] = q < m ? P.w[p0 + q] : 0.f;
        }
#pragma unroll
        for (uint32_t t = 0; t < kGather; ++t) xv[t] = q0 + t * kThreads < m ? P.xin[c[t]] : 0.f;
#pragma unroll
        for (uint32_t t = 0; t < kGather; ++t)
          if (q0 + t * kThreads < m) vals[q0 + t * kThreads] = wv[t] * xv[t];
      }
      __syncthreads();
      // rows of up to kLaneRow entries: one lane each, in row order; longer rows: one wave
      // each (lane-strided partial sums, then a butterfly), listed in row order through a
      // ballot so that every row's sum is computed the same way on every run
      const uint32_t i = r0 + threadIdx.x;
      uint32_t a = 0, b = 0;
      if (i < r1) {
        a = P.rp[i] - p0;
        b = P.rp[i + 1] - p0;
      }
      const bool longrow = b - a > kLaneRow;
      if (i < r1 && !longrow) {
        float z = 0.f;
        for (uint32_t q = a; q < b; ++q) z += vals[q];
        const float xn = __builtin_fmaf(P.damp, z, P.base[i]);
        e += (double)fabsf(xn - P.xin[i]);
        P.xout[i] = xn;
      }
      const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
      const unsigned long long lm = __ballot(longrow);
      if (lane == 0) wlong[wv] = (uint32_t)__popcll(lm);
      __syncthreads();
      uint32_t off = 0, nlong = 0;
      for (uint32_t q = 0; q < kThreads / 64; ++q) {
        off += q < wv ? wlong[q] : 0u;
        nlong += wlong[q];
      }
      if (longrow)
        lrows[off + __builtin_amdgcn_mbcnt_hi((uint32_t)(lm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)lm, 0u))] =
            threadIdx.x;
      __syncthreads();
      for (uint32_t k = wv; k < nlong; k += kThreads / 64) {
        const uint32_t ri = r0 + lrows[k];
        const uint32_t la = P.rp[ri] - p0, lb = P.rp[ri + 1] - p0;
        float z = 0.f;
        for (uint32_t q = la + lane; q < lb; q += 64) z += vals[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) z += __shfl_xor(z, o, 64);
        if (lane == 0) {
          const float xn = __builtin_fmaf(P.damp, z, P.base[ri]);
          e += (double)fabsf(xn - P.xin[ri]);
          P.xout[ri] = xn;
        }
      }
      __syncthreads();
    }
  }
  e = block_sum<kThreads>(e, red_d);
  fold_sweep_error(P, e, red_d);
}

// y = v (src null) or y = src; the per-sweep counters and errors cleared
__global__ __launch_bounds__(kThreads) void winit_kernel(uint32_t n, const float* __restrict__ src, float v,
                                                         float* __restrict__ dst, uint32_t* __restrict__ done,
                                                         double* __restrict__ err, uint64_t iters) {
  const uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t < n) dst[t] = src ? src[t] : v;
  if (t < iters) {
    done[t] = 0u;
    err[t] = 0.0;
  }
}

// base_i = (float)((1-d) * v_i) (v null: uniform 1/n)
__global__ __launch_bounds__(kThreads) void wbase_kernel(uint32_t n, const float* __restrict__ v, double omd,
                                                         float* __restrict__ base) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) base[i] = v ? (float)(omd * (double)v[i]) : (float)(omd / (double)n);
}

// Per-row sums over the CSR of v[q] = a[perm[q]] * (b ? b[perm[q]] : 1), in CSR order
// (deterministic), on the same LDS work items as the sweep.  FP32 with Kahan compensation:
// the error stays within a few ulp of the sum whatever the row length.
struct Kahan {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

__global__ __launch_bounds__(kThreads) void csr_rowsum_kernel(const uint32_t* __restrict__ rp,
                                                              const uint32_t* __restrict__ perm,
                                                              const float* __restrict__ a,
                                                              const float* __restrict__ b,
                                                              const uint32_t* __restrict__ item_r0,
                                                              const uint32_t* __restrict__ item_r1, uint32_t n_items,
                                                              float* __restrict__ out) {
  __shared__ float vals[kItemNnz];
  __shared__ uint32_t lrows[kItemRows];   // an item's long rows (offsets from its first row)
  __shared__ uint32_t wlong[kThreads / 64];
  __shared__ float red_f[kThreads / 64];
  for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const uint32_t r0 = item_r0[it], r1 = item_r1[it];
    const uint32_t p0 = rp[r0], p1 = rp[r1];
    if (p1 - p0 > kItemNnz) {
      // a single long row: strided compensated partial sums, fixed-tree reduction of the
      // partials and of their compensations
      Kahan k;
      for (uint32_t q = p0 + threadIdx.x; q < p1; q += kThreads) {
        const uint32_t e = perm[q];
        k.add(b ? a[e] * b[e] : a[e]);
      }
      const float hi = block_sum<kThreads>(k.s, red_f);
      __syncthreads();
      const float lo = block_sum<kThreads>(-k.c, red_f);
      if (threadIdx.x == 0) out[r0] = hi + lo;
    } else {
      for (uint32_t q = threadIdx.x; q < p1 - p0; q += kThreads) {
        const uint32_t e = perm[p0 + q];
        vals[q] = b ? a[e] * b[e] : a[e];
      }
      __syncthreads();
      // rows of up to kLaneRow entries: one lane each; longer rows: one wave each (lane-strided
      // compensated partial sums, butterfly sums of the partials and of their compensations),
      // listed in row order through a ballot
      const uint32_t u = r0 + threadIdx.x;
      uint32_t ra = 0, rb = 0;
      if (u < r1) {
        ra = rp[u] - p0;
        rb = rp[u + 1] - p0;
      }
      const bool longrow = rb - ra > kLaneRow;
      if (u < r1 && !longrow) {
        Kahan k;
        for (uint32_t q = ra; q < rb; ++q) k.add(vals[q]);
        out[u] = k.s - k.c;
      }
      const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
      const unsigned long long lm = __ballot(longrow);
      if (lane == 0) wlong[wv] = (uint32_t)__popcll(lm);
      __syncthreads();
      uint32_t off = 0, nlong = 0;
      for (uint32_t q = 0; q < kThreads / 64; ++q) {
        off += q < wv ? wlong[q] : 0u;
        nlong += wlong[q];
      }
      if (longrow)
        lrows[off + __builtin_amdgcn_mbcnt_hi((uint32_t)(lm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)lm, 0u))] =
            threadIdx.x;
      __syncthreads();
      for (uint32_t k2 = wv; k2 < nlong; k2 += kThreads / 64) {
        const uint32_t ri = r0 + lrows[k2];
        const uint32_t la = rp[ri] - p0, lb = rp[ri + 1] - p0;
        Kahan k;
        for (uint32_t q = la + lane; q < lb; q += 64) k.add(vals[q]);
        float hi = k.s, lo = -k.c;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          hi += __shfl_xor(hi, o, 64);
          lo += __shfl_xor(lo, o, 64);
        }
        if (lane == 0) out[ri] = hi + lo;
      }
    }
    __syncthreads();
  }
}

// Scale of a relative threshold, ||b||_1 / (1 - d) for the fixed point's constant term b
// (1 when b is null: the uniform teleport (1 - d) / n summing to 1 - d), or 1 when b is 0.
// Contiguous ranges per workgroup, partials added in workgroup order by the last workgroup
// to finish: deterministic.
constexpr uint32_t kNormBlocks = 256;
__global__ __launch_bounds__(kThreads) void l1_scale_kernel(uint32_t n, const float* __restrict__ b, double inv_omd,
                                                            double* __restrict__ part, uint32_t* __restrict__ done,
                                                            double* __restrict__ scale) {
  __shared__ double red_d[kThreads / 64];
  __shared__ uint32_t last;
  double e = 0.0;
  if (b) {
    const uint64_t per = ((uint64_t)n + gridDim.x - 1) / gridDim.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kThreads) e += (double)fabsf(b[i]);
  }
  e = block_sum<kThreads>(e, red_d);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = e;
    __threadfence();
    last = atomicAdd(done, 1u) == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (uint32_t q = 0; q < gridDim.x; ++q) s += __hip_atomic_load(part + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *scale = !b ? 1.0 : (s > 0.0 ? s * inv_omd : 1.0);
    *done = 0u;
  }
}

// Per workgroup over a contiguous range: sum and max of |b| (the bound a binned weighted sweep
// scales its fixed point by); NaN or infinity shows in the sum.
__global__ __launch_bounds__(kThreads) void abs_stats_kernel(uint32_t n, const float* __restrict__ b,
                                                             double* __restrict__ out) {
  __shared__ double red_d[kThreads / 64];
  const uint64_t per = ((uint64_t)n + gridDim.x - 1) / gridDim.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
  double sum = 0.0, mx = 0.0;
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    const double a = (double)fabsf(b[i]);
    sum += a;
    mx = a > mx ? a : mx;
  }
  sum = block_sum<kThreads>(sum, red_d);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double y = __shfl_xor(mx, o, 64);
    mx = y > mx ? y : mx;
  }
  if ((threadIdx.x & 63) == 0) red_d[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t q = 0; q < kThreads / 64; ++q) mx = red_d[q] > mx ? red_d[q] : mx;
    out[2 * blockIdx.x] = sum;
    out[2 * blockIdx.x + 1] = mx;
  }
}

// *bad = 1 when some weight is negative or not finite
__global__ __launch_bounds__(kThreads) void weight_check_kernel(uint64_t nnz, const float* __restrict__ w,
                                                                uint32_t* __restrict__ bad) {
  const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p < nnz && !(w[p] >= 0.f && w[p] <= 3.402823466e38f)) *bad = 1u;
}

// row of every CSC position (upper bound in the row pointers)
__global__ __launch_bounds__(kThreads) void row_of_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ rp,
                                                          uint32_t* __restrict__ row) {
  const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= nnz) return;
  uint32_t lo = 0, hi = n;  // first r with rp[r + 1] > p
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rp[mid + 1] <= p) lo = mid + 1;
    else hi = mid;
  }
  row[p] = lo;
}

__global__ __launch_bounds__(kThreads) void normalise_kernel(uint64_t nnz, const uint32_t* __restrict__ col,
                                                             const float* __restrict__ w_raw,
                                                             const float* __restrict__ s, float* __restrict__ w) {
  const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p < nnz) w[p] = w_raw[p] / s[col[p]];
}

__global__ __launch_bounds__(kThreads) void gather_kernel(uint64_t nnz, const uint32_t* __restrict__ perm,
                                                          const float* __restrict__ src, float* __restrict__ dst) {
  const uint64_t q = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (q < nnz) dst[q] = src[perm[q]];
}

// grad_w[p] = d * lambda[row(p)] * x[col[p]], one thread per CSC position
__global__ __launch_bounds__(kThreads) void edge_grad_kernel(uint64_t nnz, const uint32_t* __restrict__ row,
                                                             const uint32_t* __restrict__ col, float d,
                                                             const float* __restrict__ lambda,
                                                             const float* __restrict__ x, float* __restrict__ gw) {
  const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p < nnz) gw[p] = (d * lambda[row[p]]) * x[col[p]];
}

__global__ __launch_bounds__(kThreads) void chain_apply_kernel(uint64_t nnz, const uint32_t* __restrict__ col,
                                                               const float* __restrict__ t,
                                                               const float* __restrict__ s, float* __restrict__ gw) {
  const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p < nnz) {
    const uint32_t u = col[p];
    gw[p] = (gw[p] - t[u]) / s[u];
  }
}

__global__ __launch_bounds__(kThreads) void scale_kernel(uint32_t n, float a, const float* __restrict__ src,
                                                         float* __restrict__ dst) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[i] = a * src[i];
}

inline uint32_t grid1(uint64_t m) { return (uint32_t)ceil_div(m ? m : 1, kThreads); }

}  // namespace

// Work items over a compressed structure (host-side, from its row pointers).
void build_items(hipStream_t s, const uint32_t* rp_dev, uint32_t n, DevBuf<uint32_t>& r0, DevBuf<uint32_t>& r1,
                 uint32_t* count) {
  std::vector<uint32_t> rp((size_t)n + 1);
  NWG_HIP(hipMemcpyAsync(rp.data(), rp_dev, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> a, b;
  uint64_t r = 0;
  while (r < n) {
    uint64_t e = r + 1;
    if (rp[e] - rp[r] <= kItemNnz)
      while (e < n && e - r < kItemRows && rp[e + 1] - rp[r] <= kItemNnz) ++e;
    a.push_back((uint32_t)r);
    b.push_back((uint32_t)e);
    r = e;
  }
  r0.alloc(std::max<size_t>(a.size(), 1));
  r1.alloc(std::max<size_t>(b.size(), 1));
  if (!a.empty()) {
    NWG_HIP(hipMemcpyAsync(r0.p, a.data(), a.size() * 4, hipMemcpyHostToDevice, s));
    NWG_HIP(hipMemcpyAsync(r1.p, b.data(), b.size() * 4, hipMemcpyHostToDevice, s));
  }
  NWG_HIP(hipStreamSynchronize(s));
  *count = (uint32_t)a.size();
}

static uint32_t persistent_grid(uint32_t n_items) {
  int dev = 0, cus = 0;
  NWG_HIP(hipGetDevice(&dev));
  NWG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n_items, (uint64_t)cus * 8));
}

static void ensure_csr_items(hipStream_t s, nwg_graph* g) {
  build_csr_device(s, g);
  if (!g->w_items_csr_n && g->n) build_items(s, g->csr_ptr.p, g->n, g->w_items_csr_r0, g->w_items_csr_r1, &g->w_items_csr_n);
}

void weights_set(hipStream_t s, nwg_graph* g, const float* w_dev, bool normalize) {
  const uint64_t nnz = g->nnz;
  ensure_csr_items(s, g);
  g->w_raw.ensure(nnz ? nnz : 1);
  g->w_csc.ensure(nnz ? nnz : 1);
  g->w_csr.ensure(nnz ? nnz : 1);
  g->w_rowsum.ensure(g->n);
  if (nnz) NWG_HIP(hipMemcpyAsync(g->w_raw.p, w_dev, nnz * 4, hipMemcpyDeviceToDevice, s));
  if (normalize) {
    if (g->w_items_csr_n) {
      csr_rowsum_kernel<<<persistent_grid(g->w_items_csr_n), kThreads, 0, s>>>(
          g->csr_ptr.p, g->csr_to_csc.p, g->w_raw.p, nullptr, g->w_items_csr_r0.p, g->w_items_csr_r1.p,
          g->w_items_csr_n, g->w_rowsum.p);
      NWG_LAUNCH_CHECK();
    }
    normalise_kernel<<<grid1(nnz), kThreads, 0, s>>>(nnz, g->csc_idx.p, g->w_raw.p, g->w_rowsum.p, g->w_csc.p);
    NWG_LAUNCH_CHECK();
  } else {
    std::vector<float> ones(g->n, 1.0f);
    NWG_HIP(hipMemcpyAsync(g->w_rowsum.p, ones.data(), (size_t)g->n * 4, hipMemcpyHostToDevice, s));
    if (nnz) NWG_HIP(hipMemcpyAsync(g->w_csc.p, w_dev, nnz * 4, hipMemcpyDeviceToDevice, s));
  }
  gather_kernel<<<grid1(nnz), kThreads, 0, s>>>(nnz, g->csr_to_csc.p, g->w_csc.p, g->w_csr.p);
  NWG_LAUNCH_CHECK();
  // raw weights all finite and >= 0: the binned weighted sweep's fixed-point bound holds
  DevBuf<uint32_t> bad(1);
  NWG_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
  if (nnz) {
    weight_check_kernel<<<grid1(nnz), kThreads, 0, s>>>(nnz, g->w_raw.p, bad.p);
    NWG_LAUNCH_CHECK();
  }
  uint32_t hb = 0;
  NWG_HIP(hipMemcpyAsync(&hb, bad.p, sizeof hb, hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  g->has_weights = true;
  g->w_normalized = normalize;
  g->w_nonneg = hb == 0;
  if (g->wst) {  // the weighted layouts carry the old weights
    g->wst->bl_csc.reset();
    g->wst->bl_csr.reset();
  }
}

// Runs the fixed point y <- fma(d, sum w y, base) from y0 over (rp, col, w) in the cached
// buffers of `st`; returns the sweeps executed and leaves the result in `out`.
// rhs: the constant term b whose L1 norm scales a relative threshold (null: the uniform
// teleport).  L (nullable): run the binned weighted sweep over this layout, its products
// scaled by 2^qshift.  The init and the max_iters sweeps run as one HIP graph captured into
// cap (re-captured when a buffer, the layout or a parameter changes), so the host launches
// one graph per solve instead of up to four kernels per sweep; base, y0 and rhs must be
// buffers that stay the same from call to call.
static uint64_t fixed_point(nwg_ctx* ctx, WState& st, std::vector<std::unique_ptr<Captured>>& caps,
                            const uint32_t* rp, const uint32_t* col,
                            const float* w, const DevBuf<uint32_t>& r0, const DevBuf<uint32_t>& r1, uint32_t n_items,
                            uint32_t n, const float* base, const float* y0, bool y0_uniform, float damping,
                            float threshold, uint64_t max_iters, float* out, nwg_pr_info* info, const float* rhs,
                            const BinnedLayout* L = nullptr, int qshift = 0) {
  hipStream_t s = ctx->stream;
  st.y0.ensure(n);
  st.y1.ensure(n);
  st.done.ensure(std::max<uint64_t>(max_iters, 1));
  st.err.ensure(std::max<uint64_t>(max_iters, 1));
  float* y[2] = {st.y0.p, st.y1.p};
  const uint32_t grid = persistent_grid(n_items);
  st.part.ensure(std::max<uint32_t>(grid, kNormBlocks));
  const bool relative = info && (info->flags & NWG_INFO_RELATIVE_THRESHOLD);
  if (relative) {
    st.scale.ensure(1);
    st.scale_done.ensure(1);
    NWG_HIP(hipMemsetAsync(st.scale_done.p, 0, sizeof(uint32_t), s));
    l1_scale_kernel<<<kNormBlocks, kThreads, 0, s>>>(n, rhs, 1.0 / (1.0 - (double)damping), st.part.p,
                                                    st.scale_done.p, st.scale.p);
    NWG_LAUNCH_CHECK();
  }
  if (L) {
    st.qs.ensure(1);
    st.qs_host = qshift;
    NWG_HIP(hipMemcpyAsync(st.qs.p, &st.qs_host, sizeof(int), hipMemcpyHostToDevice, s));
    st.bvals.ensure(L->pairs);
    st.bpart.ensure(std::max<uint64_t>((uint64_t)L->n_slots << L->bin_bits, 1));
    st.wpart.ensure(std::max<uint64_t>((uint64_t)L->n_items2 + L->n_items3, 1));
  }
  WParams P{};
  P.rp = rp;
  P.col = col;
  P.w = w;
  P.item_r0 = r0.p;
  P.item_r1 = r1.p;
  P.n_items = n_items;
  P.base = base;
  P.damp = damping;
  P.part = st.part.p;
  P.done = st.done.p;
  P.err = st.err.p;
  P.threshold = (double)threshold;
  P.thr_scale = relative ? st.scale.p : nullptr;
  // sweeps [k0, k1) (and the init with k0 = 0) on stream q
  auto enqueue = [&](hipStream_t q, uint64_t k0, uint64_t k1) {
    if (k0 == 0) {
      winit_kernel<<<grid1(std::max<uint64_t>(n, max_iters)), kThreads, 0, q>>>(
          n, y0_uniform ? nullptr : y0, (float)(1.0 / n), y[0], st.done.p, st.err.p, max_iters);
      NWG_LAUNCH_CHECK();
    }
    if (L) {
      // binned sweep: x double-buffered as y[k & 1] -> y[(k + 1) & 1]; its L1 change goes to st.err
      BinParams B{};
      B.brow = L->brow.p;
      B.bsrc = L->bsrc.p;
      B.bpid = L->bpid.p;
      B.vals = st.bvals.p;
      B.items2 = L->items2.p;
      B.items1 = L->items1.p;
      B.items3 = L->items3.p;
      B.seg_start = L->seg_start.p;
      B.seg_flat = L->seg_flat.p;
      B.part = st.bpart.p;
      B.item_he = L->item_he.p;
      B.item_hd = L->item_hd.p;
      B.hsrc = L->hsrc.p;
      B.n_hot = L->n_hot;
      B.basev = base;
      B.damp = damping;
      B.n = n;
      B.n_items1 = L->n_items1;
      B.n_items2 = L->n_items2;
      B.bin_bits = L->bin_bits;
      B.n_items3 = L->n_items3;
      B.threshold = (double)threshold;
      B.thr_scale = P.thr_scale;
      B.bw = L->bw.p;
      B.qshift = st.qs.p;
      B.wpart = st.wpart.p;
      B.werr = st.err.p;
      for (uint64_t k = k0; k < k1; ++k) {
        B.iter = (uint32_t)k;
        B.cin = y[k & 1];
        B.x = y[(k + 1) & 1];
        binned_enqueue_weighted_sweep(q, B);
      }
    } else {
      for (uint64_t k = k0; k < k1; ++k) {
        P.iter = (uint32_t)k;
        P.xin = y[k & 1];
        P.xout = y[(k + 1) & 1];
        wsweep_kernel<<<grid, kThreads, 0, q>>>(P);
        NWG_LAUNCH_CHECK();
      }
    }
  };
  CaptureKey key;
  const void* ptrs[CaptureKey::kPtrs] = {st.y0.p, st.y1.p, base, y0, st.err.p, st.done.p, st.part.p, rp, col, w,
                                         L, L ? st.bvals.p : nullptr, L ? st.bpart.p : nullptr,
                                         L ? st.wpart.p : nullptr, L ? st.qs.p : nullptr,
                                         relative ? st.scale.p : nullptr};
  for (int q = 0; q < CaptureKey::kPtrs; ++q) key.ptrs[q] = ptrs[q];
  key.iters = max_iters;
  key.damping = damping;
  key.threshold = threshold;
  key.plan_version = n_items;
  key.extra = (relative ? 1u : 0u) | (y0_uniform ? 2u : 0u);
  // the segments' graphs, captured when missing or stale
  uint64_t nseg = 0;
  for (uint64_t k0 = 0; k0 < max_iters || (k0 == 0 && nseg == 0); k0 = segment_end(k0, max_iters)) {
    const uint64_t k1 = segment_end(k0, max_iters);
    if (caps.size() <= nseg) caps.emplace_back(new Captured());
    Captured& cap = *caps[nseg];
    ++nseg;
    if (!cap.valid || !(cap.key == key)) {
      cap.reset();
      hipStream_t cs = ctx->capture_stream;
      NWG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(cs, k0, k1);
      } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(cs, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
      }
      NWG_HIP(hipStreamEndCapture(cs, &cap.graph));
      NWG_HIP(hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0));
      cap.key = key;
      cap.valid = true;
    }
    if (k1 >= max_iters) break;
  }
  double scale_now = 1.0;
  if (relative) {
    NWG_HIP(hipMemcpyAsync(&scale_now, st.scale.p, sizeof(double), hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
  }
  const double stop_now = relative ? (double)threshold * scale_now : (double)threshold;
  NWG_HIP(hipEventRecord(ctx->ev0, s));
  uint64_t seg = 0;
  for (uint64_t k0 = 0;; k0 = segment_end(k0, max_iters), ++seg) {
    const uint64_t k1 = segment_end(k0, max_iters);
    NWG_HIP(hipGraphLaunch(caps[seg]->exec, s));
    if (k1 >= max_iters) break;
    double last_err = 0.0;  // the segment's last sweep: stop when it converged
    NWG_HIP(hipMemcpyAsync(&last_err, st.err.p + (k1 - 1), sizeof(double), hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    if (last_err < stop_now) break;
  }
  NWG_HIP(hipEventRecord(ctx->ev1, s));
  std::vector<double> h(std::max<uint64_t>(max_iters, 1));
  NWG_HIP(hipMemcpyAsync(h.data(), st.err.p, max_iters * sizeof(double), hipMemcpyDeviceToHost, s));
  double scale = 1.0;
  if (relative) NWG_HIP(hipMemcpyAsync(&scale, st.scale.p, sizeof(double), hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  const double stop = relative ? (double)threshold * scale : (double)threshold;
  uint64_t iters = max_iters;
  double last = 0.0;
  for (uint64_t k = 0; k < max_iters; ++k) {
    last = h[k];
    if (last < stop) {
      iters = k + 1;
      break;
    }
  }
  NWG_HIP(hipMemcpyAsync(out, y[iters & 1], (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  NWG_HIP(hipStreamSynchronize(s));
  if (info) {
    float ms = 0.f;
    NWG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    info->iterations = iters;
    info->final_error = max_iters ? last : 0.0;
    info->ms_total = ms;
    info->ms_per_iter = iters ? ms / (double)iters : 0.0;
    info->sweep_ms_avg = 0.0;
    info->threshold_used = stop;
    if (info->error_trace)
      for (uint64_t k = 0; k < iters && k < info->error_trace_len; ++k) info->error_trace[k] = h[k];
  }
  return iters;
}

static WState& wstate(nwg_graph* g) {
  if (!g->wst) g->wst.reset(new WState());
  return *g->wst;
}

// The binned weighted sweep runs when the graph's plan (or, before any plan, the size-based
// default) asks for the binned sweep and the weights are normalised per source from
// non-negative raw weights: then every weight is in [0, 1] and each source's out-weights sum
// to 1, which the fixed-point bound below relies on.
static bool weighted_binned(const nwg_graph* g) {
  const nwg_plan_params pp = g->plan.built ? g->plan.params : default_plan_for(g);
  return pp.binned != 0 && g->w_normalized && g->w_nonneg && g->nnz > 0;
}

// Scale of the binned weighted sweep's fixed point: 2^t with 2^t * 2M in [1/2, 1), M a bound
// of every product and partial row sum, from b (the constant term) in exact arithmetic:
//   forward  x' = b + d A x, x0 = 1/n: |sum_u w(u->i) x_u| <= ||x||_1 <= max(1, ||b||_1 / (1-d))
//            (A is column-stochastic: ||A x||_1 <= ||x||_1);
//   adjoint  l' = b + d A^T l, l0 = b: |sum_i w(u->i) l_i| <= ||l||_inf <= ||b||_inf / (1-d)
//            (each row of A^T sums to 1).
// The factor 2 covers the FP32 rounding of the iterates.  False when b is not finite (the
// pull sweep then runs and shows the NaN / infinity as the reference would).
static bool fixed_point_scale(hipStream_t s, WState& st, uint32_t n, const float* b, bool forward, float damping,
                              int* t) {
  st.stats.ensure(2 * kNormBlocks);
  abs_stats_kernel<<<kNormBlocks, kThreads, 0, s>>>(n, b, st.stats.p);
  NWG_LAUNCH_CHECK();
  std::vector<double> h(2 * kNormBlocks);
  NWG_HIP(hipMemcpyAsync(h.data(), st.stats.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  double sum = 0.0, mx = 0.0;
  for (uint32_t q = 0; q < kNormBlocks; ++q) {
    sum += h[2 * q];
    mx = std::max(mx, h[2 * q + 1]);
  }
  if (!std::isfinite(sum)) return false;
  const double omd = 1.0 - (double)damping;
  const double M = forward ? std::max(1.0, sum / omd) : mx / omd;
  *t = 0;
  if (M > 0.0) std::frexp(2.0 * M, t), *t = -*t;
  return true;
}

static const BinnedLayout* weighted_layout(hipStream_t s, nwg_graph* g, WState& st, bool csr) {
  std::unique_ptr<BinnedLayout>& L = csr ? st.bl_csr : st.bl_csc;
  const int bits = (int)(g->plan.built ? g->plan.params : default_plan_for(g)).bin_bits;
  if (L && bits && L->bin_bits != bits) L.reset();  // the plan asks for other bins
  if (!L) {
    L.reset(new BinnedLayout());
    const ActiveCsc a = csr ? ActiveCsc{g->csr_ptr.p, g->csr_idx.p, nullptr, nullptr}
                            : ActiveCsc{g->csc_ptr.p, g->csc_idx.p, g->outdeg.p, nullptr};
    try {
      binned_build(s, a, g->n, g->nnz, *L, true, csr ? g->w_csr.p : g->w_csc.p, bits);
    } catch (...) {
      L.reset();
      throw;
    }
  }
  return L.get();
}

void wpr_forward(nwg_ctx* ctx, nwg_graph* g, const float* teleport_dev, float* x_dev, float damping, float threshold,
                 uint64_t max_iters, nwg_pr_info* info) {
  hipStream_t s = ctx->stream;
  NWG_REQUIRE(g->has_weights, "graph has no weights (nwg_graph_set_weights)");
  NWG_REQUIRE(max_iters < (1ull << 31), "max_iters too large");
  if (!g->w_items_csc_n && g->n)
    build_items(s, g->csc_ptr.p, g->n, g->w_items_csc_r0, g->w_items_csc_r1, &g->w_items_csc_n);
  WState& st = wstate(g);
  st.base.ensure(g->n);
  wbase_kernel<<<grid1(g->n), kThreads, 0, s>>>(g->n, teleport_dev, 1.0 - (double)damping, st.base.p);
  NWG_LAUNCH_CHECK();
  const BinnedLayout* L = nullptr;
  int t = 0;
  if (weighted_binned(g) && fixed_point_scale(s, st, g->n, st.base.p, true, damping, &t))
    L = weighted_layout(s, g, st, false);
  fixed_point(ctx, st, st.cap_fwd, g->csc_ptr.p, g->csc_idx.p, g->w_csc.p, g->w_items_csc_r0, g->w_items_csc_r1,
              g->w_items_csc_n, g->n, st.base.p, nullptr, true, damping, threshold, max_iters, x_dev, info,
              teleport_dev ? st.base.p : nullptr, L, t);
}

void wpr_backward(nwg_ctx* ctx, nwg_graph* g, const float* x_dev, const float* gx_dev, float* gw_dev, float* gv_dev,
                  float damping, float threshold, uint64_t max_iters, bool raw, nwg_pr_info* info) {
  hipStream_t s = ctx->stream;
  NWG_REQUIRE(g->has_weights, "graph has no weights (nwg_graph_set_weights)");
  NWG_REQUIRE(max_iters < (1ull << 31), "max_iters too large");
  NWG_REQUIRE(!raw || g->w_normalized, "raw-weight gradients need weights set with normalize=1");
  ensure_csr_items(s, g);
  const uint32_t n = g->n;
  const uint64_t nnz = g->nnz;
  WState& st = wstate(g);
  st.lambda.ensure(n);
  // adjoint fixed point: lambda = g + d A^T lambda, started from g
  // grad_x into a fixed buffer, so that the captured loop serves every call
  st.gin.ensure(n);
  NWG_HIP(hipMemcpyAsync(st.gin.p, gx_dev, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, s));
  const BinnedLayout* L = nullptr;
  int t = 0;
  if (weighted_binned(g) && fixed_point_scale(s, st, n, st.gin.p, false, damping, &t))
    L = weighted_layout(s, g, st, true);
  fixed_point(ctx, st, st.cap_adj, g->csr_ptr.p, g->csr_idx.p, g->w_csr.p, g->w_items_csr_r0, g->w_items_csr_r1,
              g->w_items_csr_n, n, st.gin.p, st.gin.p, false, damping, threshold, max_iters, st.lambda.p, info,
              st.gin.p, L, t);
  if (gw_dev && nnz) {
    if (!st.row_of.n) {
      st.row_of.alloc(nnz);
      row_of_kernel<<<grid1(nnz), kThreads, 0, s>>>(n, nnz, g->csc_ptr.p, st.row_of.p);
      NWG_LAUNCH_CHECK();
    }
    edge_grad_kernel<<<grid1(nnz), kThreads, 0, s>>>(nnz, st.row_of.p, g->csc_idx.p, damping, st.lambda.p, x_dev,
                                                     gw_dev);
    NWG_LAUNCH_CHECK();
    if (raw) {
      st.t.ensure(n);
      if (g->w_items_csr_n) {
        csr_rowsum_kernel<<<persistent_grid(g->w_items_csr_n), kThreads, 0, s>>>(
            g->csr_ptr.p, g->csr_to_csc.p, gw_dev, g->w_csc.p, g->w_items_csr_r0.p, g->w_items_csr_r1.p,
            g->w_items_csr_n, st.t.p);
        NWG_LAUNCH_CHECK();
      }
      chain_apply_kernel<<<grid1(nnz), kThreads, 0, s>>>(nnz, g->csc_idx.p, st.t.p, g->w_rowsum.p, gw_dev);
      NWG_LAUNCH_CHECK();
    }
  }
  if (gv_dev) {
    scale_kernel<<<grid1(n), kThreads, 0, s>>>(n, 1.0f - damping, st.lambda.p, gv_dev);
    NWG_LAUNCH_CHECK();
  }
  NWG_HIP(hipStreamSynchronize(s));
}

}  // namespace nwg
