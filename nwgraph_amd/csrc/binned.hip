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
//   bin pass    work item = a contiguous entry range of one bin (2^13 rows): add
//               vals[e] into LDS row accumulators (4 + 2 bytes in per entry); a bin that
//               is one item finishes its rows, the items of a split bin store partial
//               rows that the combine pass adds and finishes.  Finishing a row is the
//               reference's update (page_rank.hpp:133-135): x = fma(d, z, base),
//               |x - x_old| into the sweep's L1 change, contrib = x / deg.
//
// Row sums are exact fixed-point sums of the contributions (units of 2^-60, see quantize), so
// they do not depend on how the entries are split or ordered: results are deterministic,
// and z is the float rounding of the (nearly) exact sum of the float contributions.  The L1 change is
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
constexpr uint32_t kItem1Entries = 1u << 18;  // entries per chunk-pass item
constexpr uint32_t kPadSrc = 0xFFFFu;           // source offset of a pad entry (value 0)
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;     // bin-pass item of a bin that is one item
constexpr uint32_t kGroups1 = 4;              // groups of 8 entries in flight per lane: chunk pass
constexpr uint32_t kGroups2 = 2;              // and bin pass (more would cost it its second workgroup per CU)

// Row sums are exact 64-bit fixed-point sums in units of 2^-60.  A float contribution
// c = m * 2^(e-150) (m the 24-bit significand) is converted by shifting m: exact for
// c >= 2^-37, rounded to the nearest unit (ties away from zero) below that.  As a row sum
// z <= 1 stays below 2^60 units, integer adds never overflow and the sum does not depend on
// the order or the split of the entries.  (z <= sum of all x <= 1: x starts summing to 1
// and a sweep maps a vector of sum <= 1 to one of sum <= 1.)  No FP64 arithmetic is used.
__device__ __forceinline__ unsigned long long quantize(float c) {
  const uint32_t b = __float_as_uint(c);  // c >= 0 and finite
  const uint32_t e = b >> 23;
  const unsigned long long m = (b & 0x7FFFFFu) | (e ? 0x800000u : 0u);
  const int sh = (int)(e ? e : 1u) - 90;  // c * 2^60 = m * 2^sh
  if (sh >= 0) return m << sh;
  if (sh <= -25) return 0ull;  // below half a unit
  return (m + (1ull << (-sh - 1))) >> -sh;
}
// v * 2^-60 rounded to the nearest float (ties to even), with integer operations only
__device__ __forceinline__ float dequantize(unsigned long long v) {
  if (v == 0ull) return 0.f;
  const int lz = __clzll((long long)v);
  const unsigned long long nv = v << lz;  // bit 63 set
  uint32_t mant = (uint32_t)(nv >> 40);   // 24 significant bits
  const unsigned long long rest = nv & ((1ull << 40) - 1), half = 1ull << 39;
  if (rest > half || (rest == half && (mant & 1u))) ++mant;
  int ex = 63 - lz - 60 - 23;  // value = mant * 2^ex
  if (mant == (1u << 24)) {
    mant >>= 1;
    ++ex;
  }
  return __builtin_ldexpf((float)mant, ex);  // exact: mant < 2^24, result a normal float
}

// One graph: every workgroup of sweep k exits when sweep k-1's change was below the
// threshold.  Several graphs (row_graph set): the passes always run and each graph's rows
// are frozen once that graph has converged (graph_frozen).
// One graph's change of sweep k is the sum of its kBinAccLanes lane counters; every wave
// loads and sums them (exactly, in fixed point), so all waves of a workgroup agree.
__device__ __forceinline__ bool converged_before(const BinParams& P) {
  static_assert(kBinAccLanes == 64, "one lane counter per wavefront lane");
  if (P.row_graph || P.iter == 0) return false;
  unsigned long long v =
      P.acc[(uint64_t)(P.iter - 1) * kBinAccStride + (uint64_t)(threadIdx.x & 63u) * kBinAccPad];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return fixed_to_err(v) < P.threshold;
}
// the lane counter a workgroup adds one graph's change of this sweep into
__device__ __forceinline__ unsigned long long* residual_lane(const BinParams& P) {
  const uint32_t b = blockIdx.x + blockIdx.y * gridDim.x;
  return P.acc + (uint64_t)P.iter * kBinAccStride + (uint64_t)(b % kBinAccLanes) * kBinAccPad;
}
__device__ __forceinline__ bool graph_frozen(const BinParams& P, uint32_t g) {
  return P.iter > 0 && fixed_to_err(P.acc[(uint64_t)(P.iter - 1) * P.ngraphs + g]) < P.threshold;
}


// the reference's row update (page_rank.hpp:133-135) -> |x_new - x_old|
__device__ __forceinline__ double finish_row(const BinParams& P, uint32_t i, float z) {
  const float b = P.basev ? P.basev[i] : P.base;
  const float xo = P.x[i];
  const float xn = __builtin_fmaf(P.damp, z, b);
  P.x[i] = xn;
  P.cout[i] = xn / (float)P.deg[i];
  return (double)fabsf(xn - xo);
}

constexpr uint32_t kGraphSlots = 64;  // graphs whose L1 changes a workgroup first adds in LDS

// Finishes rows r0 + v for v = t, t + nt, ... < nrows (z(v) their sums) of a workgroup with
// nt threads, and accounts their L1 changes.  One graph: a block sum, one atomic into a
// lane counter of the sweep.  Several graphs: each row's change, in fixed point, goes to an LDS slot of its
// graph (graphs g0 .. g0 + kGraphSlots - 1, g0 the graph of row r0; the union's rows are
// grouped by graph) or straight to acc, then one atomic per graph; rows of converged graphs
// are left as they are.  red: nt / 64 doubles, gerr: kGraphSlots.  nrows <= MAXR; the row
// loop has a fixed trip count (MAXR / NT, unrolled) so a thread's row loads are all in
// flight together.
// MULTI: whether P.row_graph is set (the kernels are instantiated per case, so the one-graph
// kernels carry no per-graph code).
template <int NT, int MAXR, bool MULTI, class Z>
__device__ __forceinline__ void finish_rows(const BinParams& P, uint32_t r0, uint32_t nrows, Z z, double* red,
                                            unsigned long long* gerr) {
  static_assert(MAXR % NT == 0, "whole rounds of rows");
  if constexpr (!MULTI) {
    (void)gerr;
    double e = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < MAXR / NT; ++k) {
      const uint32_t v = threadIdx.x + k * NT;
      if (v < nrows) e += finish_row(P, r0 + v, z(v));
    }
    e = wave_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (uint32_t q = 0; q < NT / 64; ++q) t += red[q];
      if (t > 0.0) atomicAdd(residual_lane(P), err_to_fixed(t));
    }
  } else {
    for (uint32_t q = threadIdx.x; q < kGraphSlots; q += NT) gerr[q] = 0ull;
    __syncthreads();
    const uint32_t g0 = nrows ? P.row_graph[r0] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < MAXR / NT; ++k) {
      const uint32_t v = threadIdx.x + k * NT;
      if (v >= nrows) continue;
      const uint32_t g = P.row_graph[r0 + v];
      if (graph_frozen(P, g)) continue;
      const double e = finish_row(P, r0 + v, z(v));
      if (e > 0.0) {
        if (g - g0 < kGraphSlots) atomicAdd(&gerr[g - g0], err_to_fixed(e));
        else atomicAdd(P.acc + (uint64_t)P.iter * P.ngraphs + g, err_to_fixed(e));
      }
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < kGraphSlots; q += NT)
      if (gerr[q]) atomicAdd(P.acc + (uint64_t)P.iter * P.ngraphs + g0 + q, gerr[q]);
  }
}


// chunk pass: vals[e] = contrib[src(e)] over the tiles of one item.  Tiles are padded to
// multiples of 8 entries (pads have source kPadSrc and get 0), so a lane moves 8 entries
// with one 16-byte id load and two 16-byte value stores.
__global__ __launch_bounds__(kBT) void pr_chunk_kernel(BinParams P) {
  if (converged_before(P)) return;
  __shared__ float4 cst4[kChunkSrc / 4];
  float* cst = reinterpret_cast<float*>(cst4);
  __shared__ uint32_t sstart[kBinMaxSegs];
  __shared__ uint32_t sflat[kBinMaxSegs + 1];
  const uint4 it = P.items1[blockIdx.x];  // {chunk, first tile, end tile, groups of 8 entries}
  const uint32_t ns = it.z - it.y, m = it.w;
  for (uint32_t k = threadIdx.x; k < ns; k += kBT) {
    sstart[k] = P.seg_start[it.y + k];
    sflat[k] = P.seg_flat[it.y + k];
  }
  if (threadIdx.x == 0) sflat[ns] = m;
  // stage the chunk's contributions: every thread's 16-byte loads issued before any is
  // waited on (one load at a time would leave the workgroup waiting on memory latency)
  const uint32_t u0 = it.x << kChunkBits;
  if ((uint64_t)u0 + kChunkSrc <= P.n) {
    constexpr uint32_t kV = kChunkSrc / 4 / kBT;
    const float4* src = reinterpret_cast<const float4*>(P.cin + u0);
    float4 v[kV];
#pragma unroll
    for (uint32_t k = 0; k < kV; ++k) v[k] = src[threadIdx.x + k * kBT];
#pragma unroll
    for (uint32_t k = 0; k < kV; ++k) cst4[threadIdx.x + k * kBT] = v[k];
  } else {
    for (uint32_t q = threadIdx.x; q < kChunkSrc; q += kBT) cst[q] = (u0 + q < P.n) ? P.cin[u0 + q] : 0.f;
  }
  __syncthreads();
  for (uint32_t g0 = 0; g0 < m; g0 += kGroups1 * kBT) {
    uint32_t p[kGroups1];
    uint4 uu[kGroups1];
#pragma unroll
    for (uint32_t h = 0; h < kGroups1; ++h) {
      const uint32_t g = g0 + h * kBT + threadIdx.x;
      if (g < m) {
        uint32_t lo = 0, hi = ns - 1;  // last tile j with sflat[j] <= g
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (sflat[mid] <= g) lo = mid;
          else hi = mid - 1;
        }
        p[h] = sstart[lo] + 8u * (g - sflat[lo]);
        uu[h] = *reinterpret_cast<const uint4*>(P.bsrc + p[h]);
      }
    }
#pragma unroll
    for (uint32_t h = 0; h < kGroups1; ++h) {
      if (g0 + h * kBT + threadIdx.x >= m) continue;
      const uint32_t w[4] = {uu[h].x, uu[h].y, uu[h].z, uu[h].w};
      float o[8];
#pragma unroll
      for (uint32_t t = 0; t < 8; ++t) {
        const uint32_t u = (w[t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
        o[t] = u == kPadSrc ? 0.f : cst[u];
      }
      *reinterpret_cast<float4*>(P.vals + p[h]) = make_float4(o[0], o[1], o[2], o[3]);
      *reinterpret_cast<float4*>(P.vals + p[h] + 4) = make_float4(o[4], o[5], o[6], o[7]);
    }
  }
}

constexpr uint32_t kNoRow = 0xFFFFFFFFu;

// Adds each lane's run (row cur, fixed-point sum run; cur == kNoRow: none) into lacc.  When
// every run in the wave belongs to one row (a hub row filling the wave), the runs are summed
// across the wave first and one lane issues the atomic, instead of 64 atomics on one address.
// Integer adds: the order does not matter.  All lanes of the wave must call it together.
__device__ __forceinline__ void wave_flush(unsigned long long* lacc, uint32_t cur, unsigned long long run,
                                           uint32_t lane) {
  const bool has = cur != kNoRow;
  const unsigned long long m = __ballot(has);
  if (!m) return;
  const int first = __builtin_ctzll(m);
  const uint32_t row0 = (uint32_t)__shfl((int)cur, first);
  if (__ballot(has && cur != row0) == 0ull) {
    unsigned long long t = has ? run : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if ((int)lane == first) atomicAdd(&lacc[row0], t);
  } else if (has) {
    atomicAdd(&lacc[cur], run);
  }
}

// Adds the entries [e0, e1) of one bin into lacc.  A wave takes 512 consecutive entries (8
// per lane, from an 8-aligned start: the arrays are padded), kGroups2 such blocks in flight;
// the loop bounds are uniform over the wave so every lane takes part in the votes of
// wave_flush.  Runs of one row are added in registers; a lane's last run goes through
// wave_flush.  HOT: the entries lie in a hot tile and their values are hot[source] (pads,
// source kPadSrc, add 0); else they are read from P.vals.
template <bool HOT>
__device__ __forceinline__ void bin_range(const BinParams& P, uint64_t e0, uint64_t e1, unsigned long long* lacc,
                                          const float* hot) {
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint64_t wq = (e0 & ~7ull) + 512u * w; wq < e1; wq += kGroups2 * 512u * (kBT / 64)) {
    uint4 rw[kGroups2], sw[kGroups2];
    float4 va[kGroups2], vb[kGroups2];
#pragma unroll
    for (uint32_t h = 0; h < kGroups2; ++h) {
      // unconditional loads (clamped into the range's first group when past its end, then
      // masked below) so that all groups' loads are issued before any is waited on
      const uint64_t q = wq + h * 512u * (kBT / 64) + 8u * lane;
      const uint64_t qc = q < e1 ? q : (e0 & ~7ull);
      rw[h] = *reinterpret_cast<const uint4*>(P.brow + qc);
      if constexpr (HOT) {
        sw[h] = *reinterpret_cast<const uint4*>(P.bsrc + qc);
      } else {
        va[h] = *reinterpret_cast<const float4*>(P.vals + qc);
        vb[h] = *reinterpret_cast<const float4*>(P.vals + qc + 4);
      }
    }
#pragma unroll
    for (uint32_t h = 0; h < kGroups2; ++h) {
      const uint64_t q = wq + h * 512u * (kBT / 64) + 8u * lane;
      uint32_t cur = kNoRow;
      unsigned long long run = 0ull;
      if (q < e1) {
        const uint32_t rows[8] = {rw[h].x & 0xFFFFu, rw[h].x >> 16, rw[h].y & 0xFFFFu, rw[h].y >> 16,
                                  rw[h].z & 0xFFFFu, rw[h].z >> 16, rw[h].w & 0xFFFFu, rw[h].w >> 16};
        float vs[8];
        if constexpr (HOT) {
          const uint32_t sv[4] = {sw[h].x, sw[h].y, sw[h].z, sw[h].w};
#pragma unroll
          for (uint32_t t = 0; t < 8; ++t) {
            const uint32_t u = (sv[t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
            vs[t] = u == kPadSrc ? 0.f : hot[u];
          }
        } else {
          vs[0] = va[h].x; vs[1] = va[h].y; vs[2] = va[h].z; vs[3] = va[h].w;
          vs[4] = vb[h].x; vs[5] = vb[h].y; vs[6] = vb[h].z; vs[7] = vb[h].w;
        }
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r) {
          const uint64_t e = q + r;
          if (e < e0 || e >= e1) continue;
          const unsigned long long f = quantize(vs[r]);
          if (rows[r] == cur) {
            run += f;
          } else {
            if (cur != kNoRow) atomicAdd(&lacc[cur], run);
            cur = rows[r];
            run = f;
          }
        }
      }
      wave_flush(lacc, cur, run, lane);
    }
  }
}

// bin pass: row sums of one entry range of one bin in LDS; a bin that is one item finishes
// its rows here, an item of a split bin stores its partial rows in its slot of P.part.
// The range's part inside the bin's hot tile (a prefix of the bin) reads its values from
// the hot contributions staged in LDS.
template <bool MULTI>
__global__ __launch_bounds__(kBT) void pr_bin_kernel(BinParams P) {
  if (converged_before(P)) return;
  __shared__ unsigned long long lacc[kBinRows];
  __shared__ double red[kBT / 64];
  __shared__ unsigned long long gerr[MULTI ? kGraphSlots : 1];
  __shared__ float hot[MULTI ? 1 : kHotSrc];
  const uint4 it = P.items2[blockIdx.x];  // {bin, first entry, end entry, partial slot or kNoSlot}
  const uint32_t b = it.x;
  for (uint32_t q = threadIdx.x; q < kBinRows; q += kBT) lacc[q] = 0ull;
  uint32_t he = it.y;  // end of the hot part of [it.y, it.z)
  if constexpr (!MULTI) {
    if (P.n_hot) {
      he = min(max(P.hot_end[b], it.y), it.z);
      if (he > it.y) {
        constexpr uint32_t kR = (kHotSrc + kBT - 1) / kBT;
        float v[kR];
#pragma unroll
        for (uint32_t k = 0; k < kR; ++k) {
          const uint32_t q = threadIdx.x + k * kBT;
          v[k] = q < P.n_hot ? P.cin[q] : 0.f;
        }
#pragma unroll
        for (uint32_t k = 0; k < kR; ++k) {
          const uint32_t q = threadIdx.x + k * kBT;
          if (q < kHotSrc) hot[q] = v[k];
        }
      }
    }
  }
  __syncthreads();
  if constexpr (!MULTI) {
    if (he > it.y) bin_range<true>(P, it.y, he, lacc, hot);
  }
  bin_range<false>(P, he, it.z, lacc, hot);
  __syncthreads();
  const uint32_t r0 = b << kBinBits;
  if (it.w != kNoSlot) {
    unsigned long long* dst = P.part + (uint64_t)it.w * kBinRows;
    for (uint32_t v = threadIdx.x; v < kBinRows; v += kBT) dst[v] = lacc[v];
    return;
  }
  const uint32_t nrows = r0 < P.n ? min(kBinRows, P.n - r0) : 0u;
  finish_rows<kBT, kBinRows, MULTI>(P, r0, nrows, [&](uint32_t v) { return dequantize(lacc[v]); }, red, gerr);
}

// rows of the bins split over several bin-pass items: add the items' partial rows (exact,
// so in any order) and finish the rows.  A workgroup (one combine unit) takes 256 / L rows
// of one split bin, each row's slots shared by L lanes (slot k on lane k mod L); the host
// picks L per bin from its slot count, so a bin with hundreds of slots is read by many
// workgroups at once and a bin with a few by few.  One workgroup per unit.
template <bool MULTI>
__global__ __launch_bounds__(256) void pr_bin_combine_kernel(BinParams P) {
  if (converged_before(P)) return;
  __shared__ double red[4];
  __shared__ unsigned long long gerr[MULTI ? kGraphSlots : 1];
  __shared__ unsigned long long psum[256];
  const uint4 it = P.items3[blockIdx.x];  // {bin, first slot, slots, first row | log2 L << 16}
  const uint32_t lanes = 1u << (it.w >> 16), rows = 256u >> (it.w >> 16);
  const uint32_t v0 = it.w & 0xFFFFu;
  const uint32_t r0 = (it.x << kBinBits) + v0;
  const uint32_t nrows = r0 < P.n ? min(rows, P.n - r0) : 0u;
  {
    const uint32_t row = threadIdx.x % rows, lane = threadIdx.x / rows;
    const unsigned long long* src = P.part + (uint64_t)it.y * kBinRows + v0 + row;
    unsigned long long s0 = 0ull, s1 = 0ull, s2 = 0ull, s3 = 0ull;
    uint32_t k = lane;
    for (; k + 3 * lanes < it.z; k += 4 * lanes) {
      s0 += src[(uint64_t)k * kBinRows];
      s1 += src[(uint64_t)(k + lanes) * kBinRows];
      s2 += src[(uint64_t)(k + 2 * lanes) * kBinRows];
      s3 += src[(uint64_t)(k + 3 * lanes) * kBinRows];
    }
    for (; k < it.z; k += lanes) s0 += src[(uint64_t)k * kBinRows];
    psum[threadIdx.x] = (s0 + s1) + (s2 + s3);
  }
  __syncthreads();
  finish_rows<256, 256, MULTI>(
      P, r0, nrows,
      [&](uint32_t v) {
        unsigned long long t = 0ull;
        for (uint32_t q = 0; q < lanes; ++q) t += psum[q * rows + v];
        return dequantize(t);
      },
      red, gerr);
}

// ---- layout construction ----
// tile key of CSC position p (row v, source u) and its packed low bits
// (tile column: 0 for a source below n_hot, else 1 + chunk; without hot tiles the chunk)
__global__ __launch_bounds__(256) void bin_key_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ rp,
                                                      const uint32_t* __restrict__ col, uint32_t n_cols,
                                                      uint32_t n_hot, uint64_t* __restrict__ keys,
                                                      uint32_t* __restrict__ low) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= nnz) return;
  uint32_t lo = 0, hi = n;  // row of p: first r with rp[r + 1] > p
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rp[mid + 1] <= p) lo = mid + 1;
    else hi = mid;
  }
  const uint32_t u = col[p];
  const bool hot = u < n_hot;
  const uint32_t c = hot ? 0u : (u >> kChunkBits) + (n_hot ? 1u : 0u);
  keys[p] = (uint64_t)(lo >> kBinBits) * n_cols + c;
  low[p] = ((lo & (kBinRows - 1)) << 16) | (hot ? u : (u & (kChunkSrc - 1)));
}

// sorted entry q of tile t goes to padded position ptp[t] + (q - tp[t])
__global__ __launch_bounds__(256) void bin_place_kernel(uint64_t nnz, const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ low,
                                                        const uint32_t* __restrict__ tp,
                                                        const uint32_t* __restrict__ ptp,
                                                        uint16_t* __restrict__ brow, uint16_t* __restrict__ bsrc) {
  const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= nnz) return;
  const uint64_t t = keys[q];
  const uint64_t np = ptp[t] + (q - tp[t]);
  brow[np] = (uint16_t)(low[q] >> 16);
  bsrc[np] = (uint16_t)(low[q] & 0xFFFFu);
}

// pads of tile t: row of the tile's last entry (so they join its run), source kPadSrc
__global__ __launch_bounds__(256) void bin_pad_kernel(uint64_t n_tiles, const uint32_t* __restrict__ tp,
                                                      const uint32_t* __restrict__ ptp, uint16_t* __restrict__ brow,
                                                      uint16_t* __restrict__ bsrc) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n_tiles) return;
  const uint32_t cnt = tp[t + 1] - tp[t];
  if (!cnt) return;
  const uint32_t last = ptp[t] + cnt - 1;
  for (uint32_t r = cnt; r < ptp[t + 1] - ptp[t]; ++r) {
    brow[ptp[t] + r] = brow[last];
    bsrc[ptp[t] + r] = (uint16_t)kPadSrc;
  }
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

__global__ __launch_bounds__(256) void residual_clear_kernel(unsigned long long* __restrict__ lanes, uint64_t words) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < words) lanes[t] = 0ull;
}

// acc[k] = sum of sweep k's lane counters (exact integer sum); one thread per sweep
__global__ __launch_bounds__(256) void residual_fold_kernel(const unsigned long long* __restrict__ lanes,
                                                            unsigned long long* __restrict__ acc, uint64_t iters) {
  const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= iters) return;
  unsigned long long t = 0ull;
  for (uint32_t l = 0; l < kBinAccLanes; ++l) t += lanes[k * kBinAccStride + (uint64_t)l * kBinAccPad];
  acc[k] = t;
}

}  // namespace

void binned_build(hipStream_t s, const ActiveCsc& a, uint32_t n, uint64_t nnz, BinnedLayout& L, bool hot) {
  NWG_REQUIRE(nnz < (1ull << 32) - 16, "binned layout needs fewer than 2^32 - 16 entries");
  L = BinnedLayout();
  L.n_bins = (uint32_t)ceil_div(n, kBinRows);
  L.n_chunks = (uint32_t)ceil_div(n, kChunkSrc);
  L.n_hot = hot ? std::min<uint32_t>(kHotSrc, n) : 0u;
  L.nnz = nnz;
  // tile columns per bin: the hot tile (if any), then one per chunk
  const uint32_t hc = L.n_hot ? 1u : 0u, TC = L.n_chunks + hc;
  const uint64_t n_tiles = (uint64_t)L.n_bins * TC;
  std::vector<uint32_t> tp(n_tiles + 1, 0);  // becomes the padded tile pointers
  uint64_t total = 0;
  if (nnz) {
    DevBuf<uint64_t> keys(nnz), keys_tmp(nnz);
    DevBuf<uint32_t> low(nnz), low_tmp(nnz);
    bin_key_kernel<<<grid256(nnz), 256, 0, s>>>(n, nnz, a.ptr, a.idx, TC, L.n_hot, keys.p, low.p);
    NWG_LAUNCH_CHECK();
    int kb = 1;
    while (kb < 64 && (1ull << kb) < n_tiles) ++kb;
    // stable: entries of one tile keep their CSC order (row, then source)
    radix_sort_pairs(s, keys.p, keys_tmp.p, low.p, low_tmp.p, nnz, kb);
    DevBuf<uint32_t> d_tp(n_tiles + 1), d_ptp(n_tiles + 1);
    tile_ptr_kernel<<<grid256(n_tiles + 1), 256, 0, s>>>(nnz, keys.p, n_tiles, d_tp.p);
    NWG_LAUNCH_CHECK();
    NWG_HIP(hipMemcpyAsync(tp.data(), d_tp.p, (n_tiles + 1) * 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> ptp(n_tiles + 1, 0);
    for (uint64_t t = 0; t < n_tiles; ++t) {
      total += (uint64_t)(tp[t + 1] - tp[t] + 7) / 8 * 8;
      NWG_REQUIRE(total < (1ull << 32) - 16, "binned layout needs fewer than 2^32 - 16 padded entries");
      ptp[t + 1] = (uint32_t)total;
    }
    L.padded = total + 8;
    L.brow.alloc(L.padded);
    L.bsrc.alloc(L.padded);
    NWG_HIP(hipMemsetAsync(L.brow.p, 0, L.padded * 2, s));
    NWG_HIP(hipMemsetAsync(L.bsrc.p, 0xFF, L.padded * 2, s));
    NWG_HIP(hipMemcpyAsync(d_ptp.p, ptp.data(), (n_tiles + 1) * 4, hipMemcpyHostToDevice, s));
    bin_place_kernel<<<grid256(nnz), 256, 0, s>>>(nnz, keys.p, low.p, d_tp.p, d_ptp.p, L.brow.p, L.bsrc.p);
    NWG_LAUNCH_CHECK();
    bin_pad_kernel<<<grid256(n_tiles), 256, 0, s>>>(n_tiles, d_tp.p, d_ptp.p, L.brow.p, L.bsrc.p);
    NWG_LAUNCH_CHECK();
    NWG_HIP(hipStreamSynchronize(s));
    tp.swap(ptp);
  } else {
    L.padded = 8;
    L.brow.alloc(L.padded);
    L.bsrc.alloc(L.padded);
    NWG_HIP(hipMemsetAsync(L.brow.p, 0, L.padded * 2, s));
    NWG_HIP(hipMemsetAsync(L.bsrc.p, 0xFF, L.padded * 2, s));
  }
  const uint32_t C = L.n_chunks;
  // hot tile ends (the hot tile is the first of its bin)
  std::vector<uint32_t> he(L.n_bins);
  for (uint32_t b = 0; b < L.n_bins; ++b) he[b] = tp[(uint64_t)b * TC + hc];
  L.hot_end.alloc(std::max<size_t>(he.size(), 1));
  if (!he.empty()) NWG_HIP(hipMemcpyAsync(L.hot_end.p, he.data(), he.size() * 4, hipMemcpyHostToDevice, s));
  // bin pass items: every bin, cut into ranges of at most kItem2Entries entries; the items
  // of a split bin get consecutive partial slots, and the bin a combine item
  std::vector<uint4> it2, it3;
  uint32_t slots = 0;
  for (uint32_t b = 0; b < L.n_bins; ++b) {
    const uint32_t p0 = tp[(uint64_t)b * TC], p1 = tp[(uint64_t)(b + 1) * TC];
    const uint32_t cnt = (uint32_t)std::max<uint64_t>(ceil_div(p1 - p0, kItem2Entries), 1);
    if (cnt > 1) {
      // combine units: 256 / L rows each, L lanes per row sharing its cnt slots
      const uint32_t lg = cnt > 32 ? 3u : cnt > 8 ? 2u : 0u;
      for (uint32_t v0 = 0; v0 < kBinRows; v0 += 256u >> lg) it3.push_back(make_uint4(b, slots, cnt, v0 | (lg << 16)));
    }
    for (uint32_t q = 0; q < cnt; ++q) {
      const uint32_t a0 = p0 + (uint32_t)((uint64_t)(p1 - p0) * q / cnt);
      const uint32_t a1 = p0 + (uint32_t)((uint64_t)(p1 - p0) * (q + 1) / cnt);
      it2.push_back(make_uint4(b, a0, a1, cnt > 1 ? slots++ : kNoSlot));
    }
  }
  // chunk pass items: the non-empty tiles of each chunk (long tiles cut into pieces), cut
  // at kItem1Entries entries or kBinMaxSegs pieces; offsets and sizes in groups of 8
  // entries (every padded tile, and kItem1Entries, is a whole number of groups)
  std::vector<uint4> it1;
  std::vector<uint32_t> sst, sfl;
  for (uint32_t c = 0; c < C; ++c) {
    uint32_t first = (uint32_t)sst.size(), flat = 0;
    for (uint32_t b = 0; b < L.n_bins; ++b) {
      const uint32_t t0 = tp[(uint64_t)b * TC + hc + c], t1 = tp[(uint64_t)b * TC + hc + c + 1];
      // a long tile is cut into pieces of at most kItem1Entries entries
      for (uint32_t q0 = t0; q0 < t1; q0 += kItem1Entries) {
        const uint32_t len = (std::min<uint32_t>(t1 - q0, kItem1Entries)) / 8;
        if ((uint32_t)sst.size() > first &&
            (flat + len > kItem1Entries / 8 || (uint32_t)sst.size() - first >= kBinMaxSegs)) {
          it1.push_back(make_uint4(c, first, (uint32_t)sst.size(), flat));
          first = (uint32_t)sst.size();
          flat = 0;
        }
        sst.push_back(q0);
        sfl.push_back(flat);
        flat += len;
      }
    }
    if ((uint32_t)sst.size() > first) it1.push_back(make_uint4(c, first, (uint32_t)sst.size(), flat));
  }
  // largest items first: with one workgroup per CU the pass ends when its last item does,
  // so the small items fill in at the end
  std::stable_sort(it1.begin(), it1.end(), [](const uint4& x, const uint4& y) { return x.w > y.w; });
  L.n_items2 = (uint32_t)it2.size();
  L.n_items3 = (uint32_t)it3.size();
  L.n_slots = slots;
  L.n_items1 = (uint32_t)it1.size();
  L.n_segs = (uint32_t)sst.size();
  L.items2.alloc(std::max<size_t>(it2.size(), 1));
  L.items3.alloc(std::max<size_t>(it3.size(), 1));
  if (!it3.empty()) NWG_HIP(hipMemcpyAsync(L.items3.p, it3.data(), it3.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
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
  if (P.row_graph) {
    pr_bin_kernel<true><<<P.n_items2, kBT, 0, s>>>(P);
    if (P.n_items3) pr_bin_combine_kernel<true><<<P.n_items3, 256, 0, s>>>(P);
  } else {
    pr_bin_kernel<false><<<P.n_items2, kBT, 0, s>>>(P);
    if (P.n_items3) pr_bin_combine_kernel<false><<<P.n_items3, 256, 0, s>>>(P);
  }
  NWG_LAUNCH_CHECK();
}

void binned_clear_residuals(hipStream_t s, unsigned long long* lanes, uint64_t iters) {
  const uint64_t words = iters * kBinAccStride;
  if (!words) return;
  residual_clear_kernel<<<(uint32_t)((words + 255) / 256), 256, 0, s>>>(lanes, words);
  NWG_LAUNCH_CHECK();
}

void binned_fold_residuals(hipStream_t s, const unsigned long long* lanes, unsigned long long* acc, uint64_t iters) {
  if (!iters) return;
  residual_fold_kernel<<<(uint32_t)((iters + 255) / 256), 256, 0, s>>>(lanes, acc, iters);
  NWG_LAUNCH_CHECK();
}

}  // namespace nwg
