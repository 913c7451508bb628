// binned.hip -- PageRank sweep by propagation blocking (binned schedule of the pull sweep).
//
// The pull sweep of nw::graph::page_rank (include/nwgraph/algorithms/page_rank.hpp:127-136)
// gathers contrib[u] for every in-edge; at 2^25 vertices every gather is its own L2
// request and most miss.  The binned sweep turns the gathers into two streaming passes
// over a static, tiled copy of the CSC (layout in nwg_internal.hpp, BinnedLayout):
//
//   chunk pass  work item = one source chunk (2^15 sources) and a run of its tiles:
//               stage contrib[chunk] in LDS, then vals[p] = contrib[src(p)] for every
//               pair p of those tiles (2-byte source offsets in, 4-byte values out)
//   bin pass    work item = a contiguous entry range of one bin (2^13 rows): add each
//               entry's pair value into LDS row accumulators (2 bytes per entry and 4 per
//               pair in); a bin that is one item finishes its rows, the items of a split
//               bin store partial rows that the combine pass adds and finishes.  Finishing
//               a row is the reference's update (page_rank.hpp:133-135): x = fma(d, z, base),
//               |x - x_old| into the sweep's L1 change, contrib = x / deg.
//
// A pair is (bin, source): the entries of one bin with one source.  At RMAT-25 with degree
// ordering there are 0.43 pairs per entry outside the hot tiles, so moving values per pair
// instead of per entry takes the streamed bytes from 12 to about 6 per entry.
//
// Row sums are exact fixed-point sums of the contributions (units of 2^-60, see quantize), so
// they do not depend on how the entries are split or ordered: results are deterministic,
// and z is the float rounding of the (nearly) exact sum of the float contributions.  The L1 change is
// accumulated in fixed point as in the pull sweep (device_utils.hpp), and every
// workgroup first checks the previous sweep's change against the threshold
// (page_rank.hpp:144-146), so a captured loop stops working after convergence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

constexpr int kBT = 1024;  // threads of a bin / chunk pass workgroup (16 waves, one workgroup per CU)
// Bins of 2^BB rows, BB = 13 or 14 (the layout's choice, BinnedLayout::bin_bits): the bin-pass
// kernels are instantiated for both.
template <int BB>
constexpr uint32_t bin_rows() { return 1u << BB; }
constexpr uint32_t kChunkSrc = 1u << kChunkBits;
constexpr uint32_t kItem2Log = 21;  // log2 of the most entries of a bin-pass item
constexpr uint32_t kItem1Pairs = 1u << 18;    // pairs per chunk-pass item
constexpr uint32_t kPadSrc = 0xFFFFu;           // source offset of a pad pair (value 0)
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;     // bin-pass item of a bin that is one item
constexpr uint32_t kGroups1 = 4;              // groups of 8 pairs in flight per lane in the chunk pass (as 8 halves)

// Inside a kBinBlock block the entries are stored slot-major: entry r of the block (in
// layout order) sits at lane r % 64, slot r / 64 of the block's 64 x 8 array, so that a lane's
// 16-byte load brings its 8 slots and slot t of the 64 lanes holds 64 consecutive entries.
// Their pairs are consecutive too, so each slot's pair loads touch one or two cache lines
// (lane-major, every slot's loads spanned the block's whole pair range, ~7 lines each).
__host__ __device__ __forceinline__ uint32_t bin_slot(uint32_t e) {
  return (e & ~(kBinBlock - 1u)) | ((e & 63u) << 3) | ((e >> 6) & 7u);
}

// Cache hints (the plan's stream_hints): NT >= 1 loads the streams each lane reads once
// (row offsets, source offsets, partial rows, staged contributions, scores, out-degrees)
// non-temporally; NT >= 2 also stores the pair values, partial rows, scores and
// contributions non-temporally.  Pair values and hot-tile source ids are always cached
// loads: the eight entry slots of a lane re-read the same lines.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <int NT, class T>
__device__ __forceinline__ T bld(const T* p) {
  if constexpr (NT >= 1) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int NT, class T>
__device__ __forceinline__ void bst(T* p, T v) {
  if constexpr (NT >= 2) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Row sums are exact 64-bit fixed-point sums in units of 2^-60.  A float contribution
// c in [0, 1] is converted with float operations that are all exact: s = c * 2^28,
// h = floor(s) (h <= 2^28; (float)h is exact), r = s - h (the fraction bits of s), and
// l = floor(r * 2^32); the fixed-point value h * 2^32 + l is c * 2^60 truncated to an
// integer, which is exact for c >= 2^-37 (every significand bit at or above 2^-60).  As a
// row sum z <= 1 stays below 2^60 units, integer adds never overflow and the sum does not
// depend on the order or the split of the entries.  (z <= sum of all x <= 1: x starts
// summing to 1 and a sweep maps a vector of sum <= 1 to one of sum <= 1.)  No FP64
// arithmetic is used.
__device__ __forceinline__ unsigned long long quantize(float c) {
  const float sc = c * 268435456.f;  // c * 2^28
  const uint32_t h = (uint32_t)sc;
  const float r = __builtin_amdgcn_fractf(sc);  // sc - floor(sc) = sc - h, exact
  const uint32_t l = (uint32_t)(r * 4294967296.f);  // floor(r * 2^32) < 2^32
  return ((unsigned long long)h << 32) | l;
}
// v * 2^-60 rounded to the nearest float (ties to even).  Below 2^53 (sums under 2^-7, nearly
// every row) v is exact as a double, so one rounding to float and an exact scale by 2^-60 give
// it; above, integer operations do.
__device__ __forceinline__ float dequantize(unsigned long long v) {
  const uint32_t hi = (uint32_t)(v >> 32);
  if (hi < (1u << 21)) return (float)__builtin_fma((double)hi, 4294967296.0, (double)(uint32_t)v) * 0x1p-60f;
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

// Weighted sweep: an entry's value w * v may be negative (the adjoint) and of any magnitude.
// The solve picks a power of two 2^t that maps every product and every partial row sum into
// [-1/2, 1/2] (a bound it proves from the weights and its right-hand side), and the scaled
// value is summed in two's complement fixed point in units of 2^-60: exact integer adds,
// independent of order and split, no overflow (|sum| < 2^59 units).
__device__ __forceinline__ unsigned long long quantize_signed(float v, int t) {
  const float sc = __builtin_ldexpf(v, t);  // a power-of-two scale: exact (tiny values may flush)
  const unsigned long long q = quantize(fabsf(sc));
  return sc < 0.f ? 0ull - q : q;
}
__device__ __forceinline__ float dequantize_signed(unsigned long long v, int t) {
  const bool neg = (long long)v < 0;
  const float f = __builtin_ldexpf(dequantize(neg ? 0ull - v : v), -t);
  return neg ? -f : f;
}

// One graph: every workgroup of sweep k exits when sweep k-1's change was below the
// threshold.  Several graphs (row_graph set): the passes always run and each graph's rows
// are frozen once that graph has converged (graph_frozen).
// One graph's change of sweep k is the sum of its kBinAccLanes lane counters; every wave
// loads and sums them (exactly, in fixed point), so all waves of a workgroup agree.
template <bool W = false>
__device__ __forceinline__ bool converged_before(const BinParams& P) {
  static_assert(kBinAccLanes == 64, "one lane counter per wavefront lane");
  if constexpr (W) return P.iter > 0 && P.werr[P.iter - 1] < P.threshold * (P.thr_scale ? *P.thr_scale : 1.0);
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
template <int NT, bool W = false>
__device__ __forceinline__ double finish_row(const BinParams& P, uint32_t i, float z) {
  if constexpr (W) {  // x' = fma(d, z, base_i) from the previous iterate cin into x
    const float xo = bld<NT>(P.cin + i);
    const float xn = __builtin_fmaf(P.damp, z, bld<NT>(P.basev + i));
    bst<NT>(P.x + i, xn);
    return (double)fabsf(xn - xo);
  }
  const float b = P.basev ? bld<NT>(P.basev + i) : P.base;
  const float xo = bld<NT>(P.x + i);
  const float xn = __builtin_fmaf(P.damp, z, b);
  bst<NT>(P.x + i, xn);
  bst<NT>(P.cout + i, xn / (float)bld<NT>(P.deg + i));
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
// W (weighted): the workgroup's L1 change is stored in P.wpart[slot] instead of a lane counter.
template <int NT, int MAXR, bool MULTI, int HINT, bool W = false, class Z>
__device__ __forceinline__ void finish_rows(const BinParams& P, uint32_t r0, uint32_t nrows, Z z, double* red,
                                            unsigned long long* gerr, uint32_t slot = 0) {
  static_assert(MAXR % NT == 0, "whole rounds of rows");
  static_assert(!(MULTI && W), "the weighted sweep solves one graph");
  if constexpr (!MULTI) {
    (void)gerr;
    double e = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < MAXR / NT; ++k) {
      const uint32_t v = threadIdx.x + k * NT;
      if (v < nrows) e += finish_row<HINT, W>(P, r0 + v, z(v));
    }
    e = wave_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (uint32_t q = 0; q < NT / 64; ++q) t += red[q];
      if constexpr (W) P.wpart[slot] = t;
      else if (t > 0.0) atomicAdd(residual_lane(P), err_to_fixed(t));
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
      const double e = finish_row<HINT>(P, r0 + v, z(v));
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


// chunk pass: vals[p] = contrib[src(p)] over the pairs of the tiles of one item.  Tiles are
// padded to multiples of 8 pairs (pads have source kPadSrc and get 0), whole groups of 8 pairs
// never cross a tile, and a lane moves half a group (4 pairs) with one 8-byte id load and one
// 16-byte value store: consecutive lanes then store consecutive 16 bytes, whole lines, which
// the nontemporal stores of the plan's stream_hints 2 need (a lane storing a whole group
// with two 16-byte stores wrote every line in two halves, 343 against 315 us at RMAT-25).
template <int NT, bool W = false>
__global__ __launch_bounds__(kBT) void pr_chunk_kernel(BinParams P) {
  if (converged_before<W>(P)) return;
  __shared__ float4 cst4[kChunkSrc / 4];
  float* cst = reinterpret_cast<float*>(cst4);
  __shared__ uint32_t sstart[kBinMaxSegs];
  __shared__ uint32_t sflat[kBinMaxSegs + 1];
  const uint4 it = P.items1[blockIdx.x];  // {chunk, first tile, end tile, groups of 8 pairs}
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
    const f32x4* src = reinterpret_cast<const f32x4*>(P.cin + u0);
    f32x4* dst = reinterpret_cast<f32x4*>(cst4);
    f32x4 v[kV];
#pragma unroll
    for (uint32_t k = 0; k < kV; ++k) v[k] = bld<NT>(src + threadIdx.x + k * kBT);
#pragma unroll
    for (uint32_t k = 0; k < kV; ++k) dst[threadIdx.x + k * kBT] = v[k];
  } else {
    for (uint32_t q = threadIdx.x; q < kChunkSrc; q += kBT) cst[q] = (u0 + q < P.n) ? P.cin[u0 + q] : 0.f;
  }
  __syncthreads();
  // half-group hg: 4 pairs of group hg >> 1
  constexpr uint32_t kH = 2 * kGroups1;
  const uint32_t mh = 2 * m;
  for (uint32_t h0 = 0; h0 < mh; h0 += kH * kBT) {
    uint32_t p[kH];
    u32x2 uu[kH];
#pragma unroll
    for (uint32_t h = 0; h < kH; ++h) {
      const uint32_t hg = h0 + h * kBT + threadIdx.x;
      if (hg < mh) {
        const uint32_t g = hg >> 1;
        uint32_t lo = 0, hi = ns - 1;  // last tile j with sflat[j] <= g
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (sflat[mid] <= g) lo = mid;
          else hi = mid - 1;
        }
        p[h] = sstart[lo] + 8u * (g - sflat[lo]) + 4u * (hg & 1u);
        uu[h] = bld<NT>(reinterpret_cast<const u32x2*>(P.bsrc + p[h]));
      }
    }
#pragma unroll
    for (uint32_t h = 0; h < kH; ++h) {
      if (h0 + h * kBT + threadIdx.x >= mh) continue;
      const uint32_t w[2] = {uu[h].x, uu[h].y};
      float o[4];
#pragma unroll
      for (uint32_t t = 0; t < 4; ++t) {
        const uint32_t u = (w[t >> 1] >> (16 * (t & 1))) & 0xFFFFu;
        o[t] = u == kPadSrc ? 0.f : cst[u];
      }
      bst<NT>(reinterpret_cast<f32x4*>(P.vals + p[h]), f32x4{o[0], o[1], o[2], o[3]});
    }
  }
}

// One wave's block of kBinBlock entries: the lane's 8 row offsets (one 16-byte load) and the
// block's first pair.
struct BinBlock {
  u32x4 e;
  uint32_t pb;
};

template <int NT>
__device__ __forceinline__ BinBlock bin_block_load(const BinParams& P, uint32_t k, uint32_t lane) {
  BinBlock b;
  b.e = bld<NT>(reinterpret_cast<const u32x4*>(P.brow + k * kBinBlock + 8u * lane));
  // the block's first pair as a vector load (every lane the same address, one request): a
  // scalar load would have to be waited on with the LDS adds (scalar loads return out of
  // order), which serialises the pipeline below
  uint32_t kv = k;
  asm volatile("" : "+v"(kv));
  b.pb = P.bpid[kv];
  return b;
}

// the weighted sweep's 8 weights of the lane in block k: 32 bytes, 32-byte aligned
struct BinWeights {
  f32x4 w[2];
};
template <int NT>
__device__ __forceinline__ void bin_weights_load(const BinParams& P, uint32_t k, uint32_t lane, BinWeights& wt) {
  const f32x4* wp = reinterpret_cast<const f32x4*>(P.bw + k * kBinBlock + 8u * lane);
  wt.w[0] = bld<NT>(wp);
  wt.w[1] = bld<NT>(wp + 1);
}

// The ballot of a 32-bit word's bit 15 over the wave (one compare of its low half as a
// signed 16-bit integer; the compiler tests the bit with two instructions).  Every lane is
// active where the bin pass calls it.
__device__ __forceinline__ unsigned long long ballot_bit15(uint32_t w) {
  unsigned long long m;
  asm("v_cmp_gt_i16_e64 %0, 0, %1" : "=s"(m) : "v"(w));
  return m;
}

// Issues the loads of a block's pair data for the lane's 8 entries (not waited on): a 32-bit
// word per entry from base[pair - sub], the pair's value (P.vals, sub 0) or in a hot tile the
// pair's source id (P.hsrc, sub the item's hot offset).  One load form for both, so that a
// pipeline slot's registers are written by one instruction whichever the block is (a branch
// between two load forms into the same registers made the compiler drain every load in
// flight before the hot one).  The pair of an entry is the block's first pair plus the pair
// starts after the block's first entry up to the entry.  Slot t of lane l is entry 64 t + l
// (bin_slot): its pair is the block's first pair, plus the starts of slots 0 .. t-1 (a
// popcount of their ballots), plus the starts of slot t in lanes 0 .. l.  With m the slot's
// ballot, the last term is bit 0 of m plus mbcnt(m >> 1) (the set bits 1 .. l of m).  The bin
// pass is bound by vector instructions: per entry this is the flag test, the two mbcnt steps
// and the address; the rest is scalar.
__device__ __forceinline__ void bin_block_pairs(const BinBlock& b, const char* base, uint32_t sub,
                                                uint32_t (&raw)[8]) {
  const uint32_t wd[4] = {b.e.x, b.e.y, b.e.z, b.e.w};
  uint32_t first = __builtin_amdgcn_readfirstlane(b.pb) - sub;
#pragma unroll
  for (uint32_t t = 0; t < 8; ++t) {
    // the pair-start flag: bit 15 of the entry's 16-bit half word
    unsigned long long m = (t & 1) ? __ballot((int32_t)wd[t >> 1] < 0) : ballot_bit15(wd[t >> 1]);
    if (t == 0) m &= ~1ull;  // the block's first entry is in pair pb by definition
    const unsigned long long m1 = m >> 1;
    const uint32_t c = __builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
    raw[t] = *reinterpret_cast<const uint32_t*>(base + ((c + first + (uint32_t)(m & 1ull)) << 2));
    first += (uint32_t)__popcll(m);
  }
}

// Adds a block's 8 entries of this lane into the LDS row accumulators (W: each times its
// weight).
template <bool HOT, bool W, int BB>
__device__ __forceinline__ void bin_block_add(const BinBlock& b, const uint32_t (&raw)[8], const BinWeights& wt,
                                              unsigned long long* lacc, const float* hot, int qs) {
  const uint32_t wd[4] = {b.e.x, b.e.y, b.e.z, b.e.w};
  float hv[8];
  if constexpr (HOT) {  // the 8 hot contributions first: one wait for the LDS reads
#pragma unroll
    for (uint32_t t = 0; t < 8; ++t) hv[t] = hot[min(raw[t], kHotSrc - 1u)];
  }
#pragma unroll
  for (uint32_t t = 0; t < 8; ++t) {
    const uint32_t row = (wd[t >> 1] >> (16 * (t & 1))) & (bin_rows<BB>() - 1u);
    float v;
    if constexpr (HOT) v = raw[t] < kHotSrc ? hv[t] : 0.f;  // pads (source kPadSrc) add 0
    else v = __uint_as_float(raw[t]);
    if constexpr (W) atomicAdd(&lacc[row], quantize_signed(wt.w[t >> 2][t & 3] * v, qs));
    else atomicAdd(&lacc[row], quantize(v));
  }
}

// Pipeline depths of the bin pass: row offsets kDE blocks ahead, pair loads kDV blocks ahead,
// in rings of kRE and kRV buffers (the loop is unrolled by lcm(kRE, kRV)).  Measured on MI355X
// at RMAT-25 (2^14-row bins), ms per sweep: depths 3/1 0.631, 2/1 0.632, 4/2 0.634, 5/2 0.636,
// 7/3 0.643, 9/4 0.650, 11/5 0.657; rings whose lcm unrolls the loop 45 or 66 times 0.71-0.77
// (code size).  A wave's loads of its next blocks overlap the other waves' adds: 16 waves per
// workgroup keep enough in flight without deep rings.  Bins of 2^13 rows (two workgroups per
// CU): 2/1 (3/1 and 3/2 were slower on the 4096^2 grid).
template <bool W, int BB>
struct BinDepth {
  static constexpr uint32_t DE = BB >= 14 ? 3 : 2, DV = 1;
  static constexpr uint32_t RE = 4, RV = 2;
  static_assert(RE > DE && RV > DV && DE > DV, "pipeline rings must hold the blocks in flight");
};
constexpr uint32_t gcd_u32(uint32_t a, uint32_t b) { return b ? gcd_u32(b, a % b) : a; }
template <class F, uint32_t... Ks>
__device__ __forceinline__ void unroll_steps(F& f, std::integer_sequence<uint32_t, Ks...>) {
  (f(std::integral_constant<uint32_t, Ks>{}), ...);
}
// blocks of a bin-pass range handed out before its block counter: kDE + 1 per wave
template <bool W, int BB>
constexpr uint32_t bin_static_blocks() {
  return (BinDepth<W, BB>::DE + 1) * (kBT / 64);
}

// Adds the entries [e0, e1) of one bin into lacc; e0, he and e1 are block boundaries (the
// layout aligns bins and hot tiles).  The blocks of [e0, he) lie in the bin's hot tile, where
// a pair's value is hot[source] (pads, source kPadSrc, add 0); those of [he, e1) read it from
// P.vals[pair].  Each wave runs its blocks through one software pipeline, so that the loads of
// several blocks overlap: while block j's entries are added, the pair loads of the next kDV
// blocks and the row offsets of the next kDE are in flight.  Wave w starts with blocks
// w, w + 16, ..., w + 16 kDE of the range; after that, with bins of 2^14 rows, each wave takes
// the next block from the workgroup's counter *next (set to the range's first block +
// bin_static_blocks() before the workgroup's barrier), one block per step and one step before
// it is loaded.  (With a fixed share of blocks per wave, the waves the SIMDs' issue order
// favours finished their shares up to a sixth of the item earlier than the others, and the
// CU's one workgroup waited for the last.)  A wave stops at its first block past the range;
// blocks past the range are loaded as the range's last block, so that the loads stay in
// bounds, and never added.  Entries are ordered by source, so one row rarely repeats inside a
// lane or an instruction: each entry is one LDS add, with no run tracking.  The sums are
// exact, so which wave adds a block does not matter.  sync() is the workgroup's set-up
// (clearing lacc, staging hot) and barrier: every wave calls it once, after issuing its first
// blocks' row offsets, which need neither.  (Their pair loads wait on the row offsets and
// follow the barrier: waiting before it held every wave there until the slowest had its row
// offsets.)
template <int NT, bool W, int BB, class Sync>
__device__ __forceinline__ void bin_range(const BinParams& P, uint32_t e0, uint32_t he, uint32_t e1, uint32_t hd,
                                          unsigned long long* lacc, const float* hot, uint32_t* next, Sync&& sync) {
  // the wave index is uniform over the wave: in a scalar register, the block indices and the
  // blocks' first pairs stay scalar too
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  constexpr uint32_t kWaves = kBT / 64;
  const uint32_t kb = e0 / kBinBlock + w, kh = he / kBinBlock, ke = e1 / kBinBlock;
  if (e0 >= e1) {  // an empty range (a bin without entries): nothing to load
    sync();
    return;
  }
  // values are addressed by 32-bit byte offsets (the layout keeps pairs below 2^30)
  const char* vbase = reinterpret_cast<const char*>(P.vals);
  const char* hbase = reinterpret_cast<const char*>(P.hsrc);
  int qs = 0;  // the weighted sweep's scale exponent (uniform: one scalar load)
  if constexpr (W) qs = __builtin_amdgcn_readfirstlane(*P.qshift);
  auto clamp = [&](uint32_t k) { return k < ke ? k : ke - 1u; };
  // The block of pipeline slot t is kq[t % kRE], its row offsets eb[t % kRE], its pair loads
  // rb[t % kRV] and (weighted) its weights wb[t % kRV]; the loop is unrolled by
  // lcm(kRE, kRV) so that every buffer keeps fixed registers (a register copy of a load still
  // in flight would wait for it).  Step t takes slot t+kDE's block (fetched from the counter
  // one step earlier, into nk), fetches the next, issues slot t+kDE's row offsets and slot
  // t+kDV's pair loads (and weights), then adds slot t; the scheduling barriers keep the
  // compiler from hoisting the adds, which wait for slot t's loads, above the newer loads.
  // Whether a block is hot is uniform over the wave.
  using D = BinDepth<W, BB>;
  constexpr uint32_t kDE = D::DE, kDV = D::DV, kRE = D::RE, kRV = D::RV;
  uint32_t kq[kRE];
  BinBlock eb[kRE];
  uint32_t rb[kRV][8];
  BinWeights wb[kRV];
  auto pairs = [&](uint32_t k, const BinBlock& e, uint32_t(&r)[8]) {
    const bool h = k < kh;  // uniform: the selects are scalar
    bin_block_pairs(e, h ? hbase : vbase, h ? hd : 0u, r);
  };
  auto add = [&](uint32_t k, const BinBlock& e, const uint32_t(&r)[8], const BinWeights& wt) {
    if (k < kh) bin_block_add<true, W, BB>(e, r, wt, lacc, hot, qs);
    else bin_block_add<false, W, BB>(e, r, wt, lacc, hot, qs);
  };
  // Bins of 2^13 rows keep two workgroups per CU, and the other fills the CU while one
  // drains: their waves keep the fixed shares w, w + 16, ... (measured: the counter costs the
  // 4096^2 grid 2 %).
  constexpr bool kDyn = BB >= 14;
  uint32_t nk = kb + kDE * kWaves;  // slot kDE's block: the wave's last static one
  // a zero the compiler cannot see as uniform: lane 0's counter fetch stays one LDS atomic
  // whose result is read a step later (an atomic at a uniform address is rewritten to one
  // whose result is broadcast at once, which waits for every LDS add in flight)
  uint32_t zv = 0u;
  asm volatile("" : "+v"(zv));
  auto step = [&](auto kc) {
    constexpr uint32_t K = decltype(kc)::value;
    constexpr uint32_t tE = (K + kDE) % kRE, tV = (K + kDV) % kRV, tVE = (K + kDV) % kRE;
    kq[tE] = __builtin_amdgcn_readfirstlane(nk);
    if constexpr (kDyn) {
      if (lane == 0) nk = atomicAdd(next + zv, 1u);
    } else {
      nk += kWaves;
    }
    eb[tE] = bin_block_load<NT>(P, clamp(kq[tE]), lane);
    if constexpr (W) bin_weights_load<NT>(P, clamp(kq[tVE]), lane, wb[tV]);
    pairs(clamp(kq[tVE]), eb[tVE], rb[tV]);
    __builtin_amdgcn_sched_barrier(0);
    if (kq[K % kRE] < ke) add(kq[K % kRE], eb[K % kRE], rb[K % kRV], wb[K % kRV]);
    __builtin_amdgcn_sched_barrier(0);
  };
  // the first blocks' row offsets are issued before the workgroup's set-up and barrier (sync)
  // and are in flight over it
#pragma unroll
  for (uint32_t d = 0; d < kDE; ++d) {
    kq[d] = kb + d * kWaves;
    eb[d] = bin_block_load<NT>(P, clamp(kq[d]), lane);
  }
  sync();
  if (kb >= ke) return;  // a wave past the range's last block (its loads were the last block's)
#pragma unroll
  for (uint32_t d = 0; d < kDV; ++d) {
    if constexpr (W) bin_weights_load<NT>(P, clamp(kq[d]), lane, wb[d]);
    pairs(clamp(kq[d]), eb[d], rb[d]);
  }
  constexpr uint32_t kU = kRE * kRV / gcd_u32(kRE, kRV);
  do unroll_steps(step, std::make_integer_sequence<uint32_t, kU>{});
  while (kq[0] < ke);
}

// The finishing loads of one graph's rows r0 + threadIdx.x + k * kBT (k < 2^BB / kBT):
// issued before the bin pass's last barrier, so that they are in flight while the other
// waves finish their adds.
template <int BB>
struct FinishIn {
  float xo[bin_rows<BB>() / kBT];
  float bv[bin_rows<BB>() / kBT];
  uint32_t dg[bin_rows<BB>() / kBT];
};
template <int NT, int BB>
__device__ __forceinline__ void finish_load(const BinParams& P, uint32_t r0, uint32_t nrows, FinishIn<BB>& in) {
#pragma unroll
  for (uint32_t k = 0; k < bin_rows<BB>() / kBT; ++k) {
    const uint32_t v = threadIdx.x + k * kBT;
    if (v < nrows) {
      in.xo[k] = bld<NT>(P.x + r0 + v);
      in.dg[k] = bld<NT>(P.deg + r0 + v);
      in.bv[k] = P.basev ? bld<NT>(P.basev + r0 + v) : P.base;
    }
  }
}
// the reference's row update (page_rank.hpp:133-135) on the prefetched rows, and the
// workgroup's L1 change into a lane counter of the sweep (as finish_rows)
template <int NT, int BB>
__device__ __forceinline__ void finish_apply(const BinParams& P, uint32_t r0, uint32_t nrows, const FinishIn<BB>& in,
                                             const unsigned long long* lacc, double* red) {
  double e = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < bin_rows<BB>() / kBT; ++k) {
    const uint32_t v = threadIdx.x + k * kBT;
    if (v < nrows) {
      const float xn = __builtin_fmaf(P.damp, dequantize(lacc[v]), in.bv[k]);
      bst<NT>(P.x + r0 + v, xn);
      bst<NT>(P.cout + r0 + v, xn / (float)in.dg[k]);
      e += (double)fabsf(xn - in.xo[k]);
    }
  }
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (uint32_t q = 0; q < kBT / 64; ++q) t += red[q];
    if (t > 0.0) atomicAdd(residual_lane(P), err_to_fixed(t));
  }
}

// bin pass: row sums of one entry range of one bin in LDS; a bin that is one item finishes
// its rows here, an item of a split bin stores its partial rows in its slot of P.part.
// The range's part inside the bin's hot tile (a prefix of the bin) reads its values from
// the hot contributions staged in LDS.  An item's hot end (P.item_he) is precomputed, so that
// no load waits on another before the hot contributions are staged.
// (8 waves per SIMD: 64 KiB of row sums allow two workgroups per CU, which 64 VGPRs keep; the
// weighted sweep's 2^13-row bins too, 3 % faster than one workgroup per CU with more
// registers at RMAT-20.  Bins of 2^14 rows take 128 KiB of LDS: one workgroup per CU.)
template <bool MULTI, int NT, bool W, int BB>
__global__ __launch_bounds__(kBT, (BB >= 14) ? 4 : 8) void pr_bin_kernel(BinParams P) {
  constexpr uint32_t kBinRows = bin_rows<BB>();
  if (converged_before<W>(P)) return;
  __shared__ unsigned long long lacc[kBinRows];
  __shared__ double red[kBT / 64];
  __shared__ unsigned long long gerr[MULTI ? kGraphSlots : 1];
  __shared__ float hot[MULTI ? 1 : kHotSrc];
  __shared__ uint32_t next;  // the range's next block to hand out
  const uint4 it = P.items2[blockIdx.x];  // {bin, first entry, end entry, partial slot or kNoSlot}
  const uint32_t b = it.x;
  // end of the hot part of [it.y, it.z); the hot contributions' loads are issued first, then
  // the first blocks' loads, and the contributions are written to LDS once the row sums are
  // cleared (bin_range's sync)
  uint32_t he = it.y, hd = 0;
  constexpr uint32_t kR = (kHotSrc + kBT - 1) / kBT;
  float hv[kR];
  if constexpr (!MULTI) {
    he = P.item_he[blockIdx.x];
    hd = P.item_hd[blockIdx.x];
    if (he > it.y) {
#pragma unroll
      for (uint32_t k = 0; k < kR; ++k) {
        const uint32_t q = threadIdx.x + k * kBT;
        hv[k] = q < P.n_hot ? bld<NT>(P.cin + q) : 0.f;
      }
    }
  }
  bin_range<NT, W, BB>(P, it.y, he, it.z, hd, lacc, hot, &next, [&]() {
    for (uint32_t q = threadIdx.x; q < kBinRows; q += kBT) lacc[q] = 0ull;
    if (threadIdx.x == 0) next = it.y / kBinBlock + bin_static_blocks<W, BB>();
    if constexpr (!MULTI) {
      if (he > it.y) {
#pragma unroll
        for (uint32_t k = 0; k < kR; ++k) {
          const uint32_t q = threadIdx.x + k * kBT;
          if (q < kHotSrc) hot[q] = hv[k];
        }
      }
    }
    __syncthreads();
  });
  const uint32_t r0 = b << BB;
  const uint32_t nrows = r0 < P.n ? min(kBinRows, P.n - r0) : 0u;
  if constexpr (!MULTI && !W) {
    if (it.w == kNoSlot) {  // one graph, a bin that is one item: prefetch the rows, then finish
      FinishIn<BB> in;
      finish_load<NT, BB>(P, r0, nrows, in);
      __syncthreads();
      finish_apply<NT, BB>(P, r0, nrows, in, lacc, red);
      return;
    }
  }
  __syncthreads();
  if (it.w != kNoSlot) {
    unsigned long long* dst = P.part + (uint64_t)it.w * kBinRows;
    for (uint32_t v = threadIdx.x; v < kBinRows; v += kBT) bst<NT>(dst + v, lacc[v]);
    if constexpr (W)
      if (threadIdx.x == 0) P.wpart[blockIdx.x] = 0.0;  // this item finishes no rows
    return;
  }
  if constexpr (W)
    finish_rows<kBT, kBinRows, MULTI, NT, true>(
        P, r0, nrows, [&, qs = *P.qshift](uint32_t v) { return dequantize_signed(lacc[v], qs); }, red, gerr,
        blockIdx.x);
  else
    finish_rows<kBT, kBinRows, MULTI, NT>(P, r0, nrows, [&](uint32_t v) { return dequantize(lacc[v]); }, red, gerr);
}

// rows of the bins split over several bin-pass items: add the items' partial rows (exact,
// so in any order) and finish the rows.  A workgroup (one combine unit) takes 256 / L rows
// of one split bin, each row's slots shared by L lanes (slot k on lane k mod L); the host
// picks L per bin from its slot count, so a bin with hundreds of slots is read by many
// workgroups at once and a bin with a few by few.  One workgroup per unit.
template <bool MULTI, int NT, bool W, int BB>
__global__ __launch_bounds__(256) void pr_bin_combine_kernel(BinParams P) {
  constexpr uint32_t kBinRows = bin_rows<BB>();
  if (converged_before<W>(P)) return;
  __shared__ double red[4];
  __shared__ unsigned long long gerr[MULTI ? kGraphSlots : 1];
  __shared__ unsigned long long psum[256];
  const uint4 it = P.items3[blockIdx.x];  // {bin, first slot, slots, first row | log2 L << 16}
  const uint32_t lanes = 1u << (it.w >> 16), rows = 256u >> (it.w >> 16);
  const uint32_t v0 = it.w & 0xFFFFu;
  const uint32_t r0 = (it.x << BB) + v0;
  const uint32_t nrows = r0 < P.n ? min(rows, P.n - r0) : 0u;
  {
    const uint32_t row = threadIdx.x % rows, lane = threadIdx.x / rows;
    const unsigned long long* src = P.part + (uint64_t)it.y * kBinRows + v0 + row;
    unsigned long long s0 = 0ull, s1 = 0ull, s2 = 0ull, s3 = 0ull;
    uint32_t k = lane;
    for (; k + 3 * lanes < it.z; k += 4 * lanes) {
      s0 += bld<NT>(src + (uint64_t)k * kBinRows);
      s1 += bld<NT>(src + (uint64_t)(k + lanes) * kBinRows);
      s2 += bld<NT>(src + (uint64_t)(k + 2 * lanes) * kBinRows);
      s3 += bld<NT>(src + (uint64_t)(k + 3 * lanes) * kBinRows);
    }
    for (; k < it.z; k += lanes) s0 += bld<NT>(src + (uint64_t)k * kBinRows);
    psum[threadIdx.x] = (s0 + s1) + (s2 + s3);
  }
  __syncthreads();
  finish_rows<256, 256, MULTI, NT, W>(
      P, r0, nrows,
      [&](uint32_t v) {
        unsigned long long t = 0ull;
        for (uint32_t q = 0; q < lanes; ++q) t += psum[q * rows + v];
        if constexpr (W) return dequantize_signed(t, *P.qshift);
        else return dequantize(t);
      },
      red, gerr, P.n_items2 + blockIdx.x);
}

// weighted sweep: the sweep's L1 change, the bin items' and combine units' partials added in
// a fixed order (strided per lane, then a fixed tree)
__global__ __launch_bounds__(256) void pr_wfold_kernel(BinParams P) {
  if (converged_before<true>(P)) return;
  __shared__ double red[4];
  const uint32_t m = P.n_items2 + P.n_items3;
  double e = 0.0;
  for (uint32_t q = threadIdx.x; q < m; q += 256) e += P.wpart[q];
  e = block_sum<256>(e, red);
  if (threadIdx.x == 0) P.werr[P.iter] = e;
}

// ---- layout construction ----
// sort key of CSC position p (row v, source u): (bin of v, u), and v's offset in its bin.
// A stable sort by this key orders each bin by source, then row (CSC order).
// With pos, low[p] is p itself (a weighted layout looks up each entry's weight and row).
__global__ __launch_bounds__(256) void bin_key_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ rp,
                                                      const uint32_t* __restrict__ col, int ub, int bb, bool pos,
                                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ low) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= nnz) return;
  uint32_t lo = 0, hi = n;  // row of p: first r with rp[r + 1] > p
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (rp[mid + 1] <= p) lo = mid + 1;
    else hi = mid;
  }
  keys[p] = ((uint64_t)(lo >> bb) << ub) | col[p];
  low[p] = pos ? (uint32_t)p : (lo & ((1u << bb) - 1u));
}

// tile of a sort key (tile column: 0 for a source below n_hot, else 1 + chunk; without hot
// tiles the chunk)
__device__ __forceinline__ uint64_t tile_of_key(uint64_t key, int ub, uint32_t TC, uint32_t n_hot) {
  const uint32_t u = (uint32_t)(key & ((1ull << ub) - 1));
  const uint32_t c = u < n_hot ? 0u : (u >> kChunkBits) + (n_hot ? 1u : 0u);
  return (key >> ub) * TC + c;
}

// flag[q] = 1 when sorted entry q starts a pair (its key differs from the previous one's)
__global__ __launch_bounds__(256) void pair_flag_kernel(uint64_t nnz, const uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ flag) {
  const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= nnz) return;
  flag[q] = (q == 0 || keys[q] != keys[q - 1]) ? 1u : 0u;
}

// tp[t] = first sorted position whose tile is >= t, tpp[t] = pairs before it (pex: exclusive
// scan of the pair flags, nnz + 1 long), for t in [0, n_tiles]
__global__ __launch_bounds__(256) void tile_ptr_kernel(uint64_t nnz, const uint64_t* __restrict__ keys, int ub,
                                                       uint32_t TC, uint32_t n_hot, uint64_t n_tiles,
                                                       const uint32_t* __restrict__ pex, uint32_t* __restrict__ tp,
                                                       uint32_t* __restrict__ tpp) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t > n_tiles) return;
  uint64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (tile_of_key(keys[mid], ub, TC, n_hot) < t) lo = mid + 1;
    else hi = mid;
  }
  tp[t] = (uint32_t)lo;
  tpp[t] = pex[lo];
}

// sorted entry q of tile t goes to entry etp[t] + (q - tp[t]) (stored at its bin_slot); its pair,
// numbered inside the tile from tpp[t], to pair ptp[t] + ...; the first entry of a kBinBlock
// block records its pair
__global__ __launch_bounds__(256) void bin_place_kernel(uint64_t nnz, const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ low,
                                                        const uint32_t* __restrict__ flag,
                                                        const uint32_t* __restrict__ pex, int ub, uint32_t TC,
                                                        uint32_t n_hot, const uint32_t* __restrict__ tp,
                                                        const uint32_t* __restrict__ tpp,
                                                        const uint32_t* __restrict__ etp,
                                                        const uint32_t* __restrict__ ptp,
                                                        uint16_t* __restrict__ brow, uint16_t* __restrict__ bsrc,
                                                        uint32_t* __restrict__ bpid, const uint32_t* __restrict__ rp,
                                                        uint32_t n, int bb, const float* __restrict__ w,
                                                        float* __restrict__ bw) {
  const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= nnz) return;
  const uint64_t key = keys[q];
  const uint64_t t = tile_of_key(key, ub, TC, n_hot);
  const uint32_t f = flag[q];
  const uint32_t ne = etp[t] + (uint32_t)(q - tp[t]);
  const uint32_t np = ptp[t] + (pex[q] + f - 1u - tpp[t]);
  uint32_t lw = low[q];
  if (w) {  // lw is the entry's CSC position: its weight, and its row's offset in the bin
    const uint32_t p = lw;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (rp[mid + 1] <= p) lo = mid + 1;
      else hi = mid;
    }
    lw = lo & ((1u << bb) - 1u);
    bw[bin_slot(ne)] = w[p];
  }
  brow[bin_slot(ne)] = (uint16_t)(lw | (f ? kPairStart : 0u));
  if (f) {
    const uint32_t u = (uint32_t)(key & ((1ull << ub) - 1));
    bsrc[np] = (uint16_t)(u < n_hot ? u : (u & (kChunkSrc - 1)));
  }
  if (ne % kBinBlock == 0) bpid[ne / kBinBlock] = np;
}

// pads of tile t: the tile's pairs are padded to a multiple of 8 with pairs of source
// kPadSrc (value 0), each given one entry (row 0, pair start) after the tile's entries.
// Entries that fill a tile up to a block boundary stay as cleared (row 0, no pair start):
// they join the last pad pair, and no block starts among them.
__global__ __launch_bounds__(256) void bin_pad_kernel(uint64_t n_tiles, const uint32_t* __restrict__ tp,
                                                      const uint32_t* __restrict__ tpp,
                                                      const uint32_t* __restrict__ etp,
                                                      const uint32_t* __restrict__ ptp, uint16_t* __restrict__ brow,
                                                      uint16_t* __restrict__ bsrc, uint32_t* __restrict__ bpid) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n_tiles) return;
  const uint32_t ne = tp[t + 1] - tp[t], np = tpp[t + 1] - tpp[t];
  const uint32_t pads = (ptp[t + 1] - ptp[t]) - np;
  for (uint32_t j = 0; j < pads; ++j) {
    const uint32_t e = etp[t] + ne + j, p = ptp[t] + np + j;
    brow[bin_slot(e)] = (uint16_t)kPairStart;
    bsrc[p] = (uint16_t)kPadSrc;
    if (e % kBinBlock == 0) bpid[e / kBinBlock] = p;
  }
}

// the hot tile of bin b (pairs ptp[b TC] ..) as 32-bit words at hsrc[hbase[b] ..]; one
// workgroup per bin
__global__ __launch_bounds__(256) void hot_src_kernel(const uint32_t* __restrict__ ptp, uint32_t TC,
                                                      const uint32_t* __restrict__ hbase,
                                                      const uint16_t* __restrict__ bsrc, uint32_t* __restrict__ hsrc) {
  const uint32_t b = blockIdx.x;
  const uint32_t p0 = ptp[(uint64_t)b * TC], p1 = ptp[(uint64_t)b * TC + 1];
  for (uint32_t j = threadIdx.x; j < p1 - p0; j += 256) hsrc[hbase[b] + j] = bsrc[p0 + j];
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

void binned_build(hipStream_t s, const ActiveCsc& a, uint32_t n, uint64_t nnz, BinnedLayout& L, bool hot,
                  const float* w, int bin_bits) {
  NWG_REQUIRE(bin_bits == 0 || bin_bits == 13 || bin_bits == 14, "binned layout: bins of 2^13 or 2^14 rows");
  NWG_REQUIRE(nnz < (1ull << 31), "binned layout needs fewer than 2^31 entries");
  L = BinnedLayout();
  // Bins of 2^14 rows have fewer pairs per entry (0.35 against 0.43 at RMAT-25, so 17 % fewer
  // pair values to store and read), but their 128 KiB of row sums leave one bin-pass
  // workgroup per CU.  That pays for the largest graphs only: measured on MI355X, RMAT-25
  // (529 M entries) 0.67 against 0.72 ms per sweep, while the 256-graph union (213 M) and
  // smaller graphs were slower with them.
  L.bin_bits = bin_bits ? bin_bits : nnz >= (1ull << 28) ? 14 : 13;
  const uint32_t bin_rows_ = 1u << L.bin_bits;
  L.n_bins = (uint32_t)ceil_div(n, bin_rows_);
  L.n_chunks = (uint32_t)ceil_div(n, kChunkSrc);
  L.n_hot = hot ? std::min<uint32_t>(kHotSrc, n) : 0u;
  L.nnz = nnz;
  // tile columns per bin: the hot tile (if any), then one per chunk
  const uint32_t hc = L.n_hot ? 1u : 0u, TC = L.n_chunks + hc;
  const uint64_t n_tiles = (uint64_t)L.n_bins * TC;
  // etp / ptp: each tile's first entry and first pair in the padded layout
  std::vector<uint32_t> etp(n_tiles + 1, 0), ptp(n_tiles + 1, 0);
  // hot pairs of bin b: pair p is hsrc[p - hdel[b]] (hot tiles stored back to back)
  std::vector<uint32_t> hdel(L.n_bins, 0);
  uint64_t etotal = 0, ptotal = 0;
  if (nnz) {
    int ub = 1;  // bits of a source id
    while (ub < 32 && (1ull << ub) < n) ++ub;
    DevBuf<uint64_t> keys(nnz), keys_tmp(nnz);
    DevBuf<uint32_t> low(nnz), low_tmp(nnz);
    bin_key_kernel<<<grid256(nnz), 256, 0, s>>>(n, nnz, a.ptr, a.idx, ub, L.bin_bits, w != nullptr, keys.p, low.p);
    NWG_LAUNCH_CHECK();
    int kb = ub;
    while (kb < 64 && (1ull << (kb - ub)) < L.n_bins) ++kb;
    // stable: the entries of one (bin, source) keep their CSC order (row ascending)
    radix_sort_pairs(s, keys.p, keys_tmp.p, low.p, low_tmp.p, nnz, kb);
    // the sort's temporaries now hold the pair flags and their exclusive scan
    uint32_t* flag = low_tmp.p;
    uint32_t* pex = reinterpret_cast<uint32_t*>(keys_tmp.p);  // nnz + 1 <= 2 nnz words
    pair_flag_kernel<<<grid256(nnz), 256, 0, s>>>(nnz, keys.p, flag);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, flag, pex, nnz);
    DevBuf<uint32_t> d_tp(n_tiles + 1), d_tpp(n_tiles + 1), d_etp(n_tiles + 1), d_ptp(n_tiles + 1);
    tile_ptr_kernel<<<grid256(n_tiles + 1), 256, 0, s>>>(nnz, keys.p, ub, TC, L.n_hot, n_tiles, pex, d_tp.p, d_tpp.p);
    NWG_LAUNCH_CHECK();
    std::vector<uint32_t> tp(n_tiles + 1), tpp(n_tiles + 1);
    NWG_HIP(hipMemcpyAsync(tp.data(), d_tp.p, (n_tiles + 1) * 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipMemcpyAsync(tpp.data(), d_tpp.p, (n_tiles + 1) * 4, hipMemcpyDeviceToHost, s));
    NWG_HIP(hipStreamSynchronize(s));
    // A bin's hot tile and the bin itself end on a kBinBlock boundary, so every bin-pass range
    // is made of whole blocks: the entries that fill up to the boundary join the tile's last
    // pad pair (value 0; the tile gets 8 pad pairs if it has none).
    for (uint64_t t = 0; t < n_tiles; ++t) {
      const uint64_t np = tpp[t + 1] - tpp[t];
      uint64_t pp = (np + 7) / 8 * 8;
      uint64_t cur = etotal + (tp[t + 1] - tp[t]) + (pp - np);
      const uint64_t col = t % TC;
      if (((hc && col == 0) || col == TC - 1) && cur % kBinBlock) {
        if (pp == np) {
          pp += 8;
          cur += 8;
        }
        cur = (cur + kBinBlock - 1) / kBinBlock * kBinBlock;
      }
      ptotal += pp;
      etotal = cur;
      // 32-bit entry indices with room for the blocks a bin-pass loop looks past its range,
      // and 32-bit byte offsets of values
      NWG_REQUIRE(etotal < (1ull << 31) && ptotal < (1ull << 30), "binned layout needs fewer than 2^31 padded entries");
      ptp[t + 1] = (uint32_t)ptotal;
      etp[t + 1] = (uint32_t)etotal;
    }
    L.padded = (etotal + kBinBlock - 1) / kBinBlock * kBinBlock + kBinBlock;
    L.pairs = ptotal + 8;
    L.brow.alloc(L.padded);
    L.bsrc.alloc(L.pairs);
    L.bpid.alloc(L.padded / kBinBlock);
    NWG_HIP(hipMemsetAsync(L.brow.p, 0, L.padded * 2, s));
    NWG_HIP(hipMemsetAsync(L.bsrc.p, 0xFF, L.pairs * 2, s));
    NWG_HIP(hipMemsetAsync(L.bpid.p, 0, L.padded / kBinBlock * 4, s));
    if (w) {
      L.bw.alloc(L.padded);
      NWG_HIP(hipMemsetAsync(L.bw.p, 0, L.padded * sizeof(float), s));
    }
    NWG_HIP(hipMemcpyAsync(d_etp.p, etp.data(), (n_tiles + 1) * 4, hipMemcpyHostToDevice, s));
    NWG_HIP(hipMemcpyAsync(d_ptp.p, ptp.data(), (n_tiles + 1) * 4, hipMemcpyHostToDevice, s));
    bin_place_kernel<<<grid256(nnz), 256, 0, s>>>(nnz, keys.p, low.p, flag, pex, ub, TC, L.n_hot, d_tp.p, d_tpp.p,
                                                  d_etp.p, d_ptp.p, L.brow.p, L.bsrc.p, L.bpid.p, a.ptr, n, L.bin_bits,
                                                  w, w ? L.bw.p : nullptr);
    NWG_LAUNCH_CHECK();
    bin_pad_kernel<<<grid256(n_tiles), 256, 0, s>>>(n_tiles, d_tp.p, d_tpp.p, d_etp.p, d_ptp.p, L.brow.p, L.bsrc.p,
                                                    L.bpid.p);
    NWG_LAUNCH_CHECK();
    // the hot tiles' sources as 32-bit words (the bin pass loads them like pair values)
    uint64_t hp = 0;
    std::vector<uint32_t> hbase(L.n_bins, 0);
    if (hc)
      for (uint32_t b = 0; b < L.n_bins; ++b) {
        const uint32_t p0 = ptp[(uint64_t)b * TC], p1 = ptp[(uint64_t)b * TC + 1];
        hbase[b] = (uint32_t)hp;
        hdel[b] = p0 - (uint32_t)hp;
        hp += p1 - p0;
      }
    L.hsrc.alloc(std::max<uint64_t>(hp, 1));
    if (hp) {
      DevBuf<uint32_t> d_hbase(L.n_bins);
      NWG_HIP(hipMemcpyAsync(d_hbase.p, hbase.data(), L.n_bins * 4, hipMemcpyHostToDevice, s));
      hot_src_kernel<<<L.n_bins, 256, 0, s>>>(d_ptp.p, TC, d_hbase.p, L.bsrc.p, L.hsrc.p);
      NWG_LAUNCH_CHECK();
      NWG_HIP(hipStreamSynchronize(s));  // before d_hbase is freed
    }
    NWG_HIP(hipStreamSynchronize(s));
  } else {
    L.hsrc.alloc(1);
    L.padded = kBinBlock;
    L.pairs = 8;
    L.brow.alloc(L.padded);
    L.bsrc.alloc(L.pairs);
    L.bpid.alloc(1);
    NWG_HIP(hipMemsetAsync(L.brow.p, 0, L.padded * 2, s));
    NWG_HIP(hipMemsetAsync(L.bsrc.p, 0xFF, L.pairs * 2, s));
    NWG_HIP(hipMemsetAsync(L.bpid.p, 0, 4, s));
    if (w) {
      L.bw.alloc(L.padded);
      NWG_HIP(hipMemsetAsync(L.bw.p, 0, L.padded * sizeof(float), s));
    }
  }
  const uint32_t C = L.n_chunks;
  // hot tile ends (the hot tile is the first of its bin)
  std::vector<uint32_t> he(L.n_bins);
  for (uint32_t b = 0; b < L.n_bins; ++b) he[b] = etp[(uint64_t)b * TC + hc];
  // bin pass items: every bin, cut into ranges of at most cap2 entries (inner cuts on
  // block boundaries, so no block is walked by two items of one bin); the items of a split bin
  // get consecutive partial slots, and the bin combine units
  // Item size: the power of two nearest to entries / CUs (entries / (4 x CUs) with 2^13-row
  // bins), between 2^15 and 2^kItem2Log.  Big graphs take big items (fewer split bins, so
  // fewer partial rows to store and combine); small ones take small items, so that their heavy
  // bins spread over all CUs.  (Measured on MI355X with 2^13-row bins: 2^19 best at RMAT-25,
  // 2^18 for the 256-graph union, 2^17 at RMAT-23, 2^16 at RMAT-22 and below.  2^14-row bins
  // store twice the partial rows per split item, and their waves take blocks from a counter:
  // RMAT-25 takes 2^21, 0.622 ms per sweep against 0.627 at 2^20, 0.657 at 2^19, 0.71 at
  // 3 x 2^20 and 0.78 at 2^22.)
  int dev = 0, cus = 0;
  NWG_HIP(hipGetDevice(&dev));
  NWG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  uint32_t cap2 = 1u << 15;
  {
    const uint64_t q = etotal / ((L.bin_bits >= 14 ? 1ull : 4ull) * (uint64_t)std::max(cus, 1));
    uint32_t lg = 0;
    while (lg < 63 && (2ull << lg) <= q) ++lg;  // floor(log2 q)
    if (q && (double)q >= 1.4142 * (double)(1ull << lg)) ++lg;
    lg = std::min<uint32_t>(std::max<uint32_t>(lg, 15u), kItem2Log);
    cap2 = 1u << lg;
  }
  std::vector<uint4> it2, it3;
  std::vector<uint32_t> ihe, ihd;
  uint32_t slots = 0;
  for (uint32_t b = 0; b < L.n_bins; ++b) {
    const uint32_t p0 = etp[(uint64_t)b * TC], p1 = etp[(uint64_t)(b + 1) * TC];
    const uint32_t cnt = (uint32_t)std::max<uint64_t>(ceil_div(p1 - p0, cap2), 1);
    if (cnt > 1) {
      // combine units: 256 / L rows each, L lanes per row sharing its cnt slots
      const uint32_t lg = cnt > 32 ? 3u : cnt > 8 ? 2u : 0u;
      for (uint32_t v0 = 0; v0 < bin_rows_; v0 += 256u >> lg) it3.push_back(make_uint4(b, slots, cnt, v0 | (lg << 16)));
    }
    auto cut = [&](uint32_t q) -> uint32_t {
      if (q == 0) return p0;
      if (q == cnt) return p1;
      const uint32_t c = p0 + (uint32_t)((uint64_t)(p1 - p0) * q / cnt);
      return std::max(p0, c / kBinBlock * kBinBlock);
    };
    for (uint32_t q = 0; q < cnt; ++q) {
      const uint32_t c0 = cut(q), c1 = cut(q + 1);
      it2.push_back(make_uint4(b, c0, c1, cnt > 1 ? slots++ : kNoSlot));
      ihe.push_back(std::min(std::max(he[b], c0), c1));  // the item's part of the hot tile ends here
      ihd.push_back(hdel[b]);
    }
  }
  // Items of bins without entries go last.  Their rows have no in-edges, so one sweep takes
  // each of them to its fixed point x = base (fma(d, 0, base)), and the next writes the same
  // contribution base / deg into the other contribution buffer; from the third sweep on
  // nothing in them changes (their L1 change is 0), and the sweeps launch only the items
  // before them.  Degree ordering puts every vertex without in-edges after the others, so
  // that at RMAT-25 half of the bins are empty and the rest hold no tiny bins.
  {
    std::vector<uint4> o2;
    std::vector<uint32_t> ohe, ohd;
    o2.reserve(it2.size());
    ohe.reserve(it2.size());
    ohd.reserve(it2.size());
    for (int pass = 0; pass < 2; ++pass)
      for (uint32_t q = 0; q < (uint32_t)it2.size(); ++q)
        if ((it2[q].z > it2[q].y) == (pass == 0)) {
          o2.push_back(it2[q]);
          ohe.push_back(ihe[q]);
          ohd.push_back(ihd[q]);
        }
    it2.swap(o2);
    ihe.swap(ohe);
    ihd.swap(ohd);
    L.n_items2_live = 0;
    while (L.n_items2_live < it2.size() && it2[L.n_items2_live].z > it2[L.n_items2_live].y) ++L.n_items2_live;
  }
  // chunk pass items: the non-empty tiles of each chunk (long tiles cut into pieces), cut
  // at cap1 pairs or kBinMaxSegs pieces; offsets and sizes in groups of 8 pairs (every
  // padded tile, and cap1, is a whole number of groups)
  // Item size: kItem1Pairs, or smaller when the graph has fewer than 2 chunks per CU, so that
  // every CU gets about 2 items.  An item stages its whole chunk (from L2 or MALL when a chunk
  // is cut into several), so more, smaller items cost staging: measured on MI355X, 2 items
  // per CU beat 4 on the 4096^2 grid (0.179 vs 0.194 ms per sweep) and on RMAT-22 (0.131 vs
  // 0.137), and beat whole chunks on RMAT-22 (0.162) and RMAT-20 (0.050 vs 0.083).
  uint32_t cap1 = kItem1Pairs;
  {
    const uint64_t want = 2ull * (uint64_t)std::max(cus, 1);
    if (C < want) {
      uint64_t chunk_pairs = 0;
      for (uint32_t b = 0; b < L.n_bins; ++b)
        chunk_pairs += ptp[(uint64_t)(b + 1) * TC] - ptp[(uint64_t)b * TC + hc];
      const uint64_t per = (ceil_div(chunk_pairs, want) + 7) / 8 * 8;
      cap1 = (uint32_t)std::min<uint64_t>(kItem1Pairs, std::max<uint64_t>(per, 8192));
    }
  }
  std::vector<uint4> it1;
  std::vector<uint32_t> sst, sfl;
  for (uint32_t c = 0; c < C; ++c) {
    uint32_t first = (uint32_t)sst.size(), flat = 0;
    for (uint32_t b = 0; b < L.n_bins; ++b) {
      const uint32_t t0 = ptp[(uint64_t)b * TC + hc + c], t1 = ptp[(uint64_t)b * TC + hc + c + 1];
      // a long tile is cut into pieces of at most cap1 pairs
      for (uint32_t q0 = t0; q0 < t1; q0 += cap1) {
        const uint32_t len = (std::min<uint32_t>(t1 - q0, cap1)) / 8;
        if ((uint32_t)sst.size() > first &&
            (flat + len > cap1 / 8 || (uint32_t)sst.size() - first >= kBinMaxSegs)) {
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
  L.item_he.alloc(std::max<size_t>(ihe.size(), 1));
  if (!ihe.empty()) NWG_HIP(hipMemcpyAsync(L.item_he.p, ihe.data(), ihe.size() * 4, hipMemcpyHostToDevice, s));
  L.item_hd.alloc(std::max<size_t>(ihd.size(), 1));
  if (!ihd.empty()) NWG_HIP(hipMemcpyAsync(L.item_hd.p, ihd.data(), ihd.size() * 4, hipMemcpyHostToDevice, s));
  if (!it1.empty()) NWG_HIP(hipMemcpyAsync(L.items1.p, it1.data(), it1.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
  if (!sst.empty()) {
    NWG_HIP(hipMemcpyAsync(L.seg_start.p, sst.data(), sst.size() * 4, hipMemcpyHostToDevice, s));
    NWG_HIP(hipMemcpyAsync(L.seg_flat.p, sfl.data(), sfl.size() * 4, hipMemcpyHostToDevice, s));
  }
  NWG_HIP(hipStreamSynchronize(s));
  L.built = true;
}

template <int NT, int BB>
static void enqueue_sweep_hinted(hipStream_t s, const BinParams& P) {
  if (P.n_items1) pr_chunk_kernel<NT><<<P.n_items1, kBT, 0, s>>>(P);
  if (P.row_graph) {
    pr_bin_kernel<true, NT, false, BB><<<P.n_items2, kBT, 0, s>>>(P);
    if (P.n_items3) pr_bin_combine_kernel<true, NT, false, BB><<<P.n_items3, 256, 0, s>>>(P);
  } else {
    // from the third sweep on, the items of bins without entries have nothing left to change
    const uint32_t items = P.iter >= 2 ? P.n_items2_live : P.n_items2;
    if (items) pr_bin_kernel<false, NT, false, BB><<<items, kBT, 0, s>>>(P);
    if (P.n_items3) pr_bin_combine_kernel<false, NT, false, BB><<<P.n_items3, 256, 0, s>>>(P);
  }
  NWG_LAUNCH_CHECK();
}

template <int BB>
static void enqueue_sweep_bits(hipStream_t s, const BinParams& P) {
  if (P.hints >= 2) enqueue_sweep_hinted<2, BB>(s, P);
  else if (P.hints == 1) enqueue_sweep_hinted<1, BB>(s, P);
  else enqueue_sweep_hinted<0, BB>(s, P);
}

void binned_enqueue_sweep(hipStream_t s, const BinParams& P) {
  NWG_REQUIRE(P.bin_bits == 13 || P.bin_bits == 14, "binned sweep: bins of 2^13 or 2^14 rows");
  if (P.bin_bits == 14) enqueue_sweep_bits<14>(s, P);
  else enqueue_sweep_bits<13>(s, P);
}

template <int BB>
static void enqueue_weighted_bits(hipStream_t s, const BinParams& P) {
  pr_bin_kernel<false, 0, true, BB><<<P.n_items2, kBT, 0, s>>>(P);
  if (P.n_items3) pr_bin_combine_kernel<false, 0, true, BB><<<P.n_items3, 256, 0, s>>>(P);
}

void binned_enqueue_weighted_sweep(hipStream_t s, const BinParams& P) {
  NWG_REQUIRE(P.bw && P.wpart && P.werr && P.basev, "weighted sweep needs weights, partials and a base");
  NWG_REQUIRE(P.bin_bits == 13 || P.bin_bits == 14, "binned sweep: bins of 2^13 or 2^14 rows");
  if (P.n_items1) pr_chunk_kernel<0, true><<<P.n_items1, kBT, 0, s>>>(P);
  if (P.bin_bits == 14) enqueue_weighted_bits<14>(s, P);
  else enqueue_weighted_bits<13>(s, P);
  pr_wfold_kernel<<<1, 256, 0, s>>>(P);
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
