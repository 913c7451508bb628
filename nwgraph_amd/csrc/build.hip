// build.hip -- on-device edge list -> CSC / CSR construction.
//
// Pipeline (reference semantics in brackets, pnnl/NWGraph):
//   1. pack key = (row << b) | entry, b = bits(n-1); with clean=1 drop u == v
//      [remove_self_loops, build.hpp:404-410] -- dropping before the sort gives the
//      same edge set as the reference's sort -> uniq -> remove order (tc.cpp:66-68)
//   2. stable LSD radix sort of the 2b-bit keys [lexical_sort_by, build.hpp:57-70]
//   3. with clean=1 drop adjacent equal keys [uniq, build.hpp:392-402]
//   4. row pointers from the sorted keys and entries = low bits
//      [fill_directed<idx> + sort_to_be_indexed, build.hpp:209-232,
//       compressed.hpp:491-502: each row ascending makes the layout unique]
//   5. out-degree = histogram of sources [build_degrees, bench/gapbs/common.hpp:110-121]
#include <algorithm>

#include "device_utils.hpp"
#include "nwg_internal.hpp"

namespace nwg {

namespace {

inline int bits_for(uint32_t n) {
  int b = 1;
  while (b < 32 && (1ull << b) < (uint64_t)n) ++b;
  return b;
}

// flag[i] = keep edge i; also records an id out of range.
__global__ __launch_bounds__(256) void edge_flag_kernel(const uint32_t* __restrict__ src,
                                                        const uint32_t* __restrict__ dst, uint64_t ne, uint32_t n,
                                                        int clean, uint32_t* __restrict__ flag,
                                                        uint32_t* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const uint32_t u = src[i], v = dst[i];
  const bool self = (u == v);
  uint32_t keep = 1;
  if (clean && self) keep = 0;
  if (keep && (u >= n || v >= n)) {
    atomicOr(bad, 1u);
    keep = 0;
  }
  flag[i] = keep;
}

// keys[pos[i]] = (row << b) | entry for kept edges (row = dst for idx=1, src for idx=0)
__global__ __launch_bounds__(256) void pack_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                                   uint64_t ne, int b, int idx, const uint32_t* __restrict__ flag,
                                                   const uint32_t* __restrict__ pos, uint64_t* __restrict__ keys) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne || !flag[i]) return;
  const uint64_t r = idx ? dst[i] : src[i];
  const uint64_t e = idx ? src[i] : dst[i];
  keys[pos[i]] = (r << b) | e;
}

__global__ __launch_bounds__(256) void uniq_flag_kernel(const uint64_t* __restrict__ keys, uint64_t m,
                                                        uint32_t* __restrict__ flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void compact_keys_kernel(const uint64_t* __restrict__ in, uint64_t m,
                                                           const uint32_t* __restrict__ flag,
                                                           const uint32_t* __restrict__ pos,
                                                           uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && flag[i]) out[pos[i]] = in[i];
}

// Row pointers: every key i fills ptr[r] for rows (row(i-1), row(i)]; the last key
// also fills rows (row(m-1), n].  entries = low b bits.
__global__ __launch_bounds__(256) void csx_kernel(const uint64_t* __restrict__ keys, uint64_t m, uint32_t n, int b,
                                                  uint32_t* __restrict__ ptr, uint32_t* __restrict__ ent) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t k = keys[i];
  const uint64_t r = k >> b;
  ent[i] = (uint32_t)(k & ((1ull << b) - 1));
  const int64_t prev = i ? (int64_t)(keys[i - 1] >> b) : -1;
  for (int64_t rr = prev + 1; rr <= (int64_t)r; ++rr) ptr[rr] = (uint32_t)i;
  if (i == m - 1)
    for (uint64_t rr = r + 1; rr <= n; ++rr) ptr[rr] = (uint32_t)m;
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Out-degree histogram of the CSC's source ids [build_degrees, common.hpp:110-121].  A
// grid of 2 workgroups per CU walks contiguous ranges of the ids with 16-byte loads, four
// in flight per lane (16 ids), and adds one per id with a non-returning atomic.  Ids below
// kDegLds are counted in a per-workgroup LDS histogram and flushed once per workgroup: on
// skewed graphs the hubs take most entries, and a million global atomics on one hub's
// counter serialise.  (RMAT's hubs have low ids; in the degree-ordered numbering every hub
// does.)  The ids are read once, so they are loaded non-temporally.
constexpr uint32_t kDegLds = 16384;
constexpr uint32_t kDegThreads = 512;
constexpr uint32_t kDegVec = 4;  // 16-byte loads in flight per lane
__device__ __forceinline__ void degree_count(uint32_t u, uint32_t n, uint32_t* h, uint32_t* deg) {
  if (u < kDegLds) {
    if (u < n) atomicAdd(&h[u], 1u);
  } else if (u < n) {
    atomicAdd(&deg[u], 1u);
  }
}
__global__ __launch_bounds__(kDegThreads) void degree_kernel(const uint32_t* __restrict__ col, uint64_t nnz,
                                                             uint32_t n, uint64_t vec_per_block,
                                                             uint32_t* __restrict__ deg) {
  __shared__ uint32_t h[kDegLds];
  for (uint32_t q = threadIdx.x; q < kDegLds; q += kDegThreads) h[q] = 0u;
  __syncthreads();
  const uint64_t nv = nnz / 4;  // whole 16-byte groups; the nnz % 4 tail is counted by workgroup 0
  const u32x4* c4 = reinterpret_cast<const u32x4*>(col);
  const uint64_t v0 = (uint64_t)blockIdx.x * vec_per_block;
  const uint64_t v1 = v0 + vec_per_block < nv ? v0 + vec_per_block : nv;
  for (uint64_t v = v0 + threadIdx.x; v < v1; v += (uint64_t)kDegThreads * kDegVec) {
    u32x4 a[kDegVec];
#pragma unroll
    for (uint32_t t = 0; t < kDegVec; ++t) {
      const uint64_t q = v + (uint64_t)t * kDegThreads;
      a[t] = q < v1 ? __builtin_nontemporal_load(c4 + q) : u32x4{n, n, n, n};  // id n counts nothing
    }
#pragma unroll
    for (uint32_t t = 0; t < kDegVec; ++t) {
      degree_count(a[t].x, n, h, deg);
      degree_count(a[t].y, n, h, deg);
      degree_count(a[t].z, n, h, deg);
      degree_count(a[t].w, n, h, deg);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < nnz - nv * 4) degree_count(col[nv * 4 + threadIdx.x], n, h, deg);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < kDegLds && q < n; q += kDegThreads)
    if (h[q]) atomicAdd(&deg[q], h[q]);
}

// Out-degrees of a large CSC without global atomics: the ids are grouped by id range
// (2^kDegRangeBits ids per range) with a radix sort of a copy over the range bits only,
// then every range is counted in an LDS histogram by the workgroups of its work items
// (kDegItem entries each); a range held by one item stores its counts, the items of a range
// split over several add theirs.  At RMAT-25 (5.3e8 ids) the global-atomic histogram is
// bound by the device's atomic rate (~1.3e10 per second); this path streams.
constexpr int kDegRangeBits = 14;
constexpr uint32_t kDegRange = 1u << kDegRangeBits;
constexpr uint64_t kDegSortMin = 1ull << 24;  // entries from which the range path is taken
constexpr uint32_t kDegItem = 1u << 20;
constexpr uint32_t kDegCountThreads = 512;

// start[r] = first position whose id's range is >= r, r in [0, nr]; every r written once
__global__ __launch_bounds__(256) void range_start_kernel(const uint32_t* __restrict__ ids, uint64_t nnz, uint32_t nr,
                                                          uint32_t* __restrict__ start) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p > nnz) return;
  const uint32_t a = p == 0 ? 0u : (ids[p - 1] >> kDegRangeBits) + 1u;
  uint32_t b = p == nnz ? nr : (ids[p] >> kDegRangeBits);
  if (b > nr) b = nr;  // ids are below n (checked by the callers); never write past the table
  for (uint32_t r = a; r <= b; ++r) start[r] = (uint32_t)p;
}

// one work item {range, first, end, split}: LDS counts of the range's ids in [first, end)
__global__ __launch_bounds__(kDegCountThreads) void range_count_kernel(const uint32_t* __restrict__ ids,
                                                                       const uint4* __restrict__ items, uint32_t n,
                                                                       uint32_t* __restrict__ deg) {
  __shared__ uint32_t h[kDegRange];
  const uint4 it = items[blockIdx.x];
  for (uint32_t q = threadIdx.x; q < kDegRange; q += kDegCountThreads) h[q] = 0u;
  __syncthreads();
  constexpr uint32_t kU = 4;
  for (uint32_t p = it.y + threadIdx.x; p < it.z; p += kDegCountThreads * kU) {
    uint32_t u[kU];
#pragma unroll
    for (uint32_t t = 0; t < kU; ++t) {
      const uint32_t q = p + t * kDegCountThreads;
      u[t] = q < it.z ? __builtin_nontemporal_load(ids + q) : 0u;
    }
#pragma unroll
    for (uint32_t t = 0; t < kU; ++t)
      if (p + t * kDegCountThreads < it.z) atomicAdd(&h[u[t] & (kDegRange - 1)], 1u);
  }
  __syncthreads();
  const uint32_t v0 = it.x << kDegRangeBits;
  const uint32_t cnt = n - v0 < kDegRange ? n - v0 : kDegRange;
  for (uint32_t q = threadIdx.x; q < cnt; q += kDegCountThreads) {
    if (!it.w) deg[v0 + q] = h[q];
    else if (h[q]) atomicAdd(&deg[v0 + q], h[q]);
  }
}

// *bad = 1 when some given out-degree is below the number of entries of its vertex
__global__ __launch_bounds__(256) void outdeg_check_kernel(uint32_t n, const uint32_t* __restrict__ given,
                                                           const uint32_t* __restrict__ counted,
                                                           uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n && given[i] < counted[i]) *bad = 1u;
}

// CSR keys from a CSC: (source << b) | target, payload = CSC position; one thread per CSC row.
__global__ __launch_bounds__(256) void transpose_keys_kernel(uint32_t n, const uint32_t* __restrict__ ptr,
                                                             const uint32_t* __restrict__ col, int b,
                                                             uint64_t* __restrict__ keys,
                                                             uint32_t* __restrict__ pos) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (uint32_t p = ptr[r]; p < ptr[r + 1]; ++p) {
    keys[p] = ((uint64_t)col[p] << b) | r;
    pos[p] = p;
  }
}

inline uint32_t grid_for(uint64_t n) { return (uint32_t)ceil_div(n ? n : 1, 256); }

uint32_t read_back_u32(hipStream_t s, const uint32_t* d) {
  uint32_t h = 0;
  NWG_HIP(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace

int key_bits_for(uint32_t n) { return bits_for(n); }

void sorted_keys_to_csx(hipStream_t s, const uint64_t* keys, uint64_t m, uint32_t n, int b, uint32_t* ptr,
                        uint32_t* ent) {
  if (m) {
    csx_kernel<<<grid_for(m), 256, 0, s>>>(keys, m, n, b, ptr, ent);
  } else {
    fill_u32_kernel<<<grid_for((uint64_t)n + 1), 256, 0, s>>>(ptr, (uint64_t)n + 1, 0u);
  }
  NWG_LAUNCH_CHECK();
}

// flags a CSC given in device memory whose row pointers are not monotone from 0 to nnz or
// whose entries are not below n (a plain store of 1: every flagging thread writes the same)
__global__ __launch_bounds__(256) void csc_check_kernel(uint32_t n, uint64_t nnz, const uint32_t* __restrict__ ptr,
                                                        const uint32_t* __restrict__ idx, uint32_t* __restrict__ bad) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  bool b = false;
  if (t < nnz) b = idx[t] >= n;
  if (t < n) b = b || ptr[t] > ptr[t + 1];
  if (t == 0) b = b || ptr[0] != 0u || ptr[n] != (uint32_t)nnz;
  if (b) *bad = 1u;
}

void validate_device_csc(hipStream_t s, uint32_t n, uint64_t nnz, const uint32_t* ptr, const uint32_t* idx,
                         uint32_t* flag) {
  DevBuf<uint32_t> own;
  if (!flag) {
    own.alloc(1);
    flag = own.p;
  }
  fill_u32_kernel<<<1, 256, 0, s>>>(flag, 1, 0u);
  csc_check_kernel<<<grid_for(std::max<uint64_t>(nnz, n)), 256, 0, s>>>(n, nnz, ptr, idx, flag);
  NWG_LAUNCH_CHECK();
  if (read_back_u32(s, flag)) throw Error(NWG_EINVAL, "device CSC: row pointers not monotone from 0 to nnz, or an entry >= n");
}

// The range path for a large CSC (ids below n < 2^32 assumed checked): a sorted copy, range
// starts read back, work items built on the host.
static void out_degrees_by_range(hipStream_t s, uint32_t n, uint64_t nnz, const uint32_t* col, uint32_t* deg) {
  const uint32_t nr = (uint32_t)ceil_div(n, kDegRange);
  DevBuf<uint32_t> ids(nnz), tmp(nnz), start((uint64_t)nr + 1);
  NWG_HIP(hipMemcpyAsync(ids.p, col, nnz * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  radix_sort_keys_u32(s, ids.p, tmp.p, nnz, kDegRangeBits, std::max(kDegRangeBits, bits_for(n)));
  range_start_kernel<<<grid_for(nnz + 1), 256, 0, s>>>(ids.p, nnz, nr, start.p);
  NWG_LAUNCH_CHECK();
  std::vector<uint32_t> hs((size_t)nr + 1);
  NWG_HIP(hipMemcpyAsync(hs.data(), start.p, hs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NWG_HIP(hipStreamSynchronize(s));
  std::vector<uint4> items;
  for (uint32_t r = 0; r < nr; ++r) {
    const uint32_t a = hs[r], b = hs[r + 1];
    const uint32_t split = b - a > kDegItem ? 1u : 0u;
    for (uint64_t q = a; q < b; q += kDegItem)
      items.push_back(make_uint4(r, (uint32_t)q, (uint32_t)std::min<uint64_t>(b, q + kDegItem), split));
  }
  if (items.empty()) return;
  DevBuf<uint4> ditems(items.size());
  NWG_HIP(hipMemcpyAsync(ditems.p, items.data(), items.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
  range_count_kernel<<<(uint32_t)items.size(), kDegCountThreads, 0, s>>>(ids.p, ditems.p, n, deg);
  NWG_LAUNCH_CHECK();
  NWG_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
}

void out_degrees_device(hipStream_t s, uint32_t n, uint64_t nnz, const uint32_t* col, uint32_t* deg,
                        bool may_allocate) {
  NWG_REQUIRE(reinterpret_cast<uintptr_t>(col) % 16 == 0, "column ids must be 16-byte aligned");
  fill_u32_kernel<<<grid_for(n), 256, 0, s>>>(deg, n, 0u);
  NWG_LAUNCH_CHECK();
  if (may_allocate && nnz >= kDegSortMin && nnz < (1ull << 32)) {
    out_degrees_by_range(s, n, nnz, col, deg);
    return;
  }
  if (nnz) {
    int dev = 0, cus = 0;
    NWG_HIP(hipGetDevice(&dev));
    NWG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // ranges of whole lane rounds (kDegThreads x kDegVec groups of 4 ids), about 2 per CU
    const uint64_t round = (uint64_t)kDegThreads * kDegVec;
    const uint64_t nv = nnz / 4;
    const uint64_t per = std::max<uint64_t>(1, ceil_div(ceil_div(nv, 2ull * std::max(cus, 1)), round)) * round;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, ceil_div(nv, per));
    degree_kernel<<<blocks, kDegThreads, 0, s>>>(col, nnz, n, per, deg);
    NWG_LAUNCH_CHECK();
  }
}

void check_outdeg_device(hipStream_t s, uint32_t n, uint64_t nnz, const uint32_t* col, const uint32_t* outdeg,
                         uint32_t* counted, uint32_t* flag) {
  DevBuf<uint32_t> own_counted, own_flag;
  if (!counted) {
    own_counted.alloc(n);
    counted = own_counted.p;
  }
  if (!flag) {
    own_flag.alloc(1);
    flag = own_flag.p;
  }
  out_degrees_device(s, n, nnz, col, counted, own_counted.p != nullptr);
  fill_u32_kernel<<<1, 256, 0, s>>>(flag, 1, 0u);
  outdeg_check_kernel<<<grid_for(n), 256, 0, s>>>(n, outdeg, counted, flag);
  NWG_LAUNCH_CHECK();
  if (read_back_u32(s, flag))
    throw Error(NWG_EINVAL, "outdeg: some vertex's out-degree is below its number of entries in the CSC");
}

// Sorted compressed structure from packed keys (consumes keys/tmp).
static void keys_to_compressed(hipStream_t s, uint32_t n, int b, DevBuf<uint64_t>& keys, DevBuf<uint64_t>& tmp,
                               uint64_t m, bool clean, DevBuf<uint32_t>& ptr, DevBuf<uint32_t>& ent,
                               uint64_t* nnz_out) {
  radix_sort_pairs(s, keys.p, tmp.p, nullptr, nullptr, m, 2 * b);
  if (clean && m) {
    DevBuf<uint32_t> flag(m), pos(m + 1);
    uniq_flag_kernel<<<grid_for(m), 256, 0, s>>>(keys.p, m, flag.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, flag.p, pos.p, m);
    const uint32_t mu = read_back_u32(s, pos.p + m);
    compact_keys_kernel<<<grid_for(m), 256, 0, s>>>(keys.p, m, flag.p, pos.p, tmp.p);
    NWG_LAUNCH_CHECK();
    std::swap(keys, tmp);
    m = mu;
    NWG_HIP(hipStreamSynchronize(s));
  }
  ptr.alloc((uint64_t)n + 1);
  ent.alloc(m ? m : 1);
  if (m) {
    csx_kernel<<<grid_for(m), 256, 0, s>>>(keys.p, m, n, b, ptr.p, ent.p);
  } else {
    fill_u32_kernel<<<grid_for((uint64_t)n + 1), 256, 0, s>>>(ptr.p, (uint64_t)n + 1, 0u);
  }
  NWG_LAUNCH_CHECK();
  *nnz_out = m;
}

void build_csc_device(hipStream_t s, uint32_t n, uint64_t ne, uint32_t* src, uint32_t* dst, bool clean,
                      nwg_graph* g) {
  NWG_REQUIRE(n >= 1, "graph needs at least one vertex");
  NWG_REQUIRE(ne < (1ull << 32), "edge count must be below 2^32 (uint32 indices)");
  const int b = bits_for(n);
  uint64_t m = 0;
  DevBuf<uint64_t> keys(ne ? ne : 1), tmp(ne ? ne : 1);
  if (ne) {
    DevBuf<uint32_t> flag(ne), pos(ne + 1), bad(1);
    fill_u32_kernel<<<1, 256, 0, s>>>(bad.p, 1, 0u);
    edge_flag_kernel<<<grid_for(ne), 256, 0, s>>>(src, dst, ne, n, clean ? 1 : 0, flag.p, bad.p);
    NWG_LAUNCH_CHECK();
    exclusive_scan_u32(s, flag.p, pos.p, ne);
    if (read_back_u32(s, bad.p)) throw Error(NWG_EINVAL, "edge list holds a vertex id >= n");
    m = read_back_u32(s, pos.p + ne);
    pack_kernel<<<grid_for(ne), 256, 0, s>>>(src, dst, ne, b, 1, flag.p, pos.p, keys.p);
    NWG_LAUNCH_CHECK();
    NWG_HIP(hipStreamSynchronize(s));
  }
  uint64_t nnz = 0;
  keys_to_compressed(s, n, b, keys, tmp, m, clean, g->csc_ptr, g->csc_idx, &nnz);
  g->n = n;
  g->nnz = nnz;
  g->outdeg.alloc(n);
  out_degrees_device(s, n, nnz, g->csc_idx.p, g->outdeg.p);
  NWG_HIP(hipStreamSynchronize(s));
}

void build_csr_device(hipStream_t s, nwg_graph* g) {
  if (g->has_csr) return;
  const uint32_t n = g->n;
  const uint64_t nnz = g->nnz;
  const int b = bits_for(n);
  DevBuf<uint64_t> keys(nnz ? nnz : 1), tmp(nnz ? nnz : 1);
  DevBuf<uint32_t> pos(nnz ? nnz : 1), pos_tmp(nnz ? nnz : 1);
  if (nnz) {
    transpose_keys_kernel<<<grid_for(n), 256, 0, s>>>(n, g->csc_ptr.p, g->csc_idx.p, b, keys.p, pos.p);
    NWG_LAUNCH_CHECK();
    radix_sort_pairs(s, keys.p, tmp.p, pos.p, pos_tmp.p, nnz, 2 * b);
  }
  g->csr_ptr.alloc((uint64_t)n + 1);
  g->csr_idx.alloc(nnz ? nnz : 1);
  if (nnz) {
    csx_kernel<<<grid_for(nnz), 256, 0, s>>>(keys.p, nnz, n, b, g->csr_ptr.p, g->csr_idx.p);
  } else {
    fill_u32_kernel<<<grid_for((uint64_t)n + 1), 256, 0, s>>>(g->csr_ptr.p, (uint64_t)n + 1, 0u);
  }
  NWG_LAUNCH_CHECK();
  g->csr_to_csc = std::move(pos);
  NWG_HIP(hipStreamSynchronize(s));
  g->has_csr = true;
}

}  // namespace nwg
