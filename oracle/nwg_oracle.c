/*
 * nwg_oracle.c -- CPU oracle for the PageRank path (TEST INFRASTRUCTURE ONLY).
 *
 * Plain C restatement of the reference algorithms; the checker for the HIP
 * library, never part of it.  Built by oracle/Makefile into libnwg_oracle.so
 * and loaded by tests/, the smoke entry and the benchmark's cpu_baseline leg.
 *
 * What pins each function (tests/test_oracle.py):
 *   nwo_philox4x32_10  Random123 known-answer vectors (three published triples).
 *   nwo_rmat_edges     parity unpinned against any reference (the reference's RMAT is
 *                      an un-vendored submodule); pinned by invariants: quadrant
 *                      frequencies vs a/b/c, seed determinism, range, and device == host.
 *   nwo_grid_edges     closed form: lattice neighbours only, keep-rate, symmetry,
 *                      long-range count; device == host.
 *   nwo_clean_edges    reference KATs build_test.cpp:62-83,120-157 and numpy
 *                      lexsort/unique brute force.
 *   nwo_fill_sorted    reference KATs adjacency_test.cpp:26-45,240-257 and a
 *                      scipy.sparse CSC/CSR conversion of the same edges.
 *   nwo_out_degrees    numpy bincount of the sources.
 *   nwo_pr_v13         the karate golden values (scores / N, threshold 1e-9); with threshold
 *                      0 the dense Neumann partial sums sum_t (d P^T)^t (1-d) 1 (numpy
 *                      matrix powers); the directed cycle's closed form (residual
 *                      (1-d) d^t, absorbed while above the threshold)
 *   nwo_pr_v1          the reference's karate golden values (page_rank_test.cpp:71-75,
 *                      0.5 % relative), scipy dense solve.
 *   nwo_pr_v10         closed forms (cycle / regular digraph -> 1/N; isolated vertex ->
 *                      base), v1 in double, numpy dense power iteration.
 *   nwo_pr_v11_blocks  nblocks=N equals v10 bit-for-bit; nblocks=1 equals a
 *                      pure-Python sequential Gauss-Seidel on small graphs.
 *   nwo_fp64_jacobi / nwo_fp64_adjoint  numpy.linalg.solve of the dense system.
 *   nwo_read_mm        the reference's test/data/karate.mtx (78 symmetric entries ->
 *                      156 directed edges, matching page_rank_test.cpp:41-61).
 *   nwo_read_nw        byte layout written by a numpy writer that follows
 *                      edge_list.hpp:237-280; round trip.
 */
#include "nwg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Philox4x32-10                                                              */
/* ------------------------------------------------------------------------- */
void nwo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_u64(uint64_t a, uint32_t b, uint32_t tag, uint64_t seed, uint32_t out[4]) {
  uint32_t ctr[4] = {(uint32_t)a, (uint32_t)(a >> 32), b, tag};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  nwo_philox4x32_10(ctr, key, out);
}

#define TAG_RMAT 0x524D4154u /* 'RMAT' */
#define TAG_GRID 0x47524944u /* 'GRID' */
#define TAG_PPRS 0x50505253u /* 'PPRS' */
#define TAG_WGHT 0x57474854u /* 'WGHT' */
#define TAG_NORM 0x4E4F524Du /* 'NORM' */

static uint64_t per10000(uint32_t p) { return ((uint64_t)p << 32) / 10000u; }

/* ------------------------------------------------------------------------- */
/* RMAT: edge e picks one quadrant per level, most significant bit first.      */
/* Level l uses word l%4 of Philox(ctr = {e.lo, e.hi, l/4, 'RMAT'}, key = seed) */
/* quadrant A=(0,0) if r < tA, B=(0,1) if r < tAB, C=(1,0) if r < tABC, else D  */
/* with (src bit, dst bit); no noise, no vertex permutation.                   */
/* ------------------------------------------------------------------------- */
void nwo_rmat_edges(uint32_t scale, uint32_t a, uint32_t b, uint32_t c, uint64_t seed, uint64_t first,
                    uint64_t count, uint32_t* src, uint32_t* dst) {
  const uint64_t tA = per10000(a), tAB = per10000(a + b), tABC = per10000(a + b + c);
  for (uint64_t k = 0; k < count; ++k) {
    uint64_t e = first + k;
    uint32_t s = 0, d = 0, r[4] = {0, 0, 0, 0};
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 3u) == 0) philox_u64(e, l >> 2, TAG_RMAT, seed, r);
      uint64_t x = r[l & 3u];
      uint32_t sb = (x >= tAB) ? 1u : 0u;
      uint32_t db = (x < tA) ? 0u : (x < tAB) ? 1u : (x < tABC) ? 0u : 1u;
      s |= sb << (scale - 1 - l);
      d |= db << (scale - 1 - l);
    }
    src[k] = s;
    dst[k] = d;
  }
}

/* ------------------------------------------------------------------------- */
/* Grid: W x W lattice, id = r*W + c.  Lattice edge k (horizontal k < W(W-1):  */
/* (r,c)-(r,c+1) with k = r(W-1)+c; vertical k = W(W-1) + rW + c: (r,c)-(r+1,c))*/
/* is kept when Philox({k.lo,k.hi,0,'GRID'})[0] < keep and stored in slots     */
/* 2k (a->b) and 2k+1 (b->a).  Vertex v gets one long-range edge in slot 2S+v  */
/* when Philox({v,0,1,'GRID'})[0] < long; its target is                        */
/* (v + 1 + mulhi(word1, W*W-1)) mod W*W (never v).  Dropped slots hold         */
/* (UINT32_MAX, UINT32_MAX), removed later as self-loops.                      */
/* ------------------------------------------------------------------------- */
uint64_t nwo_grid_slots(uint32_t width) {
  uint64_t W = width;
  return 4 * W * (W - 1) + W * W;
}

void nwo_grid_edges(uint32_t width, uint32_t keep_per10000, uint32_t long_per10000, uint64_t seed, uint32_t* src,
                    uint32_t* dst) {
  const uint64_t W = width, N = W * W, H = W * (W - 1), S = 2 * H;
  const uint64_t tk = per10000(keep_per10000), tl = per10000(long_per10000);
  uint32_t r[4];
  for (uint64_t k = 0; k < S; ++k) {
    uint64_t a, b;
    if (k < H) {
      uint64_t row = k / (W - 1), col = k % (W - 1);
      a = row * W + col;
      b = a + 1;
    } else {
      uint64_t kk = k - H, row = kk / W, col = kk % W;
      a = row * W + col;
      b = a + W;
    }
    philox_u64(k, 0, TAG_GRID, seed, r);
    if (r[0] < tk) {
      src[2 * k] = (uint32_t)a; dst[2 * k] = (uint32_t)b;
      src[2 * k + 1] = (uint32_t)b; dst[2 * k + 1] = (uint32_t)a;
    } else {
      src[2 * k] = dst[2 * k] = src[2 * k + 1] = dst[2 * k + 1] = UINT32_MAX;
    }
  }
  for (uint64_t v = 0; v < N; ++v) {
    philox_u64(v, 1, TAG_GRID, seed, r);
    if (r[0] < tl) {
      uint64_t off = ((uint64_t)r[1] * (N - 1)) >> 32;
      src[2 * S + v] = (uint32_t)v;
      dst[2 * S + v] = (uint32_t)((v + 1 + off) % N);
    } else {
      src[2 * S + v] = dst[2 * S + v] = UINT32_MAX;
    }
  }
}

/* PPR sources: vector j draws ids mulhi(Philox({j, t, 2, 'PPRS'})[0], n) for t = 0,1,...
 * keeping the first s distinct ones. */
void nwo_ppr_sources(uint32_t n, uint32_t k, uint32_t s, uint64_t seed, uint32_t* ids) {
  uint32_t r[4];
  for (uint32_t j = 0; j < k; ++j) {
    uint32_t got = 0;
    for (uint64_t t = 0; got < s; ++t) {
      uint32_t ctr[4] = {j, (uint32_t)t, 2u, TAG_PPRS};
      uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
      nwo_philox4x32_10(ctr, key, r);
      uint32_t id = (uint32_t)(((uint64_t)r[0] * n) >> 32);
      int dup = 0;
      for (uint32_t q = 0; q < got; ++q) dup |= (ids[(uint64_t)j * s + q] == id);
      if (!dup) ids[(uint64_t)j * s + got++] = id;
    }
  }
}

/* w(u,v) = fma(0.9, t, 0.1) with t = (Philox({u, v, 3, 'WGHT'})[0] >> 8) * 2^-24 */
float nwo_edge_weight(uint32_t u, uint32_t v, uint64_t seed) {
  uint32_t r[4];
  uint32_t ctr[4] = {u, v, 3u, TAG_WGHT};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  nwo_philox4x32_10(ctr, key, r);
  float t = (float)(r[0] >> 8) * 0x1p-24f;
  return fmaf(0.9f, t, 0.1f);
}

void nwo_normal_vector(uint64_t n, uint64_t seed, double* out) {
  uint32_t r[4];
  for (uint64_t i = 0; i < n; i += 2) {
    philox_u64(i / 2, 4u, TAG_NORM, seed, r);
    double u1 = ((double)(r[0] >> 8) + 1.0) * 0x1p-24;
    double u2 = (double)(r[1] >> 8) * 0x1p-24;
    double rad = sqrt(-2.0 * log(u1));
    out[i] = rad * cos(2.0 * M_PI * u2);
    if (i + 1 < n) out[i + 1] = rad * sin(2.0 * M_PI * u2);
  }
}

/* ------------------------------------------------------------------------- */
/* sorting helpers: LSD radix sort of 64-bit keys, 16-bit digits              */
/* ------------------------------------------------------------------------- */
static void radix_sort_u64(uint64_t* keys, uint64_t n) {
  if (n < 2) return;
  uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
  uint64_t* cnt = (uint64_t*)malloc(65536 * sizeof(uint64_t));
  uint64_t *a = keys, *b = tmp;
  uint64_t all_or = 0, all_and = ~0ull;
  for (uint64_t i = 0; i < n; ++i) { all_or |= keys[i]; all_and &= keys[i]; }
  for (int shift = 0; shift < 64; shift += 16) {
    if ((((all_or ^ all_and) >> shift) & 0xFFFFull) == 0) continue; /* digit constant */
    memset(cnt, 0, 65536 * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cnt[(a[i] >> shift) & 0xFFFF]++;
    uint64_t sum = 0;
    for (int d = 0; d < 65536; ++d) { uint64_t t = cnt[d]; cnt[d] = sum; sum += t; }
    for (uint64_t i = 0; i < n; ++i) b[cnt[(a[i] >> shift) & 0xFFFF]++] = a[i];
    uint64_t* t = a; a = b; b = t;
  }
  if (a != keys) memcpy(keys, a, n * sizeof(uint64_t));
  free(tmp);
  free(cnt);
}

uint64_t nwo_clean_edges(uint64_t ne, uint32_t* src, uint32_t* dst) {
  /* lexical_sort_by<0> (build.hpp:57-70): order by (u, v) */
  uint64_t* k = (uint64_t*)malloc((ne ? ne : 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < ne; ++i) k[i] = ((uint64_t)src[i] << 32) | dst[i];
  radix_sort_u64(k, ne);
  /* uniq (build.hpp:392-402): drop adjacent equal (u, v) */
  uint64_t m = 0;
  for (uint64_t i = 0; i < ne; ++i)
    if (i == 0 || k[i] != k[i - 1]) k[m++] = k[i];
  /* remove_self_loops (build.hpp:404-410): stable removal of u == v */
  uint64_t out = 0;
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t u = (uint32_t)(k[i] >> 32), v = (uint32_t)k[i];
    if (u != v) { src[out] = u; dst[out] = v; ++out; }
  }
  free(k);
  return out;
}

void nwo_fill_sorted(int idx, uint32_t n, uint64_t ne, const uint32_t* src, const uint32_t* dst, uint32_t* indices,
                     uint32_t* col) {
  /* fill_directed<idx> (build.hpp:209-232): rows are column idx of the edge list,
   * entries the other column; sort_to_be_indexed (compressed.hpp:491-502) sorts
   * each row ascending, which makes the result unique. */
  const uint32_t* rows = idx ? dst : src;
  const uint32_t* ents = idx ? src : dst;
  uint64_t* k = (uint64_t*)malloc((ne ? ne : 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < ne; ++i) k[i] = ((uint64_t)rows[i] << 32) | ents[i];
  radix_sort_u64(k, ne);
  memset(indices, 0, ((uint64_t)n + 1) * sizeof(uint32_t));
  for (uint64_t i = 0; i < ne; ++i) indices[(k[i] >> 32) + 1]++;
  for (uint64_t r = 0; r < n; ++r) indices[r + 1] += indices[r];
  for (uint64_t i = 0; i < ne; ++i) col[i] = (uint32_t)k[i];
  free(k);
}

void nwo_out_degrees(uint32_t n, const uint32_t* indices, const uint32_t* col, uint32_t* deg) {
  /* build_degrees (common.hpp:110-121): count every source id of the CSC */
  memset(deg, 0, (uint64_t)n * sizeof(uint32_t));
  for (uint64_t p = 0; p < indices[n]; ++p) deg[col[p]]++;
}

/* ------------------------------------------------------------------------- */
/* PageRank restatements                                                     */
/* ------------------------------------------------------------------------- */
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Scalars of page_rank.hpp:92-94: init = (Real)(1.0/N); base = (Real)((1.0 - (double)d)/N).
 * The score update base + d*z (page_rank.hpp:133) is written as fmaf: the reference's
 * Release build (-Ofast -march=native, CMakeLists.txt:65) contracts it on FMA hardware. */
uint64_t nwo_pr_v11_blocks(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg, float* pr,
                           float damping, float threshold, uint64_t max_iters, uint32_t nblocks, double* err_trace) {
  const float init = (float)(1.0 / n);
  const float base = (float)((1.0 - (double)damping) / n);
  if (nblocks < 1) nblocks = 1;
  if (nblocks > n) nblocks = n;
  float* contrib = (float*)malloc((uint64_t)n * sizeof(float));
  float* snap = nblocks > 1 ? (float*)malloc((uint64_t)n * sizeof(float)) : NULL;
  for (uint32_t i = 0; i < n; ++i) pr[i] = init;                     /* :100-104 */
  for (uint32_t i = 0; i < n; ++i) contrib[i] = pr[i] / (float)deg[i]; /* :114-118 */
  const uint64_t bs = ((uint64_t)n + nblocks - 1) / nblocks;
  uint64_t iter = 0;
  for (; iter < max_iters;) {
    if (snap) memcpy(snap, contrib, (uint64_t)n * sizeof(float));
    double error = 0.0;
    for (uint64_t b0 = 0; b0 < n; b0 += bs) {
      uint64_t b1 = b0 + bs < n ? b0 + bs : n;
      double partial = 0.0;
      for (uint64_t i = b0; i < b1; ++i) {                          /* :127-136 */
        float z = 0.0f;
        for (uint64_t p = indices[i]; p < indices[i + 1]; ++p) {
          uint32_t u = col[p];
          z += (!snap || (u >= b0 && u < b1)) ? contrib[u] : snap[u];
        }
        float old = pr[i];
        pr[i] = fmaf(damping, z, base);
        partial += (double)fabsf(pr[i] - old);
        contrib[i] = pr[i] / (float)deg[i];
      }
      error += partial;                                              /* std::plus join */
    }
    if (err_trace) err_trace[iter] = error;
    ++iter;
    if (error < (double)threshold) break;                            /* :144-146 */
  }
  free(contrib);
  free(snap);
  return iter;
}

uint64_t nwo_pr_v11_parallel(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg, float* pr,
                             float damping, float threshold, uint64_t max_iters, int nthreads, double* err_trace,
                             double* sweep_seconds) {
  const float init = (float)(1.0 / n);
  const float base = (float)((1.0 - (double)damping) / n);
  float* contrib = (float*)malloc((uint64_t)n * sizeof(float));
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) pr[i] = init;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) contrib[i] = pr[i] / (float)deg[i];
  uint64_t iter = 0;
  for (; iter < max_iters;) {
    double t0 = now_s();
    double error = 0.0;
    /* static contiguous chunks, contrib written in place and read by other threads
     * without synchronisation, as the reference's parallel_reduce body does
     * (page_rank.hpp:124-139; benign on x86 for aligned 4-byte floats). */
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : error)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      float z = 0.0f;
      for (uint64_t p = indices[i]; p < indices[i + 1]; ++p) z += ((volatile float*)contrib)[col[p]];
      float old = pr[i];
      float x = fmaf(damping, z, base);
      pr[i] = x;
      error += (double)fabsf(x - old);
      ((volatile float*)contrib)[i] = x / (float)deg[i];
    }
    double t1 = now_s();
    if (err_trace) err_trace[iter] = error;
    if (sweep_seconds) sweep_seconds[iter] = t1 - t0;
    ++iter;
    if (error < (double)threshold) break;
  }
  free(contrib);
  return iter;
}

#define V10_CHUNKS 1024u

uint64_t nwo_pr_v10(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg, float* pr,
                    float damping, float threshold, uint64_t max_iters, const float* teleport, int nthreads,
                    double* err_trace, double* sweep_seconds) {
  const float init = (float)(1.0 / n);
  const float base = (float)((1.0 - (double)damping) / n);
  const double omd = 1.0 - (double)damping;
  float* contrib = (float*)malloc((uint64_t)n * sizeof(float));
  double partial[V10_CHUNKS];
  if (nthreads < 1) nthreads = 1;
  const uint64_t cs = ((uint64_t)n + V10_CHUNKS - 1) / V10_CHUNKS;
  for (uint32_t i = 0; i < n; ++i) pr[i] = init; /* experimental/page_rank.hpp:389-393 */
  uint64_t iter = 0;
  for (; iter < max_iters;) {
    double t0 = now_s();
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) contrib[i] = pr[i] / (float)deg[i]; /* :400-406 */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t c = 0; c < (int64_t)V10_CHUNKS; ++c) {                          /* :408-425 */
      uint64_t i0 = (uint64_t)c * cs, i1 = i0 + cs < n ? i0 + cs : n;
      double ps = 0.0;
      for (uint64_t i = i0; i < i1; ++i) {
        float z = 0.0f;
        for (uint64_t p = indices[i]; p < indices[i + 1]; ++p) z += contrib[col[p]];
        float b = teleport ? (float)(omd * (double)teleport[i]) : base;
        float old = pr[i];
        pr[i] = fmaf(damping, z, b);
        ps += (double)fabsf(pr[i] - old);
      }
      partial[c] = ps;
    }
    double error = 0.0;
    for (uint32_t c = 0; c < V10_CHUNKS; ++c) error += partial[c];
    double t1 = now_s();
    if (err_trace) err_trace[iter] = error;
    if (sweep_seconds) sweep_seconds[iter] = t1 - t0;
    ++iter;
    if (error < (double)threshold) break;
  }
  free(contrib);
  return iter;
}

uint64_t nwo_pr_v1(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg, double* pr,
                   double damping, double threshold, uint64_t max_iters) {
  /* experimental/page_rank.hpp:82-119, Real = double, sequential Jacobi */
  const double init = 1.0 / n, base = (1.0 - damping) / n;
  double* contrib = (double*)malloc((uint64_t)n * sizeof(double));
  for (uint32_t i = 0; i < n; ++i) pr[i] = init;
  uint64_t iter = 0;
  for (; iter < max_iters;) {
    double error = 0.0;
    for (uint32_t i = 0; i < n; ++i) contrib[i] = pr[i] / (double)(deg[i] + 0);
    for (uint32_t i = 0; i < n; ++i) {
      double z = 0.0;
      for (uint64_t p = indices[i]; p < indices[i + 1]; ++p) z += contrib[col[p]];
      double old = pr[i];
      pr[i] = base + damping * z;
      error += fabs(pr[i] - old);
    }
    ++iter;
    if (error < threshold) break;
  }
  free(contrib);
  return iter;
}

uint64_t nwo_pr_v13(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg, float* pr,
                    float damping, float threshold, uint64_t max_iters, uint64_t* pushed) {
  /* experimental/page_rank.hpp:560-625 (page_rank_v13, Real = float).  Every vertex starts
   * with residual 1 - d (:572-576: the double 1.0 - d stored as float).  An iteration first
   * absorbs each residual above the threshold into the score and turns it into a delta
   * r * d / deg (:580-594), then sets every row's residual to the sum of its in-neighbours'
   * positive deltas, or 0 (:597-612): residuals at or below the threshold that were not
   * absorbed are dropped.  The loop stops after the first iteration in which no vertex with
   * out-edges pushed (:616-620). */
  float* delta = (float*)malloc((uint64_t)(n ? n : 1) * sizeof(float));
  float* residual = (float*)malloc((uint64_t)(n ? n : 1) * sizeof(float));
  for (uint32_t i = 0; i < n; ++i) {
    pr[i] = 0.0f;
    delta[i] = 0.0f;
    residual[i] = (float)(1.0 - (double)damping);
  }
  uint64_t iter = 0;
  while (iter < max_iters) {
    uint64_t changed = 0;
    for (uint32_t src = 0; src < n; ++src) {
      delta[src] = 0.0f;
      if (residual[src] > threshold) {
        const float old = residual[src];
        pr[src] += old;
        residual[src] = 0.0f;
        if (deg[src] > 0) {
          delta[src] = old * damping / (float)deg[src];
          ++changed;
        }
      }
    }
    for (uint32_t i = 0; i < n; ++i) {
      float sum = 0.0f;
      for (uint64_t p = indices[i]; p < indices[i + 1]; ++p)
        if (delta[col[p]] > 0.0f) sum += delta[col[p]];
      residual[i] = sum > 0.0f ? sum : 0.0f;
    }
    if (pushed) pushed[iter] = changed;
    ++iter;
    if (!changed) break;
  }
  free(delta);
  free(residual);
  return iter;
}

/* ------------------------------------------------------------------------- */
/* FP64 fixed points                                                          */
/* ------------------------------------------------------------------------- */
static double edge_coef(const double* w, const uint32_t* deg, uint64_t p, uint32_t u) {
  return w ? w[p] : 1.0 / (double)deg[u];
}

uint64_t nwo_fp64_jacobi(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg, const double* w,
                         const double* v, double damping, double tol, uint64_t max_iters, double* x) {
  double* y = (double*)malloc((uint64_t)n * sizeof(double));
  for (uint32_t i = 0; i < n; ++i) x[i] = 1.0 / n;
  uint64_t iter = 0;
  while (iter < max_iters) {
    double err = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      double z = 0.0;
      for (uint64_t p = indices[i]; p < indices[i + 1]; ++p) z += edge_coef(w, deg, p, col[p]) * x[col[p]];
      y[i] = (1.0 - damping) * (v ? v[i] : 1.0 / n) + damping * z;
      err += fabs(y[i] - x[i]);
    }
    memcpy(x, y, (uint64_t)n * sizeof(double));
    ++iter;
    if (err < tol) break;
  }
  free(y);
  return iter;
}

uint64_t nwo_fp64_adjoint(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                          const double* w, const double* c, double damping, double tol, uint64_t max_iters,
                          double* lambda) {
  double* mu = (double*)malloc((uint64_t)n * sizeof(double));
  memcpy(lambda, c, (uint64_t)n * sizeof(double));
  uint64_t iter = 0;
  while (iter < max_iters) {
    memcpy(mu, c, (uint64_t)n * sizeof(double));
    for (uint32_t i = 0; i < n; ++i)
      for (uint64_t p = indices[i]; p < indices[i + 1]; ++p)
        mu[col[p]] += damping * edge_coef(w, deg, p, col[p]) * lambda[i];
    double err = 0.0;
    for (uint32_t i = 0; i < n; ++i) err += fabs(mu[i] - lambda[i]);
    memcpy(lambda, mu, (uint64_t)n * sizeof(double));
    ++iter;
    if (err < tol) break;
  }
  free(mu);
  return iter;
}

/* ------------------------------------------------------------------------- */
/* readers                                                                    */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint32_t *s, *d;
  uint64_t n, cap;
} edge_buf;

static int eb_push(edge_buf* b, uint32_t s, uint32_t d) {
  if (b->n == b->cap) {
    uint64_t nc = b->cap ? 2 * b->cap : 1024;
    uint32_t* ns = (uint32_t*)realloc(b->s, nc * sizeof(uint32_t));
    if (!ns) return -1;
    b->s = ns;
    uint32_t* nd = (uint32_t*)realloc(b->d, nc * sizeof(uint32_t));
    if (!nd) return -1;
    b->d = nd;
    b->cap = nc;
  }
  b->s[b->n] = s;
  b->d[b->n] = d;
  b->n++;
  return 0;
}

int nwo_read_mm(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst) {
  /* read_mm (mmio.hpp:171-208) + mm_fill (mmio.hpp:86-103) */
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[4096];
  char h[5][64];
  if (!fgets(line, sizeof line, f)) { fclose(f); return -2; }
  if (sscanf(line, "%63s %63s %63s %63s %63s", h[0], h[1], h[2], h[3], h[4]) != 5 ||
      strcmp(h[0], "%%MatrixMarket") != 0) { fclose(f); return -2; }
  int sym;
  if (strcmp(h[4], "symmetric") == 0) sym = 1;
  else if (strcmp(h[4], "general") == 0) sym = 0;
  else { fclose(f); return -3; }
  while (fgets(line, sizeof line, f))
    if (line[0] != '%') break;
  unsigned long long n0, n1, nnz;
  if (sscanf(line, "%llu %llu %llu", &n0, &n1, &nnz) != 3) { fclose(f); return -4; }
  edge_buf b = {0, 0, 0, 0};
  uint64_t N = n0;
  for (uint64_t i = 0; i < nnz; ++i) {
    unsigned long long d0, d1;
    if (!fgets(line, sizeof line, f) || sscanf(line, "%llu %llu", &d0, &d1) != 2) { fclose(f); free(b.s); free(b.d); return -5; }
    if (eb_push(&b, (uint32_t)(d0 - 1), (uint32_t)(d1 - 1))) { fclose(f); return -6; }
    uint64_t m = (d0 > d1 ? d0 : d1);
    if (m > N) N = m;
    if (sym && d0 != d1)
      if (eb_push(&b, (uint32_t)(d1 - 1), (uint32_t)(d0 - 1))) { fclose(f); return -6; }
  }
  fclose(f);
  *n = (uint32_t)N;
  *ne = b.n;
  *src = b.s;
  *dst = b.d;
  return 0;
}

int nwo_read_nw(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst) {
  /* edge_list::deserialize (edge_list.hpp:262-280): magic[27], directedness (int),
   * graph_base: size_t N, bool is_open (graph_base.hpp:72-87); struct_of_arrays:
   * size_t nedges, size_t narrays, then per array size_t n, size_t elsize, data
   * (soa.hpp:343-375). */
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  char magic[27];
  int dirn;
  uint64_t N, st, na, cnt, es;
  unsigned char open_flag;
  if (fread(magic, 1, 27, f) != 27 || strncmp(magic, "NW Graph index_edge_list", 27) != 0) { fclose(f); return -2; }
  if (fread(&dirn, sizeof dirn, 1, f) != 1 || fread(&N, 8, 1, f) != 1 || fread(&open_flag, 1, 1, f) != 1 ||
      fread(&st, 8, 1, f) != 1 || fread(&na, 8, 1, f) != 1 || na < 2) { fclose(f); return -3; }
  uint32_t* arr[2] = {(uint32_t*)malloc((st ? st : 1) * 4), (uint32_t*)malloc((st ? st : 1) * 4)};
  for (int a = 0; a < 2; ++a) {
    if (fread(&cnt, 8, 1, f) != 1 || fread(&es, 8, 1, f) != 1 || es != 4 || cnt != st ||
        fread(arr[a], 4, st, f) != st) { fclose(f); free(arr[0]); free(arr[1]); return -4; }
  }
  fclose(f);
  *n = (uint32_t)N;
  *ne = st;
  *src = arr[0];
  *dst = arr[1];
  return 0;
}

void nwo_free(void* p) { free(p); }
