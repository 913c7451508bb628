/*
 * nwg_oracle.h -- CPU oracle for the PageRank path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C restatement of the reference algorithms, used as the
 * checker for the HIP library.  Only tests, the smoke entry and the CPU
 * baseline leg of the benchmark may load it; the product library never links
 * it and never falls back to it.
 *
 * Reference citations are into pnnl/NWGraph:
 *   page_rank (v11) ........ include/nwgraph/algorithms/page_rank.hpp:89-148
 *   page_rank_v10 (Jacobi) . include/nwgraph/experimental/algorithms/page_rank.hpp:377-432
 *   page_rank_v1 ........... include/nwgraph/experimental/algorithms/page_rank.hpp:82-119
 *   page_rank_v13 (delta) .. include/nwgraph/experimental/algorithms/page_rank.hpp:560-625
 *   fill_directed<idx> ..... include/nwgraph/build.hpp:209-232
 *   lexical_sort_by/uniq/remove_self_loops  include/nwgraph/build.hpp:57-70,392-410
 *   sort_to_be_indexed ..... include/nwgraph/containers/compressed.hpp:491-502
 *   build_degrees .......... bench/gapbs/common.hpp:110-121
 *   read_mm / mm_fill ...... include/nwgraph/io/mmio.hpp:86-103,171-217
 *   NW binary edge list .... include/nwgraph/edge_list.hpp:237-280, graph_base.hpp:72-87, containers/soa.hpp:332-375
 *
 * Pinning status of every function is listed in the header comment of
 * nwg_oracle.c (each is checked by tests/test_oracle.py against golden
 * vectors, closed forms or brute force).
 */
#ifndef NWG_ORACLE_H
#define NWG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---- */
void nwo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- synthetic inputs (no reference counterpart; the reference's RMAT lives in an
 *      un-vendored GAP submodule).  Definitions are shared bit-for-bit with the
 *      device generators and are documented in nwg_oracle.c. ---- */
/* raw RMAT edges [first, first+count) of a graph with 2^scale vertices. a,b,c in 1/10000. */
void nwo_rmat_edges(uint32_t scale, uint32_t a, uint32_t b, uint32_t c, uint64_t seed,
                    uint64_t first, uint64_t count, uint32_t* src, uint32_t* dst);
/* raw grid slots: returns number of slots (2*S + W*W); dropped slots hold (UINT32_MAX, UINT32_MAX). */
uint64_t nwo_grid_slots(uint32_t width);
void nwo_grid_edges(uint32_t width, uint32_t keep_per10000, uint32_t long_per10000, uint64_t seed,
                    uint32_t* src, uint32_t* dst);
/* PPR teleport sources: k vectors x s distinct ids in [0,n) */
void nwo_ppr_sources(uint32_t n, uint32_t k, uint32_t s, uint64_t seed, uint32_t* ids);
/* raw edge weight of edge (u,v) in [0.1, 1) */
float nwo_edge_weight(uint32_t u, uint32_t v, uint64_t seed);
/* standard normal vector (Box-Muller over Philox) */
void nwo_normal_vector(uint64_t n, uint64_t seed, double* out);

/* ---- edge-list cleaning and compressed build ---- */
/* lexical_sort_by<0> + uniq + remove_self_loops, in place; returns the new edge count */
uint64_t nwo_clean_edges(uint64_t ne, uint32_t* src, uint32_t* dst);
/* fill_directed<idx> followed by sort_to_be_indexed: idx=1 -> CSC (rows = targets,
 * entries = sources); idx=0 -> CSR.  indices has n+1 entries, col has ne. */
void nwo_fill_sorted(int idx, uint32_t n, uint64_t ne, const uint32_t* src, const uint32_t* dst,
                     uint32_t* indices, uint32_t* col);
/* build_degrees: out-degree of every vertex from a CSC (count of each source id) */
void nwo_out_degrees(uint32_t n, const uint32_t* indices, const uint32_t* col, uint32_t* deg);

/* ---- PageRank restatements (float scores, double error) ---- */
/* v11 as a deterministic block Gauss-Seidel: rows split in nblocks contiguous blocks;
 * inside a block the sweep is sequential and in place (page_rank.hpp:127-136); a block
 * sees other blocks' contributions from the start of the sweep.  nblocks=1 is the
 * reference run with one thread; nblocks=n is Jacobi. err_trace (nullable) gets one
 * double per executed sweep.  Returns sweeps executed. */
uint64_t nwo_pr_v11_blocks(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                           float* pr, float damping, float threshold, uint64_t max_iters,
                           uint32_t nblocks, double* err_trace);
/* v11 with nthreads OpenMP threads over static contiguous chunks, contributions updated
 * in place and read without synchronisation exactly as the reference's TBB body does
 * (schedule dependent).  Used only as the timed CPU baseline. sweep_seconds (nullable). */
uint64_t nwo_pr_v11_parallel(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                             float* pr, float damping, float threshold, uint64_t max_iters,
                             int nthreads, double* err_trace, double* sweep_seconds);
/* v10: Jacobi twin of v11 (contributions recomputed before each sweep).  The result does
 * not depend on nthreads: rows are independent and the error is summed per chunk in a
 * fixed order.  teleport (nullable): personalised base (1-d)*teleport[i] instead of (1-d)/N. */
uint64_t nwo_pr_v10(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                    float* pr, float damping, float threshold, uint64_t max_iters,
                    const float* teleport, int nthreads, double* err_trace, double* sweep_seconds);
/* v1: sequential Jacobi in double (the function the karate golden values pin). */
uint64_t nwo_pr_v1(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                   double* pr, double damping, double threshold, uint64_t max_iters);
/* v13: delta (residual) PageRank in float, un-normalised scores (sum over vertices ~ N).
 * Returns the iterations executed; pushed[k] (optional) gets the number of vertices
 * whose residual was above the threshold and had out-edges in iteration k. */
uint64_t nwo_pr_v13(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                    float* pr, float damping, float threshold, uint64_t max_iters, uint64_t* pushed);

/* ---- FP64 references ---- */
/* Jacobi in double on x = (1-d) v + d A x, A_iu = w(u->i) (w NULL => 1/outdeg(u));
 * v NULL => uniform 1/N.  Starts from 1/N, runs until the L1 change < tol or max_iters. */
uint64_t nwo_fp64_jacobi(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                         const double* w, const double* v, double damping, double tol,
                         uint64_t max_iters, double* x);
/* adjoint lambda = c + d A^T lambda by Jacobi in double (push over the CSC) */
uint64_t nwo_fp64_adjoint(uint32_t n, const uint32_t* indices, const uint32_t* col, const uint32_t* deg,
                          const double* w, const double* c, double damping, double tol,
                          uint64_t max_iters, double* lambda);
/* ---- readers ---- */
/* read_mm<directed>: returns 0 on success; *n = max(header rows, max id + 1);
 * edges are 0-based; a symmetric file mirrors off-diagonal entries.  Caller frees with free(). */
int nwo_read_mm(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst);
/* NW binary edge list (edge_list<directed>::deserialize) */
int nwo_read_nw(const char* path, uint32_t* n, uint64_t* ne, uint32_t** src, uint32_t** dst);
void nwo_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
