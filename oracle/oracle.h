/*
 * CPU oracle for the hybrid traversal path — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check in __graft_entry__.py and bench.py's cpu_baseline leg may load it.
 * The product library (sep-graph_amd/, libsepgraph.so) never links it and
 * has no CPU fallback.
 *
 * Every routine is a plain, single-threaded restatement of the reference's
 * host-side verification code (cited per function) or, where the reference
 * defines no host routine for a quantity (FP64 PageRank with uniform dangling
 * redistribution, multi-source Brandes), of the textbook algorithm.
 *
 * Conventions shared with the product: vertex ids and CSR/CSC offsets are
 * uint32 (the reference's index_t), BFS levels are int32 with -1 for
 * unreached (the int32 view of the reference's UINT32_MAX sentinel), SSSP
 * distances are uint32 with UINT32_MAX for unreached.
 *
 * How each routine is pinned (tests/test_oracle.py):
 *   or_csc_from_csr      brute-force transpose + ascending-source order
 *   or_bfs               closed forms (path, star, grid, cycle) + all-pairs brute force
 *   or_sssp_fifo         equals or_sssp_dijkstra (independent restatement), closed forms
 *   or_pagerank          dense linear solve of the Google matrix, closed forms, sum = 1
 *   or_pagerank_residual linear solve of the dangling-dropped system
 *   or_bc                brute-force pair enumeration, closed forms (path, star, clique)
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CSR -> CSC transpose with sources ascending within each column.
 * Follows csr_tocsc, include/groute/graphs/csr_graph.h:330-397. w/w_out may be NULL. */
int or_csc_from_csr(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *w,
                    uint32_t *col_start, uint32_t *src, uint32_t *w_out);

/* FIFO BFS. Follows BFSHost, samples/hybrid_bfs/hybrid_bfs_host.cpp:33-60. */
int or_bfs(uint32_t n, const uint32_t *row, const uint32_t *col, uint32_t source, int32_t *levels);

/* FIFO label-correcting SSSP with uint32 arithmetic.
 * Follows SSSPHostNaive, samples/hybrid_sssp/hybrid_sssp_host.cpp:34-60. */
int or_sssp_fifo(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *w,
                 uint32_t source, uint32_t *dist);

/* Dijkstra with a binary heap (textbook; independent second restatement). */
int or_sssp_dijkstra(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *w,
                     uint32_t source, uint32_t *dist);

/* FP64 power-iteration PageRank: x0 = 1/n, x' = (1-d)/n + d*(P^T x + dangling(x)/n),
 * stop after the first iteration whose L1 change is below tol (or max_iter).
 * Not the reference's definition (see or_pagerank_residual); this is the
 * definition the MI355X build reports against. Returns iterations via *iters. */
int or_pagerank(uint32_t n, const uint32_t *row, const uint32_t *col, double damping, double tol,
                int max_iter, double *x, int *iters, double *last_residual);

/* Reference-semantics residual PageRank in FP64: value starts 0, residual is
 * seeded with (1-a)*a/outdeg pushes, vertices are processed FIFO until no
 * residual crosses `error`; dangling mass is dropped and ranks are not normalised.
 * Follows PageRankHost, samples/hybrid_pr/hybrid_pr_host.cpp:43-124. */
int or_pagerank_residual(uint32_t n, const uint32_t *row, const uint32_t *col, double alpha,
                         double error, double *rank);

/* Single-source Brandes dependency (sigma FP64, delta FP64), with the reference's
 * final *0.5. Follows BetweennessCentralityHost, samples/hybrid_bc/hybrid_bc_host.cpp:16-92.
 * delta may be NULL; sigma may be NULL. */
int or_bc_single(uint32_t n, const uint32_t *row, const uint32_t *col, uint32_t source,
                 double *half_delta, double *sigma);

/* Multi-source BC: bc[v] = 0.5 * sum over sources s of delta_s(v), FP64. */
int or_bc(uint32_t n, const uint32_t *row, const uint32_t *col, const uint32_t *sources, int k,
          double *bc);

#ifdef __cplusplus
}
#endif
#endif
