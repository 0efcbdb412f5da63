/*
 * CPU oracle for the GraphIt traversal path — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or the timed CPU baseline). The
 * product path (easy_graphit_amd / libgraphit_amd.so) never links or calls it.
 *
 * Every function restates the reference's algorithm in plain serial C, citing
 * the reference file:line it follows (paths relative to the reference root).
 * How each one is pinned (golden vectors, closed forms, library routines,
 * brute force) is listed in the header comment of oracle.c.
 */
#ifndef GRAPHIT_ORACLE_H
#define GRAPHIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded generators (shared spec with the GPU generators) ---------- */
uint64_t or_splitmix64_mix(uint64_t z);
uint64_t or_hash64(uint64_t seed, uint64_t stream, uint64_t counter);
uint32_t or_permute_id(uint32_t x, int scale, uint64_t seed);
void or_rmat_edges(int scale, int64_t num_edges, uint64_t seed, int permute,
                   int32_t *src, int32_t *dst);
int64_t or_grid_edges(int width, int64_t num_random, uint64_t seed,
                      int32_t *src, int32_t *dst);
void or_uniform_edges(int32_t num_vertices, int64_t num_edges, uint64_t seed,
                      int32_t *src, int32_t *dst);
int64_t or_uniform_batch_edges(int32_t num_graphs, int32_t vmin, int32_t vmax,
                               int32_t avg_degree, uint64_t seed, int32_t *vertex_offsets,
                               int32_t *src, int32_t *dst);
int32_t or_edge_weight(uint64_t seed, int32_t u, int32_t v, int32_t wmin, int32_t wmax);

/* ---- CSR construction (GAPBS builder semantics) ----------------------- */
int64_t or_build_csr(int32_t num_vertices, int64_t num_pairs, const int32_t *src,
                     const int32_t *dst, int symmetrize, int32_t *offsets,
                     int32_t *nbrs);
void or_fill_weights(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                     uint64_t seed, int32_t wmin, int32_t wmax, int32_t *weights);
void or_transpose(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                  const int32_t *w, int32_t *t_offsets, int32_t *t_nbrs, int32_t *t_w);
int32_t or_pick_sources(int32_t num_vertices, const int32_t *offsets, uint64_t seed,
                        int32_t count, int32_t *out);

/* ---- algorithms ------------------------------------------------------- */
int32_t or_bfs_depth(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                     int32_t source, int32_t *depth);
int32_t or_bfs_depth_omp(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                         int32_t source, int32_t symmetric, int32_t nthreads, int32_t *depth,
                         int32_t *threads_used);
int32_t or_bfs_multi_depth(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                           const int32_t *sources, int32_t num_sources, int32_t *depth);
int32_t or_bfs_parent_valid(int32_t num_vertices, const int32_t *offsets,
                            const int32_t *nbrs, int32_t source, const int32_t *depth,
                            const int32_t *parent);
void or_sssp_dijkstra(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                      const int32_t *w, int32_t source, int32_t *dist);
int64_t or_sssp_delta_stepping(int32_t num_vertices, const int32_t *offsets,
                               const int32_t *nbrs, const int32_t *w, int32_t source,
                               int32_t delta, int32_t *dist);
int64_t or_sssp_delta_stepping_omp(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                                   const int32_t *w, int32_t source, int32_t delta, int32_t nthreads,
                                   int32_t *dist, int32_t *threads_used);
int32_t or_cc_omp(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs, int32_t nthreads,
                  int32_t *labels, int32_t *threads_used);
void or_cc_labels(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                  int32_t *labels);
int32_t or_cc_label_propagation(int32_t num_vertices, const int32_t *offsets,
                                const int32_t *nbrs, int32_t *labels);
void or_set_threads(int32_t n);
int32_t or_max_threads(void);
void or_pagerank(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                 int32_t iters, double damping, double *rank);
void or_bc(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
           const int32_t *sources, int32_t num_sources, double *bc);
int64_t or_bfs_pull_examined(int32_t num_vertices, const int32_t *offsets,
                             const int32_t *nbrs, const int32_t *depth, int32_t level);

/* ---- MLP schedule selector (FP64 forward / backward) ------------------ */
void or_mlp_forward(int32_t batch, int32_t n_in, int32_t n_hidden, int32_t n_out,
                    const double *x, const double *w1, const double *b1,
                    const double *w2, const double *b2, const double *w3,
                    const double *b3, double *h1, double *h2, double *y);
double or_mlp_loss_backward(int32_t batch, int32_t n_in, int32_t n_hidden, int32_t n_out,
                            const double *x, const double *target, const double *mask,
                            const double *w1, const double *b1, const double *w2,
                            const double *b2, const double *w3, const double *b3,
                            double *gw1, double *gb1, double *gw2, double *gb2,
                            double *gw3, double *gb3);

#ifdef __cplusplus
}
#endif
#endif
