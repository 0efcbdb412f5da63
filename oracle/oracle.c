/*
 * CPU oracle for the GraphIt traversal path — TEST INFRASTRUCTURE ONLY.
 * (Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
 * never by the product library.)
 *
 * Serial C restatement of what the reference computes, plus the bench's
 * OpenMP CPU baseline (or_bfs_depth_omp). Citations are reference-relative
 * file:line. How each function is pinned:
 *
 *   or_splitmix64_mix / or_hash64  SplitMix64 published output vectors
 *                                  (seed 0 and seed 1234567) in tests/test_oracle.py.
 *   or_permute_id                  bijectivity over every scale 1..16 (brute force).
 *   or_rmat_edges                  quadrant frequencies vs (a,b,c,d) (chi-square),
 *                                  hand-derived draws for a scale-2 edge.
 *   or_grid_edges / uniform        closed-form edge counts, degree formulas.
 *   or_build_csr                   GAPBS's own builder (BuilderBase::MakeGraphFromEL +
 *                                  SquishGraph, builder.h:134-246) compiled from the
 *                                  reference into oracle/_ref, on the same edge lists;
 *                                  scipy.sparse canonical CSR when _ref is absent.
 *   or_transpose                   scipy.sparse transpose; double transpose = identity.
 *   or_bfs_depth                   scipy.sparse.csgraph.breadth_first_order and
 *                                  shortest_path(unweighted); fixtures in tests/golden.
 *   or_bfs_depth_omp               equal to or_bfs_depth on every test graph and
 *                                  thread count (the CPU baseline of the bench; OpenMP).
 *   or_sssp_dijkstra               scipy.sparse.csgraph.dijkstra; or_sssp_delta_stepping.
 *   or_cc_labels                   scipy connected_components (mapped to min id);
 *                                  or_cc_label_propagation (apps/cc.cpp restated).
 *   or_cc_omp, or_sssp_delta_stepping_omp  (CPU baselines; OpenMP) equal to or_cc_labels /
 *                                  or_sssp_dijkstra on every test graph at 1, 2 and 8 threads.
 *   or_pagerank                    closed forms (directed cycle: uniform 1/N; star);
 *                                  networkx.pagerank on dangling-free graphs at
 *                                  convergence; dense-matrix power iteration; the
 *                                  serial push loop restated in Python (bit-equal).
 *   or_bc                          brute-force all-shortest-path enumeration on tiny
 *                                  graphs; networkx single-source dependencies; equal
 *                                  (to FP64 rounding) at 1 and 8 OpenMP threads.
 *   or_mlp_*                       central finite differences of the loss.
 */
#include "oracle.h"

#include <limits.h>
#include <omp.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9e3779b97f4a7c15ULL

/* Threads of the OpenMP functions (or_pagerank, or_bc); 0 = OpenMP's default. */
void or_set_threads(int32_t n) {
    if (n > 0) omp_set_num_threads(n);
}

int32_t or_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------
 * Seeded generators. The reference generator (runtime/infra_gapbs/generator.h:
 * 74-107) draws R-MAT quadrants with std::mt19937 per block of 2^18 edges and
 * permutes ids with std::shuffle; both are libstdc++-specific. This project's
 * spec keeps the algorithm (bit-recursive quadrant choice with a=0.57 b=0.19
 * c=0.19, then a vertex-id permutation) but draws from a counter-based hash so
 * every edge is independent and the GPU generator reproduces it bit for bit.
 * ---------------------------------------------------------------------- */
uint64_t or_splitmix64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* Output #counter of a SplitMix64 generator whose state starts at
 * seed ^ mix(stream); stream 0 is the plain SplitMix64 sequence of seed. */
uint64_t or_hash64(uint64_t seed, uint64_t stream, uint64_t counter) {
    uint64_t state = seed ^ or_splitmix64_mix(stream);
    return or_splitmix64_mix(state + (counter + 1) * GOLDEN);
}

static uint32_t uniform_below(uint64_t h, uint32_t n) {
    return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

/* Bijection on [0, 2^scale): four rounds of odd-multiply-add and xorshift,
 * each invertible modulo 2^scale. Plays the role of PermuteIDs
 * (generator.h:45-55). */
uint32_t or_permute_id(uint32_t x, int scale, uint64_t seed) {
    if (scale <= 0) return 0;
    uint32_t mask = (scale >= 32) ? 0xffffffffu : ((1u << scale) - 1u);
    int sh = (scale + 1) / 2;
    for (int r = 0; r < 4; ++r) {
        uint32_t m = (uint32_t)or_hash64(seed, 2, 2 * r) | 1u;
        uint32_t c = (uint32_t)or_hash64(seed, 2, 2 * r + 1);
        x = (x * m + c) & mask;
        x ^= x >> sh;
    }
    return x;
}

#define RMAT_T_A 2448131358u   /* floor(0.57 * 2^32) */
#define RMAT_T_AB 3264175144u  /* floor(0.76 * 2^32) */
#define RMAT_T_ABC 4080218931u /* floor(0.95 * 2^32) */

/* generator.h:74-107: per level, draw r; r < a+b keeps src bit 0 (dst bit 1
 * when r > a); otherwise src bit 1 (dst bit 1 when r > a+b+c). */
void or_rmat_edges(int scale, int64_t num_edges, uint64_t seed, int permute,
                   int32_t *src, int32_t *dst) {
    int64_t words = (scale + 1) / 2;
    for (int64_t e = 0; e < num_edges; ++e) {
        uint32_t s = 0, d = 0;
        uint64_t h = 0;
        for (int l = 0; l < scale; ++l) {
            if ((l & 1) == 0) h = or_hash64(seed, 1, (uint64_t)e * words + l / 2);
            uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
            s <<= 1;
            d <<= 1;
            if (r < RMAT_T_AB) {
                if (r >= RMAT_T_A) d++;
            } else {
                s++;
                if (r >= RMAT_T_ABC) d++;
            }
        }
        if (permute) {
            s = or_permute_id(s, scale, seed);
            d = or_permute_id(d, scale, seed);
        }
        src[e] = (int32_t)s;
        dst[e] = (int32_t)d;
    }
}

/* width x width 4-neighbour grid (right and down edge of each vertex, in id
 * order) followed by num_random uniform long-range pairs. */
int64_t or_grid_edges(int width, int64_t num_random, uint64_t seed, int32_t *src,
                      int32_t *dst) {
    int64_t k = 0;
    uint32_t nv = (uint32_t)width * (uint32_t)width;
    for (int r = 0; r < width; ++r)
        for (int c = 0; c < width; ++c) {
            int32_t v = r * width + c;
            if (c + 1 < width) { src[k] = v; dst[k] = v + 1; ++k; }
            if (r + 1 < width) { src[k] = v; dst[k] = v + width; ++k; }
        }
    for (int64_t i = 0; i < num_random; ++i) {
        src[k] = (int32_t)uniform_below(or_hash64(seed, 4, 2 * i), nv);
        dst[k] = (int32_t)uniform_below(or_hash64(seed, 4, 2 * i + 1), nv);
        ++k;
    }
    return k;
}

/* Uniform random graph (the uniform=true branch of generator.h:57-72). */
void or_uniform_edges(int32_t num_vertices, int64_t num_edges, uint64_t seed,
                      int32_t *src, int32_t *dst) {
    for (int64_t i = 0; i < num_edges; ++i) {
        src[i] = (int32_t)uniform_below(or_hash64(seed, 5, 2 * i), (uint32_t)num_vertices);
        dst[i] = (int32_t)uniform_below(or_hash64(seed, 5, 2 * i + 1), (uint32_t)num_vertices);
    }
}

/* Many small uniform graphs packed with contiguous id ranges: graph i has
 * n_i vertices (piecewise-linear log-uniform over the octaves of
 * [vmin, vmax] from one 32-bit draw of stream 7) and n_i*avg_degree/2
 * uniform pairs inside its range (stream 8). Writes vertex_offsets[G+1];
 * returns the number of pairs written (src/dst must hold sum n_i*avg/2). */
int64_t or_uniform_batch_edges(int32_t num_graphs, int32_t vmin, int32_t vmax,
                               int32_t avg_degree, uint64_t seed, int32_t *vertex_offsets,
                               int32_t *src, int32_t *dst) {
    int32_t octaves = 0;
    while (((int64_t)vmin << (octaves + 1)) <= vmax) ++octaves;
    int64_t V = 0, e = 0;
    for (int32_t i = 0; i < num_graphs; ++i) {
        int64_t n = vmin;
        if (octaves > 0) {
            uint32_t a = uniform_below(or_hash64(seed, 7, (uint64_t)i), (uint32_t)octaves * 65536u);
            int64_t lo = (int64_t)vmin << (a / 65536u);
            n = lo + (lo * (int64_t)(a % 65536u)) / 65536;
            if (n > vmax) n = vmax;
        }
        vertex_offsets[i] = (int32_t)V;
        int64_t m = n * avg_degree / 2;
        for (int64_t k = 0; k < m; ++k, ++e) {
            if (src) {
                src[e] = (int32_t)(V + uniform_below(or_hash64(seed, 8, 2 * (uint64_t)e), (uint32_t)n));
                dst[e] = (int32_t)(V + uniform_below(or_hash64(seed, 8, 2 * (uint64_t)e + 1), (uint32_t)n));
            }
        }
        V += n;
    }
    vertex_offsets[num_graphs] = (int32_t)V;
    return e;
}

/* Weight of the undirected edge {u,v}, uniform in [wmin, wmax]; a function of
 * the endpoints only, so both directions and every duplicate agree
 * (the role of InsertWeights, generator.h:124-139). */
int32_t or_edge_weight(uint64_t seed, int32_t u, int32_t v, int32_t wmin, int32_t wmax) {
    uint32_t a = (uint32_t)(u < v ? u : v), b = (uint32_t)(u < v ? v : u);
    uint64_t h = or_hash64(seed, 3, ((uint64_t)a << 32) | b);
    return wmin + (int32_t)uniform_below(h, (uint32_t)(wmax - wmin + 1));
}

/* SourcePicker::PickNext (benchmark.h:56-64): uniform draws, skipping
 * degree-0 vertices. Returns the number written, or -1 if no vertex has an
 * edge. */
int32_t or_pick_sources(int32_t num_vertices, const int32_t *offsets, uint64_t seed,
                        int32_t count, int32_t *out) {
    if (offsets[num_vertices] == 0) return -1;
    uint64_t draw = 0;
    int32_t got = 0;
    while (got < count) {
        int32_t v = (int32_t)uniform_below(or_hash64(seed, 6, draw++), (uint32_t)num_vertices);
        if (offsets[v + 1] > offsets[v]) out[got++] = v;
    }
    return got;
}

/* ------------------------------------------------------------------------
 * CSR construction: MakeCSR (builder.h:188-203: count degrees, prefix sum,
 * scatter, both directions when symmetrizing) then SquishCSR (builder.h:
 * 134-163: per-vertex sort, unique, remove self-loop).
 * ---------------------------------------------------------------------- */
static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

int64_t or_build_csr(int32_t num_vertices, int64_t num_pairs, const int32_t *src,
                     const int32_t *dst, int symmetrize, int32_t *offsets,
                     int32_t *nbrs) {
    int64_t n = num_vertices;
    int64_t *pos = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < num_pairs; ++e) {
        pos[src[e] + 1]++;
        if (symmetrize) pos[dst[e] + 1]++;
    }
    for (int64_t v = 0; v < n; ++v) pos[v + 1] += pos[v];
    int64_t total = pos[n];
    int32_t *tmp = (int32_t *)malloc((size_t)(total ? total : 1) * sizeof(int32_t));
    int64_t *fill = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    memcpy(fill, pos, (size_t)(n + 1) * sizeof(int64_t));
    for (int64_t e = 0; e < num_pairs; ++e) {
        tmp[fill[src[e]]++] = dst[e];
        if (symmetrize) tmp[fill[dst[e]]++] = src[e];
    }
    free(fill);
    int64_t out = 0;
    offsets[0] = 0;
    for (int64_t v = 0; v < n; ++v) {
        int32_t *seg = tmp + pos[v];
        int64_t len = pos[v + 1] - pos[v];
        if (len > 1) qsort(seg, (size_t)len, sizeof(int32_t), cmp_i32);
        int32_t prev = -1;
        int have = 0;
        for (int64_t i = 0; i < len; ++i) {
            int32_t x = seg[i];
            if (have && x == prev) continue;
            have = 1;
            prev = x;
            if (x == (int32_t)v) continue;
            nbrs[out++] = x;
        }
        offsets[v + 1] = (int32_t)out;
    }
    free(tmp);
    free(pos);
    return out;
}

void or_fill_weights(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                     uint64_t seed, int32_t wmin, int32_t wmax, int32_t *weights) {
    for (int32_t v = 0; v < num_vertices; ++v)
        for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e)
            weights[e] = or_edge_weight(seed, v, nbrs[e], wmin, wmax);
}

/* builtin_transpose (runtime/infra_gpu/graph.h:165-234): count in-degrees,
 * prefix, then place edges in CSR order, so each row lists its sources in
 * increasing order. */
void or_transpose(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                  const int32_t *w, int32_t *t_offsets, int32_t *t_nbrs, int32_t *t_w) {
    int64_t n = num_vertices, m = offsets[num_vertices];
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < m; ++e) cnt[nbrs[e] + 1]++;
    for (int64_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
    for (int64_t v = 0; v <= n; ++v) t_offsets[v] = (int32_t)cnt[v];
    for (int64_t u = 0; u < n; ++u)
        for (int64_t e = offsets[u]; e < offsets[u + 1]; ++e) {
            int64_t p = cnt[nbrs[e]]++;
            t_nbrs[p] = (int32_t)u;
            if (w && t_w) t_w[p] = w[e];
        }
    free(cnt);
}

/* ------------------------------------------------------------------------
 * BFS (apps/bfs.cpp:23-33, 44-60). The reference records parent[dst]=src
 * through a first-wins CAS; the depth each vertex gets is the same for every
 * interleaving, so depth is the bit-exact output and parents are checked for
 * validity.
 * ---------------------------------------------------------------------- */
int32_t or_bfs_depth(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                     int32_t source, int32_t *depth) {
    for (int32_t v = 0; v < num_vertices; ++v) depth[v] = -1;
    if (source < 0 || source >= num_vertices) return 0;
    int32_t *queue = (int32_t *)malloc((size_t)num_vertices * sizeof(int32_t));
    int64_t head = 0, tail = 0;
    queue[tail++] = source;
    depth[source] = 0;
    int32_t levels = 1;
    while (head < tail) {
        int32_t u = queue[head++];
        for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
            int32_t v = nbrs[e];
            if (depth[v] == -1) {
                depth[v] = depth[u] + 1;
                if (depth[v] + 1 > levels) levels = depth[v] + 1;
                queue[tail++] = v;
            }
        }
    }
    free(queue);
    return levels;
}

/* The bench's CPU baseline: direction-optimizing BFS (Beamer, Asanovic and
 * Patterson, SC'12: top-down while the frontier's out-edges are at most
 * 1/alpha of the unexplored edges, bottom-up until the frontier shrinks
 * below V/beta; alpha 15, beta 18, the published defaults GAPBS uses) over
 * OpenMP threads. Top-down claims with compare-and-swap on depth and
 * per-thread queue blocks; bottom-up scans the unvisited vertices' lists for
 * a frontier member (a bitmap). Bottom-up needs in-neighbours, so a directed
 * graph (symmetric == 0) runs top-down only. nthreads <= 0: OpenMP's default.
 * Returns the number of levels; *threads_used gets the team size. */
int32_t or_bfs_depth_omp(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                         int32_t source, int32_t symmetric, int32_t nthreads, int32_t *depth,
                         int32_t *threads_used) {
    const int32_t V = num_vertices;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int32_t team = 1;
#pragma omp parallel
    {
#pragma omp single
        team = omp_get_num_threads();
    }
    if (threads_used) *threads_used = team;
#pragma omp parallel for schedule(static)
    for (int32_t v = 0; v < V; ++v) depth[v] = -1;
    if (source < 0 || source >= V) return 0;
    const int64_t words = ((int64_t)V + 63) / 64;
    int32_t *queue = (int32_t *)malloc((size_t)V * sizeof(int32_t));
    uint64_t *front = (uint64_t *)calloc((size_t)words, sizeof(uint64_t));
    uint64_t *next = (uint64_t *)calloc((size_t)words, sizeof(uint64_t));
    int64_t qn = 1;  /* frontier size (queue form during top-down) */
    queue[0] = source;
    depth[source] = 0;
    int64_t edges_left = (int64_t)offsets[V] - (offsets[source + 1] - offsets[source]);
    int64_t mf = offsets[source + 1] - offsets[source];
    int32_t level = 0;
    int bottom_up = 0;
    while (qn > 0) {
        if (!bottom_up && symmetric && mf > edges_left / 15) {
            /* switch: queue -> bitmap */
            memset(front, 0, (size_t)words * sizeof(uint64_t));
            for (int64_t i = 0; i < qn; ++i) front[queue[i] >> 6] |= 1ull << (queue[i] & 63);
            bottom_up = 1;
        } else if (bottom_up && qn < V / 18) {
            /* switch back: bitmap -> queue */
            int64_t n = 0;
            for (int64_t w = 0; w < words; ++w)
                for (uint64_t m = front[w]; m; m &= m - 1) queue[n++] = (int32_t)(w * 64 + __builtin_ctzll(m));
            qn = n;
            bottom_up = 0;
        }
        int64_t next_n = 0, next_mf = 0;
        if (bottom_up) {
            memset(next, 0, (size_t)words * sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : next_n, next_mf)
            for (int32_t v = 0; v < V; ++v) {
                if (depth[v] != -1) continue;
                for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e) {
                    const int32_t u = nbrs[e];
                    if ((front[u >> 6] >> (u & 63)) & 1ull) {
                        depth[v] = level + 1;
                        __atomic_fetch_or(&next[v >> 6], 1ull << (v & 63), __ATOMIC_RELAXED);
                        next_n += 1;
                        next_mf += offsets[v + 1] - offsets[v];
                        break;
                    }
                }
            }
            uint64_t *t = front;
            front = next;
            next = t;
        } else {
            int32_t *out = (int32_t *)malloc((size_t)(V > 0 ? V : 1) * sizeof(int32_t));
            int64_t out_n = 0;
#pragma omp parallel reduction(+ : next_mf)
            {
                int32_t local[1024];
                int32_t ln = 0;
#pragma omp for schedule(dynamic, 64) nowait
                for (int64_t i = 0; i < qn; ++i) {
                    const int32_t u = queue[i];
                    for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
                        const int32_t v = nbrs[e];
                        int32_t expect = -1;
                        if (__atomic_load_n(&depth[v], __ATOMIC_RELAXED) == -1 &&
                            __atomic_compare_exchange_n(&depth[v], &expect, level + 1, 0, __ATOMIC_RELAXED,
                                                        __ATOMIC_RELAXED)) {
                            next_mf += offsets[v + 1] - offsets[v];
                            local[ln++] = v;
                            if (ln == 1024) {
                                const int64_t at = __atomic_fetch_add(&out_n, ln, __ATOMIC_RELAXED);
                                memcpy(out + at, local, sizeof(int32_t) * (size_t)ln);
                                ln = 0;
                            }
                        }
                    }
                }
                if (ln) {
                    const int64_t at = __atomic_fetch_add(&out_n, ln, __ATOMIC_RELAXED);
                    memcpy(out + at, local, sizeof(int32_t) * (size_t)ln);
                }
            }
            memcpy(queue, out, sizeof(int32_t) * (size_t)out_n);
            free(out);
            next_n = out_n;
        }
        edges_left -= next_mf;
        mf = next_mf;
        qn = next_n;
        ++level;
    }
    free(queue);
    free(front);
    free(next);
    return level;
}

/* Multi-source BFS: distance to the nearest source (queue seeded with every
 * source at depth 0). On a disjoint union of graphs with one source per
 * graph it is every graph's own BFS. Returns the number of levels. */
int32_t or_bfs_multi_depth(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                           const int32_t *sources, int32_t num_sources, int32_t *depth) {
    for (int32_t v = 0; v < num_vertices; ++v) depth[v] = -1;
    int32_t *queue = (int32_t *)malloc((size_t)num_vertices * sizeof(int32_t) + 4);
    int64_t head = 0, tail = 0;
    int32_t levels = 0;
    for (int32_t i = 0; i < num_sources; ++i) {
        int32_t s = sources[i];
        if (s < 0 || s >= num_vertices || depth[s] == 0) continue;
        depth[s] = 0;
        queue[tail++] = s;
        levels = 1;
    }
    while (head < tail) {
        int32_t u = queue[head++];
        for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
            int32_t v = nbrs[e];
            if (depth[v] == -1) {
                depth[v] = depth[u] + 1;
                if (depth[v] + 1 > levels) levels = depth[v] + 1;
                queue[tail++] = v;
            }
        }
    }
    free(queue);
    return levels;
}

static int has_edge(const int32_t *offsets, const int32_t *nbrs, int32_t u, int32_t v) {
    int32_t lo = offsets[u], hi = offsets[u + 1];
    while (lo < hi) {
        int32_t mid = lo + (hi - lo) / 2;
        if (nbrs[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo < offsets[u + 1] && nbrs[lo] == v;
}

/* Number of vertices whose parent is not a valid BFS-tree parent. */
int32_t or_bfs_parent_valid(int32_t num_vertices, const int32_t *offsets,
                            const int32_t *nbrs, int32_t source, const int32_t *depth,
                            const int32_t *parent) {
    int32_t bad = 0;
    for (int32_t v = 0; v < num_vertices; ++v) {
        if (v == source) { bad += parent[v] != source; continue; }
        if (depth[v] < 0) { bad += parent[v] != -1; continue; }
        int32_t p = parent[v];
        if (p < 0 || p >= num_vertices || depth[p] != depth[v] - 1 ||
            !has_edge(offsets, nbrs, p, v))
            ++bad;
    }
    return bad;
}

/* Σ over vertices unvisited before the pull step that builds level+1 of the
 * neighbours examined in CSR order up to and including the first one at
 * depth == level (deg(v) when none is). This is the pull-step edge work. */
int64_t or_bfs_pull_examined(int32_t num_vertices, const int32_t *offsets,
                             const int32_t *nbrs, const int32_t *depth, int32_t level) {
    int64_t total = 0;
    for (int32_t v = 0; v < num_vertices; ++v) {
        if (depth[v] != -1 && depth[v] <= level) continue;
        int32_t k = 0;
        for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e) {
            ++k;
            if (depth[nbrs[e]] == level) break;
        }
        total += k;
    }
    return total;
}

/* ------------------------------------------------------------------------
 * SSSP: the reference relaxes SP[dst] = min(SP[dst], SP[src] + w)
 * (apps/sssp.cpp:24-31, graphit_types.h:426-430) inside delta windows
 * (gpu_priority_queue.h:115-153). Final distances are the shortest-path
 * distances whatever the window order; Dijkstra with 64-bit sums is the pin.
 * Unreached vertices keep INT_MAX (apps/sssp.cpp:29-31).
 * ---------------------------------------------------------------------- */
typedef struct { int64_t d; int32_t v; } heap_item;

static void heap_push(heap_item *h, int64_t *n, heap_item it) {
    int64_t i = (*n)++;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (h[p].d <= it.d) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = it;
}

static heap_item heap_pop(heap_item *h, int64_t *n) {
    heap_item top = h[0], last = h[--(*n)];
    int64_t i = 0;
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= *n) break;
        if (c + 1 < *n && h[c + 1].d < h[c].d) ++c;
        if (last.d <= h[c].d) break;
        h[i] = h[c];
        i = c;
    }
    h[i] = last;
    return top;
}

void or_sssp_dijkstra(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                      const int32_t *w, int32_t source, int32_t *dist) {
    int64_t *d = (int64_t *)malloc((size_t)num_vertices * sizeof(int64_t));
    for (int32_t v = 0; v < num_vertices; ++v) d[v] = INT64_MAX;
    int64_t cap = (int64_t)offsets[num_vertices] + 1, n = 0;
    heap_item *h = (heap_item *)malloc((size_t)cap * sizeof(heap_item));
    d[source] = 0;
    heap_item s0 = {0, source};
    heap_push(h, &n, s0);
    while (n > 0) {
        heap_item it = heap_pop(h, &n);
        if (it.d != d[it.v]) continue;
        for (int32_t e = offsets[it.v]; e < offsets[it.v + 1]; ++e) {
            int64_t nd = it.d + w[e];
            if (nd < d[nbrs[e]]) {
                d[nbrs[e]] = nd;
                heap_item x = {nd, nbrs[e]};
                heap_push(h, &n, x);
            }
        }
    }
    for (int32_t v = 0; v < num_vertices; ++v)
        dist[v] = d[v] == INT64_MAX ? INT_MAX : (int32_t)d[v];
    free(h);
    free(d);
}

/* Serial delta-stepping in the structure of ordered_processing.h:46-142
 * (bucket index = dist / delta, process the lowest non-empty bucket until it
 * stays empty). Returns the number of buckets processed. */
int64_t or_sssp_delta_stepping(int32_t num_vertices, const int32_t *offsets,
                               const int32_t *nbrs, const int32_t *w, int32_t source,
                               int32_t delta, int32_t *dist) {
    for (int32_t v = 0; v < num_vertices; ++v) dist[v] = INT_MAX;
    dist[source] = 0;
    int64_t nb = 1024;
    int64_t **bucket = (int64_t **)calloc((size_t)nb, sizeof(int64_t *));
    int64_t *bsz = (int64_t *)calloc((size_t)nb, sizeof(int64_t));
    int64_t *bcap = (int64_t *)calloc((size_t)nb, sizeof(int64_t));
#define PUSHB(b, v)                                                          \
    do {                                                                     \
        int64_t _b = (b);                                                    \
        while (_b >= nb) {                                                   \
            int64_t nn = nb * 2;                                             \
            bucket = (int64_t **)realloc(bucket, (size_t)nn * sizeof(int64_t *)); \
            bsz = (int64_t *)realloc(bsz, (size_t)nn * sizeof(int64_t));    \
            bcap = (int64_t *)realloc(bcap, (size_t)nn * sizeof(int64_t));  \
            for (int64_t q = nb; q < nn; ++q) { bucket[q] = 0; bsz[q] = 0; bcap[q] = 0; } \
            nb = nn;                                                         \
        }                                                                    \
        if (bsz[_b] == bcap[_b]) {                                           \
            bcap[_b] = bcap[_b] ? bcap[_b] * 2 : 16;                         \
            bucket[_b] = (int64_t *)realloc(bucket[_b], (size_t)bcap[_b] * sizeof(int64_t)); \
        }                                                                    \
        bucket[_b][bsz[_b]++] = (v);                                         \
    } while (0)
    PUSHB(0, source);
    int64_t processed = 0;
    for (int64_t b = 0; b < nb; ++b) {
        if (bsz[b] == 0) continue;
        ++processed;
        while (bsz[b] > 0) {
            int64_t n = bsz[b];
            int64_t *items = bucket[b];
            bucket[b] = 0; bsz[b] = 0; bcap[b] = 0;
            for (int64_t i = 0; i < n; ++i) {
                int32_t u = (int32_t)items[i];
                if ((int64_t)dist[u] / delta != b) continue; /* stale entry */
                for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
                    int64_t nd = (int64_t)dist[u] + w[e];
                    int32_t v = nbrs[e];
                    if (nd < dist[v]) {
                        dist[v] = (int32_t)nd;
                        PUSHB(nd / delta, v);
                    }
                }
            }
            free(items);
        }
    }
#undef PUSHB
    for (int64_t b = 0; b < nb; ++b) free(bucket[b]);
    free(bucket); free(bsz); free(bcap);
    return processed;
}

/* ------------------------------------------------------------------------
 * OpenMP delta-stepping (the CPU baseline of the SSSP benchmarks), after the
 * structure of the reference's ordered_processing.h:46-142: the current
 * bucket's vertices are relaxed in parallel, each thread appends improved
 * vertices to thread-local bins by bucket index, and the next round takes
 * the least non-empty bucket index over all threads' bins. writeMin is a
 * compare-and-swap loop. Distances equal or_sssp_dijkstra's for every
 * thread count. Returns the number of rounds.
 * ---------------------------------------------------------------------- */
typedef struct {
    int32_t *items;
    int64_t n, cap;
} i32vec;

static void i32vec_push(i32vec *v, int32_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->items = (int32_t *)realloc(v->items, (size_t)v->cap * sizeof(int32_t));
    }
    v->items[v->n++] = x;
}

int64_t or_sssp_delta_stepping_omp(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                                   const int32_t *w, int32_t source, int32_t delta, int32_t nthreads,
                                   int32_t *dist, int32_t *threads_used) {
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    if (threads_used) *threads_used = nt;
    for (int32_t v = 0; v < num_vertices; ++v) dist[v] = INT_MAX;
    dist[source] = 0;
    /* a pending distance lies below (current bucket + 1) * delta + the largest
     * weight, so nb rotating bins never hold two bucket indices at once */
    int64_t maxw = 1;
    const int64_t m = num_vertices ? offsets[num_vertices] : 0;
    for (int64_t e = 0; e < m; ++e) maxw = w[e] > maxw ? w[e] : maxw;
    const int64_t nb = maxw / delta + 2;
    i32vec **bins = (i32vec **)calloc((size_t)nt, sizeof(i32vec *));
    for (int t = 0; t < nt; ++t) bins[t] = (i32vec *)calloc((size_t)nb, sizeof(i32vec));
    int64_t cap = 1024, fsize = 1;
    int32_t *frontier = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
    frontier[0] = source;
    int64_t cur = 0, rounds = 0;
    while (fsize > 0) {
        ++rounds;
#pragma omp parallel num_threads(nt)
        {
            i32vec *mine = bins[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 64)
            for (int64_t i = 0; i < fsize; ++i) {
                const int32_t u = frontier[i];
                const int32_t du = __atomic_load_n(&dist[u], __ATOMIC_RELAXED);
                if ((int64_t)du / delta != cur) continue; /* stale: it was improved into an earlier round */
                for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
                    const int32_t v = nbrs[e];
                    const int64_t nd64 = (int64_t)du + w[e];
                    if (nd64 >= INT_MAX) continue;
                    const int32_t nd = (int32_t)nd64;
                    int32_t old = __atomic_load_n(&dist[v], __ATOMIC_RELAXED);
                    while (nd < old) {
                        if (__atomic_compare_exchange_n(&dist[v], &old, nd, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                            i32vec_push(&mine[(nd / delta) % nb], v);
                            break;
                        }
                    }
                }
            }
        }
        /* the next bucket: the current one when it was refilled, else the least non-empty */
        int64_t next = -1;
        for (int64_t b = cur; b < cur + nb && next < 0; ++b)
            for (int t = 0; t < nt; ++t)
                if (bins[t][b % nb].n) {
                    next = b;
                    break;
                }
        if (next < 0) break;
        cur = next;
        int64_t total = 0;
        for (int t = 0; t < nt; ++t) total += bins[t][cur % nb].n;
        if (total > cap) {
            cap = total;
            frontier = (int32_t *)realloc(frontier, sizeof(int32_t) * (size_t)cap);
        }
        fsize = 0;
        for (int t = 0; t < nt; ++t) {
            i32vec *bv = &bins[t][cur % nb];
            memcpy(frontier + fsize, bv->items, (size_t)bv->n * sizeof(int32_t));
            fsize += bv->n;
            bv->n = 0;
        }
    }
    for (int t = 0; t < nt; ++t) {
        for (int64_t b = 0; b < nb; ++b) free(bins[t][b].items);
        free(bins[t]);
    }
    free(bins);
    free(frontier);
    return rounds;
}

/* OpenMP label propagation with shortcutting (the CPU baseline of the CC
 * benchmark): every vertex takes the least label among itself and its
 * neighbours, then labels jump to their label's label, until nothing
 * changes; the fixpoint is the least vertex id of each component, as
 * or_cc_labels. Returns the rounds. */
int32_t or_cc_omp(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs, int32_t nthreads,
                  int32_t *labels, int32_t *threads_used) {
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    if (threads_used) *threads_used = nt;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int32_t v = 0; v < num_vertices; ++v) labels[v] = v;
    int32_t rounds = 0;
    int changed = 1;
    while (changed) {
        changed = 0;
        ++rounds;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024) reduction(| : changed)
        for (int32_t v = 0; v < num_vertices; ++v) {
            int32_t m = __atomic_load_n(&labels[v], __ATOMIC_RELAXED);
            const int32_t m0 = m;
            for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e) {
                const int32_t l = __atomic_load_n(&labels[nbrs[e]], __ATOMIC_RELAXED);
                if (l < m) m = l;
            }
            if (m < m0) {
                __atomic_store_n(&labels[v], m, __ATOMIC_RELAXED);
                changed = 1;
            }
        }
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int32_t v = 0; v < num_vertices; ++v) {
            int32_t l = __atomic_load_n(&labels[v], __ATOMIC_RELAXED);
            int32_t ll = __atomic_load_n(&labels[l], __ATOMIC_RELAXED);
            while (ll < l) {
                l = ll;
                ll = __atomic_load_n(&labels[l], __ATOMIC_RELAXED);
            }
            __atomic_store_n(&labels[v], l, __ATOMIC_RELAXED);
        }
    }
    return rounds;
}

/* ------------------------------------------------------------------------
 * CC: apps/cc.cpp hooks ID[ID[s]] min= ID[ID[d]] both ways, then pointer
 * jumps until stable; labels only decrease and a component's minimum vertex
 * stays its own root, so the fixpoint is "min vertex id of the component".
 * ---------------------------------------------------------------------- */
static int32_t uf_find(int32_t *p, int32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

void or_cc_labels(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                  int32_t *labels) {
    int32_t *p = (int32_t *)malloc((size_t)num_vertices * sizeof(int32_t));
    for (int32_t v = 0; v < num_vertices; ++v) p[v] = v;
    for (int32_t u = 0; u < num_vertices; ++u)
        for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
            int32_t a = uf_find(p, u), b = uf_find(p, nbrs[e]);
            if (a != b) { if (a < b) p[b] = a; else p[a] = b; }
        }
    /* union by smaller id keeps every root the minimum of its set */
    for (int32_t v = 0; v < num_vertices; ++v) labels[v] = uf_find(p, v);
    free(p);
}

/* apps/cc.cpp:23-49,60-72 executed serially. Returns hook sweeps. */
int32_t or_cc_label_propagation(int32_t num_vertices, const int32_t *offsets,
                                const int32_t *nbrs, int32_t *ID) {
    for (int32_t v = 0; v < num_vertices; ++v) ID[v] = v;
    int32_t sweeps = 0, upd0 = 1;
    while (upd0) {
        upd0 = 0;
        ++sweeps;
        for (int32_t s = 0; s < num_vertices; ++s)
            for (int32_t e = offsets[s]; e < offsets[s + 1]; ++e) {
                int32_t d = nbrs[e];
                int32_t src_id = ID[s], dst_id = ID[d];
                int32_t p_src = ID[src_id], p_dst = ID[dst_id];
                if (p_dst < ID[src_id]) ID[src_id] = p_dst;
                if (p_src < ID[dst_id]) ID[dst_id] = p_src;
                if (p_dst != ID[dst_id] || p_src != ID[src_id]) upd0 = 1;
            }
        int32_t upd1 = 1;
        while (upd1) {
            upd1 = 0;
            for (int32_t v = 0; v < num_vertices; ++v) {
                int32_t y = ID[v], x = ID[y];
                if (x != y) { ID[v] = x; upd1 = 1; }
            }
        }
    }
    return sweeps;
}

/* ------------------------------------------------------------------------
 * PageRank exactly as apps/pr.cpp:28-46,57-64 in FP64: init 1/N, then per
 * iteration contrib = old/outdeg, new[dst] += contrib[src] over every edge,
 * new = (1-d)/N + d*new, old = new. No dangling-mass redistribution. The
 * reference's damp is the float constant 0.85f (apps/pr.cpp:23).
 * ---------------------------------------------------------------------- */
void or_pagerank(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
                 int32_t iters, double damping, double *rank) {
    /* The serial loop adds contrib[u] into new[v] for u ascending. Pulling
     * over the in-neighbours of v in ascending order (the transpose built
     * below lists them so) performs the same additions in the same order,
     * so every vertex's sum is bit-identical to the serial push, and the
     * vertices can be summed in parallel (OpenMP). */
    int64_t n = num_vertices;
    int64_t m = n ? offsets[n] : 0;
    int32_t *toff = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    int32_t *tsrc = (int32_t *)malloc((size_t)(m > 0 ? m : 1) * sizeof(int32_t));
    int32_t *fill = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    for (int64_t e = 0; e < m; ++e) toff[nbrs[e] + 1]++;
    for (int64_t v = 0; v < n; ++v) toff[v + 1] += toff[v];
    for (int64_t u = 0; u < n; ++u) /* u ascending: each in-list sorted by source */
        for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
            int32_t v = nbrs[e];
            tsrc[toff[v] + fill[v]++] = (int32_t)u;
        }
    free(fill);
    double *contrib = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int64_t v = 0; v < n; ++v) rank[v] = 1.0 / (double)n;
    double beta = (1.0 - damping) / (double)n;
    for (int32_t it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < n; ++v) {
            int32_t deg = offsets[v + 1] - offsets[v];
            contrib[v] = deg ? rank[v] / deg : 0.0; /* never read when deg==0 */
        }
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t v = 0; v < n; ++v) {
            double acc = 0.0;
            for (int32_t e = toff[v]; e < toff[v + 1]; ++e) acc += contrib[tsrc[e]];
            rank[v] = beta + damping * acc;
        }
    }
    free(toff);
    free(tsrc);
    free(contrib);
}

/* ------------------------------------------------------------------------
 * Betweenness, apps/bc.cpp:28-63,92-117 per source: forward levels with
 * sigma (double) accumulated from the previous level; backward from the
 * deepest level, D(v) = 1/sigma(v) + sum of D over out-neighbours not yet
 * processed; final (D - 1/sigma) * sigma, 0 where sigma == 0. The source
 * keeps its own term. Scores are summed over sources.
 * ---------------------------------------------------------------------- */
static void bc_one_source(int64_t n, const int32_t *offsets, const int32_t *nbrs, int32_t s,
                          double *sigma, double *D, int32_t *depth, int32_t *order,
                          int32_t *lvl_start, unsigned char *done, double *bc) {
    for (int64_t v = 0; v < n; ++v) { sigma[v] = 0.0; D[v] = 0.0; depth[v] = -1; }
    sigma[s] = 1.0;
    depth[s] = 0;
    int64_t head = 0, tail = 0;
    int32_t nl = 0;
    order[tail++] = s;
    lvl_start[0] = 0;
    /* level-synchronous forward pass */
    while (head < tail) {
        int64_t lend = tail;
        lvl_start[++nl] = (int32_t)lend;
        for (int64_t i = head; i < lend; ++i) {
            int32_t u = order[i];
            for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e) {
                int32_t v = nbrs[e];
                if (depth[v] == -1) { depth[v] = depth[u] + 1; order[tail++] = v; }
                if (depth[v] == depth[u] + 1) sigma[v] += sigma[u];
            }
        }
        head = lend;
    }
    /* levels 0..nl-1 occupy order[lvl_start[l] .. lvl_start[l+1]) */
    memset(done, 0, (size_t)n);
    for (int32_t l = nl - 1; l >= 0; --l) {
        if (l < nl - 1) {
            /* push D from level l+1 to its not-yet-processed out-neighbours */
            for (int32_t i = lvl_start[l + 1]; i < lvl_start[l + 2]; ++i) {
                int32_t u = order[i];
                for (int32_t e = offsets[u]; e < offsets[u + 1]; ++e)
                    if (!done[nbrs[e]]) D[nbrs[e]] += D[u];
            }
        }
        for (int32_t i = lvl_start[l]; i < lvl_start[l + 1]; ++i) {
            int32_t u = order[i];
            done[u] = 1;
            D[u] += 1.0 / sigma[u];
        }
    }
    /* only the reached vertices have sigma != 0 */
    for (int64_t i = 0; i < tail; ++i) {
        int32_t v = order[i];
        bc[v] += (D[v] - 1.0 / sigma[v]) * sigma[v];
    }
}

/* Sources run in parallel (OpenMP, a static split so a thread count fixes
 * the order): each thread sums its sources into a private array in source
 * order, and the private arrays are added in thread order. With one thread
 * this is the serial loop. */
void or_bc(int32_t num_vertices, const int32_t *offsets, const int32_t *nbrs,
           const int32_t *sources, int32_t num_sources, double *bc) {
    int64_t n = num_vertices;
    for (int64_t v = 0; v < n; ++v) bc[v] = 0.0;
    int nt = omp_get_max_threads();
    if (nt > num_sources) nt = num_sources > 0 ? num_sources : 1;
    double **part = (double **)calloc((size_t)nt, sizeof(double *));
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        double *sigma = (double *)malloc((size_t)n * sizeof(double));
        double *D = (double *)malloc((size_t)n * sizeof(double));
        int32_t *depth = (int32_t *)malloc((size_t)n * sizeof(int32_t));
        int32_t *order = (int32_t *)malloc((size_t)n * sizeof(int32_t));
        int32_t *lvl_start = (int32_t *)malloc((size_t)(n + 2) * sizeof(int32_t));
        unsigned char *done = (unsigned char *)malloc((size_t)n);
        double *mine = t == 0 ? bc : (double *)calloc((size_t)n, sizeof(double));
        part[t] = mine;
#pragma omp for schedule(static)
        for (int32_t si = 0; si < num_sources; ++si)
            bc_one_source(n, offsets, nbrs, sources[si], sigma, D, depth, order, lvl_start, done, mine);
        free(sigma); free(D); free(depth); free(order); free(lvl_start); free(done);
    }
    for (int t = 1; t < nt; ++t) {
        for (int64_t v = 0; v < n; ++v) bc[v] += part[t][v];
        free(part[t]);
    }
    free(part);
}

/* ------------------------------------------------------------------------
 * Schedule-selector MLP in FP64: y = W3 relu(W2 relu(W1 x + b1) + b2) + b3,
 * weights row-major [in][out]; loss = masked mean squared error.
 * ---------------------------------------------------------------------- */
static void dense(int32_t batch, int32_t n_in, int32_t n_out, const double *x,
                  const double *w, const double *b, double *y, int relu) {
    for (int32_t i = 0; i < batch; ++i)
        for (int32_t j = 0; j < n_out; ++j) {
            double acc = b[j];
            for (int32_t k = 0; k < n_in; ++k) acc += x[i * n_in + k] * w[k * n_out + j];
            y[i * n_out + j] = (relu && acc < 0.0) ? 0.0 : acc;
        }
}

void or_mlp_forward(int32_t batch, int32_t n_in, int32_t n_hidden, int32_t n_out,
                    const double *x, const double *w1, const double *b1,
                    const double *w2, const double *b2, const double *w3,
                    const double *b3, double *h1, double *h2, double *y) {
    dense(batch, n_in, n_hidden, x, w1, b1, h1, 1);
    dense(batch, n_hidden, n_hidden, h1, w2, b2, h2, 1);
    dense(batch, n_hidden, n_out, h2, w3, b3, y, 0);
}

static void dense_backward(int32_t batch, int32_t n_in, int32_t n_out, const double *x,
                           const double *w, const double *gy, double *gw, double *gb,
                           double *gx) {
    for (int32_t k = 0; k < n_in; ++k)
        for (int32_t j = 0; j < n_out; ++j) {
            double acc = 0.0;
            for (int32_t i = 0; i < batch; ++i) acc += x[i * n_in + k] * gy[i * n_out + j];
            gw[k * n_out + j] = acc;
        }
    for (int32_t j = 0; j < n_out; ++j) {
        double acc = 0.0;
        for (int32_t i = 0; i < batch; ++i) acc += gy[i * n_out + j];
        gb[j] = acc;
    }
    if (gx)
        for (int32_t i = 0; i < batch; ++i)
            for (int32_t k = 0; k < n_in; ++k) {
                double acc = 0.0;
                for (int32_t j = 0; j < n_out; ++j) acc += gy[i * n_out + j] * w[k * n_out + j];
                gx[i * n_in + k] = acc;
            }
}

double or_mlp_loss_backward(int32_t batch, int32_t n_in, int32_t n_hidden, int32_t n_out,
                            const double *x, const double *target, const double *mask,
                            const double *w1, const double *b1, const double *w2,
                            const double *b2, const double *w3, const double *b3,
                            double *gw1, double *gb1, double *gw2, double *gb2,
                            double *gw3, double *gb3) {
    double *h1 = (double *)malloc((size_t)batch * n_hidden * sizeof(double));
    double *h2 = (double *)malloc((size_t)batch * n_hidden * sizeof(double));
    double *y = (double *)malloc((size_t)batch * n_out * sizeof(double));
    double *gy = (double *)malloc((size_t)batch * n_out * sizeof(double));
    double *g2 = (double *)malloc((size_t)batch * n_hidden * sizeof(double));
    double *g1 = (double *)malloc((size_t)batch * n_hidden * sizeof(double));
    or_mlp_forward(batch, n_in, n_hidden, n_out, x, w1, b1, w2, b2, w3, b3, h1, h2, y);
    double msum = 0.0, loss = 0.0;
    for (int64_t i = 0; i < (int64_t)batch * n_out; ++i) msum += mask[i];
    if (msum <= 0.0) msum = 1.0;
    for (int64_t i = 0; i < (int64_t)batch * n_out; ++i) {
        double d = y[i] - target[i];
        loss += mask[i] * d * d;
        gy[i] = 2.0 * mask[i] * d / msum;
    }
    loss /= msum;
    dense_backward(batch, n_hidden, n_out, h2, w3, gy, gw3, gb3, g2);
    for (int64_t i = 0; i < (int64_t)batch * n_hidden; ++i) if (h2[i] <= 0.0) g2[i] = 0.0;
    dense_backward(batch, n_hidden, n_hidden, h1, w2, g2, gw2, gb2, g1);
    for (int64_t i = 0; i < (int64_t)batch * n_hidden; ++i) if (h1[i] <= 0.0) g1[i] = 0.0;
    dense_backward(batch, n_in, n_hidden, x, w1, g1, gw1, gb1, 0);
    free(h1); free(h2); free(y); free(gy); free(g2); free(g1);
    return loss;
}
