// Core runtime: errors, device info, allocation, fills, scans, and graph
// construction / transpose / blocking on the device.
//
// Replaces GraphT + load_graph + builtin_transpose + block_graph_edges
// (runtime/infra_gpu/graph.h:17-338) and the GAPBS CSR build
// (runtime/infra_gapbs/builder.h:134-246): where the reference builds on the
// host and copies, every step here runs on the GPU (radix sort of packed
// (row, col) keys, adjacent-duplicate and self-loop removal, degree count,
// scan).
#include <cstdarg>
#include <mutex>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "gi_bfs_fused.h"
#include "gi_hash.h"
#include "gi_internal.h"

namespace gi {

// ------------------------------------------------------------------ errors
static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

// ------------------------------------------------------------------ device
const DeviceInfo &device_info() {
    static DeviceInfo info;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
                info.num_cus = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
                info.max_threads_per_cu = p.maxThreadsPerMultiProcessor > 0
                                              ? p.maxThreadsPerMultiProcessor
                                              : 2048;
            }
        }
    });
    return info;
}

int persistent_grid(int block, int per_cu) {
    const DeviceInfo &d = device_info();
    int per = d.max_threads_per_cu / block;
    if (per < 1) per = 1;
    if (per > per_cu) per = per_cu;
    return d.num_cus * per;
}

int dmalloc(void **p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(GI_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return GI_OK;
}

void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// ------------------------------------------------------------------ fills
template <class T>
__global__ void fill_kernel(T *p, int64_t n, T v) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <class T>
static int fill_any(T *p, int64_t n, T v, hipStream_t s) {
    if (n <= 0) return GI_OK;
    int block = 256;
    int64_t want = ceil_div(n, block);
    int grid = (int)(want < persistent_grid(block) ? want : persistent_grid(block));
    fill_kernel<T><<<grid, block, 0, s>>>(p, n, v);
    GI_LAUNCH_CHECK();
    return GI_OK;
}

int fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s) { return fill_any(p, n, v, s); }
int fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s) { return fill_any(p, n, v, s); }
int fill_u8(uint8_t *p, int64_t n, uint8_t v, hipStream_t s) { return fill_any(p, n, v, s); }
int fill_f32(float *p, int64_t n, float v, hipStream_t s) { return fill_any(p, n, v, s); }
int fill_f64(double *p, int64_t n, double v, hipStream_t s) { return fill_any(p, n, v, s); }

// ------------------------------------------------------------------ scans
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
    int get(size_t need, void **out) {
        if (need > bytes) {
            dfree(ptr);
            ptr = nullptr;
            bytes = 0;
            GI_TRY(dmalloc(&ptr, need));
            bytes = need;
        }
        *out = ptr;
        return GI_OK;
    }
};
static thread_local Workspace g_scan_ws;

int exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s) {
    // out has n+1 entries; the scan runs over n+1 inputs with the last input
    // treated as zero via a transform-free trick: scan n, then add total.
    size_t bytes = 0;
    GI_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n,
                                   rocprim::plus<int32_t>(), s));
    void *ws = nullptr;
    GI_TRY(g_scan_ws.get(bytes, &ws));
    GI_HIP(rocprim::exclusive_scan(ws, bytes, in, out, 0, (size_t)n, rocprim::plus<int32_t>(),
                                   s));
    return GI_OK;
}

int read_count(const int32_t *d_count, int32_t *out, hipStream_t s) {
    GI_HIP(hipMemcpyAsync(out, d_count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GI_HIP(hipStreamSynchronize(s));
    return GI_OK;
}

// ------------------------------------------------------------------ graph build
static int bits_for(int64_t n) {
    int b = 0;
    while ((1ll << b) < n) ++b;
    return b < 1 ? 1 : b;
}

__global__ void rmat_keys_kernel(int scale, uint64_t seed, int64_t m, int permute,
                                 IdPermutation perm, int symmetrize, int b, uint64_t *keys) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
        uint32_t s, d;
        rmat_edge(scale, seed, (uint64_t)e, &s, &d);
        if (permute) {
            s = perm(s);
            d = perm(d);
        }
        keys[e] = ((uint64_t)s << b) | d;
        if (symmetrize) keys[m + e] = ((uint64_t)d << b) | s;
    }
}

__global__ void uniform_keys_kernel(uint32_t nv, uint64_t seed, int64_t m, int symmetrize,
                                    int b, uint64_t *keys) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
        uint32_t s = uniform_below(hash64(seed, 5, 2 * e), nv);
        uint32_t d = uniform_below(hash64(seed, 5, 2 * e + 1), nv);
        keys[e] = ((uint64_t)s << b) | d;
        if (symmetrize) keys[m + e] = ((uint64_t)d << b) | s;
    }
}

// Grid edges in the oracle's order: vertex id order, right edge then down
// edge; then num_random uniform pairs. Symmetrized.
__global__ void grid_keys_kernel(int width, int64_t num_grid, int64_t num_random,
                                 uint64_t seed, int b, uint64_t *keys) {
    int64_t m = num_grid + num_random;
    uint32_t nv = (uint32_t)width * (uint32_t)width;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
        uint32_t s, d;
        if (e < num_grid) {
            // rows 0..W-2 emit 2 edges per vertex except the last column (1);
            // the last row emits only right edges.
            int64_t per_row = 2 * (int64_t)width - 1;
            int64_t full_rows = (int64_t)width - 1;
            if (e < full_rows * per_row) {
                int64_t r = e / per_row, k = e % per_row;
                int64_t c = k / 2;
                if (k == per_row - 1) {  // last column: down edge only
                    c = width - 1;
                    s = (uint32_t)(r * width + c);
                    d = s + width;
                } else if ((k & 1) == 0) {
                    s = (uint32_t)(r * width + c);
                    d = s + 1;
                } else {
                    s = (uint32_t)(r * width + c);
                    d = s + width;
                }
            } else {
                int64_t k = e - full_rows * per_row;
                s = (uint32_t)((int64_t)(width - 1) * width + k);
                d = s + 1;
            }
        } else {
            int64_t i = e - num_grid;
            s = uniform_below(hash64(seed, 4, 2 * i), nv);
            d = uniform_below(hash64(seed, 4, 2 * i + 1), nv);
        }
        keys[e] = ((uint64_t)s << b) | d;
        keys[m + e] = ((uint64_t)d << b) | s;
    }
}

__global__ void pairs_keys_kernel(const int32_t *src, const int32_t *dst, int64_t m,
                                  int symmetrize, int b, uint64_t *keys) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
        uint64_t s = (uint32_t)src[e], d = (uint32_t)dst[e];
        keys[e] = (s << b) | d;
        if (symmetrize) keys[m + e] = (d << b) | s;
    }
}

// keep[i] = 1 when key i starts a new (row, col) run and is not a self-loop
// (SquishCSR: sort, unique, remove self, builder.h:147-149).
__global__ void squish_flags_kernel(const uint64_t *keys, int64_t n, int b, int32_t *keep) {
    uint64_t mask = (1ull << b) - 1;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t k = keys[i];
        bool first = (i == 0) || keys[i - 1] != k;
        keep[i] = (first && (k >> b) != (k & mask)) ? 1 : 0;
    }
}

__global__ void squish_scatter_kernel(const uint64_t *keys, const int32_t *keep,
                                      const int32_t *pos, int64_t n, int b, int32_t *nbr,
                                      int32_t *deg) {
    uint64_t mask = (1ull << b) - 1;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (!keep[i]) continue;
        uint64_t k = keys[i];
        nbr[pos[i]] = (int32_t)(k & mask);
        atomicAdd(&deg[(int32_t)(k >> b)], 1);
    }
}

static int grid_for(int64_t n, int block = 256) {
    int64_t want = ceil_div(n, block);
    int cap = persistent_grid(block);
    return (int)(want < 1 ? 1 : (want < cap ? want : cap));
}

// Sort packed keys, squish, and produce a CSR into g (keys are consumed).
static int build_from_keys(gi_graph *g, int32_t V, uint64_t *keys, int64_t n, int b,
                           hipStream_t s) {
    uint64_t *alt = nullptr;
    GI_TRY(dalloc(&alt, (size_t)n));
    size_t bytes = 0;
    rocprim::double_buffer<uint64_t> db(keys, alt);
    GI_HIP(rocprim::radix_sort_keys(nullptr, bytes, db, (size_t)n, 0, 2 * b, s));
    void *tmp = nullptr;
    int rc = dmalloc(&tmp, bytes);
    if (rc != GI_OK) { dfree(alt); return rc; }
    hipError_t he = rocprim::radix_sort_keys(tmp, bytes, db, (size_t)n, 0, 2 * b, s);
    dfree(tmp);
    if (he != hipSuccess) { dfree(alt); return fail(GI_ERR_HIP, "radix sort failed"); }
    uint64_t *sorted = db.current();
    int32_t *keep = nullptr, *pos = nullptr;
    rc = dalloc(&keep, (size_t)n + 1);
    if (rc == GI_OK) rc = dalloc(&pos, (size_t)n + 1);
    if (rc != GI_OK) { dfree(alt); dfree(keep); return rc; }
    squish_flags_kernel<<<grid_for(n), 256, 0, s>>>(sorted, n, b, keep);
    fill_i32(keep + n, 1, 0, s);
    rc = exclusive_scan_i32(keep, pos, n + 1, s);
    int32_t total = 0;
    if (rc == GI_OK) rc = read_count(pos + n, &total, s);
    if (rc != GI_OK) { dfree(alt); dfree(keep); dfree(pos); return rc; }
    int32_t *deg = nullptr;
    rc = dalloc(&g->d_nbr, (size_t)total + 1);
    if (rc == GI_OK) rc = dalloc(&g->d_off, (size_t)V + 2);
    if (rc == GI_OK) rc = dalloc(&deg, (size_t)V + 2);
    if (rc != GI_OK) { dfree(alt); dfree(keep); dfree(pos); dfree(deg); return rc; }
    fill_i32(deg, V + 1, 0, s);
    squish_scatter_kernel<<<grid_for(n), 256, 0, s>>>(sorted, keep, pos, n, b, g->d_nbr, deg);
    rc = exclusive_scan_i32(deg, g->d_off, V + 1, s);
    GI_HIP(hipStreamSynchronize(s));
    dfree(alt);
    dfree(keep);
    dfree(pos);
    dfree(deg);
    if (rc != GI_OK) return rc;
    g->V = V;
    g->E = total;
    // keys buffer belongs to the caller
    return GI_OK;
}

static int check_sizes(int64_t V, int64_t entries) {
    if (V <= 0 || V >= (1ll << 31) - 1) return fail(GI_ERR_RANGE, "vertex count %lld out of range", (long long)V);
    if (entries >= (1ll << 31) - 1)
        return fail(GI_ERR_RANGE, "%lld edge entries exceed int32 CSR offsets", (long long)entries);
    return GI_OK;
}

static int finish_build(gi_graph *g, int32_t V, uint64_t *keys, int64_t n, int b,
                        int directed, gi_graph **out) {
    hipStream_t s = 0;
    int rc = build_from_keys(g, V, keys, n, b, s);
    dfree(keys);
    if (rc != GI_OK) { gi_graph_destroy(g); return rc; }
    g->directed = directed;
    *out = g;
    return GI_OK;
}

// ------------------------------------------------------------------ COO, degree
__global__ void coo_kernel(const int32_t *off, int32_t V, int32_t E, int32_t *esrc) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        // last vertex whose offset <= e
        int32_t r = upper_bound_i32(off, 0, V + 1, (int32_t)e) - 1;
        esrc[e] = r;
    }
}

int ensure_coo(gi_graph *g, hipStream_t s) {
    if (g->d_esrc) return GI_OK;
    if (!g->d_off) return fail(GI_ERR_INVALID, "graph has no CSR offsets");
    GI_TRY(dalloc(&g->d_esrc, (size_t)g->E + 1));
    coo_kernel<<<grid_for(g->E), 256, 0, s>>>(g->d_off, g->V, (int32_t)g->E, g->d_esrc);
    GI_LAUNCH_CHECK();
    return GI_OK;
}

__global__ void max_degree_kernel(const int32_t *off, int32_t V, int32_t *out) {
    int32_t best = 0;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        int32_t d = off[v + 1] - off[v];
        best = d > best ? d : best;
    }
    for (int o = 32; o > 0; o >>= 1) {
        int32_t x = __shfl_xor(best, o);
        best = x > best ? x : best;
    }
    if (lane_id() == 0) atomicMax(out, best);
}

int ensure_max_degree(gi_graph *g, hipStream_t s) {
    if (g->max_degree >= 0) return GI_OK;
    int32_t *d = nullptr;
    GI_TRY(dalloc(&d, 1));
    fill_i32(d, 1, 0, s);
    max_degree_kernel<<<grid_for(g->V), 256, 0, s>>>(g->d_off, g->V, d);
    int rc = read_count(d, &g->max_degree, s);
    dfree(d);
    return rc;
}

int ensure_scratch(gi_graph *g, int64_t len) {
    if (g->scratch_len >= len) return GI_OK;
    dfree(g->d_scratch);
    g->d_scratch = nullptr;
    g->scratch_len = 0;
    GI_TRY(dalloc(&g->d_scratch, (size_t)len));
    g->scratch_len = len;
    return GI_OK;
}

__global__ void weights_kernel(const int32_t *esrc, const int32_t *nbr, int32_t E,
                               uint64_t seed, int32_t wmin, int32_t wmax, int32_t *w) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride)
        w[e] = edge_weight(seed, esrc[e], nbr[e], wmin, wmax);
}

// ------------------------------------------------------------------ transpose
__global__ void transpose_keys_kernel(const int32_t *esrc, const int32_t *nbr, int32_t E,
                                      uint64_t *keys, int32_t *vals) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        keys[e] = ((uint64_t)(uint32_t)nbr[e] << 32) | (uint32_t)esrc[e];
        vals[e] = (int32_t)e;
    }
}

__global__ void transpose_scatter_kernel(const uint64_t *keys, const int32_t *vals,
                                         const int32_t *w_in, int32_t E, int32_t *nbr,
                                         int32_t *w_out, int32_t *deg) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride) {
        uint64_t k = keys[i];
        nbr[i] = (int32_t)(uint32_t)k;
        if (w_out) w_out[i] = w_in[vals[i]];
        atomicAdd(&deg[(int32_t)(k >> 32)], 1);
    }
}

// ------------------------------------------------------------------ blocking
__global__ void block_keys_kernel(const int32_t *nbr, int32_t E, int32_t block_size,
                                  int32_t *keys, int32_t *vals) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        keys[e] = nbr[e] / block_size;
        vals[e] = (int32_t)e;
    }
}

__global__ void block_gather_kernel(const int32_t *perm, const int32_t *esrc_in,
                                    const int32_t *nbr_in, const int32_t *w_in, int32_t E,
                                    int32_t *esrc, int32_t *nbr, int32_t *w) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride) {
        int32_t e = perm[i];
        esrc[i] = esrc_in[e];
        nbr[i] = nbr_in[e];
        if (w) w[i] = w_in[e];
    }
}

__global__ void bucket_end_kernel(const int32_t *sorted_keys, int32_t E, int32_t nb,
                                  int32_t *bucket_end) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += stride)
        bucket_end[b] = upper_bound_i32(sorted_keys, 0, E, (int32_t)b);
}

}  // namespace gi

using namespace gi;

// ====================================================================== C ABI
extern "C" {

const char *gi_last_error(void) { return gi::g_error.c_str(); }
int32_t gi_version(void) { return 100; }

void gi_schedule_init(gi_schedule *s) {
    if (!s) return;
    s->direction = GI_PUSH;
    s->pull_frontier_rep = GI_PULL_BOOLMAP;
    s->frontier_creation = GI_FC_SPARSE;
    s->deduplication = GI_DEDUP_DISABLED;
    s->load_balancing = GI_LB_VERTEX_BASED;
    s->edge_blocking = GI_UNBLOCKED;
    s->block_size = -1;
    s->kernel_fusion = GI_FUSION_DISABLED;
    s->max_cta = 0;  // auto: fill every CU (the reference's 60 suits a V100)
    s->cta_size = 512;
}

int gi_graph_from_csr(int32_t V, int64_t E, const int32_t *offsets, const int32_t *nbrs,
                      const int32_t *weights, int32_t directed, gi_graph **out) {
    if (!out || !offsets || (E > 0 && !nbrs)) return fail(GI_ERR_INVALID, "null argument");
    *out = nullptr;
    GI_TRY(check_sizes(V, E));
    if (offsets[0] != 0 || offsets[V] != E)
        return fail(GI_ERR_INVALID, "offsets[0] must be 0 and offsets[V] must equal E");
    for (int64_t v = 0; v < V; ++v)
        if (offsets[v + 1] < offsets[v]) return fail(GI_ERR_INVALID, "offsets not monotone at %lld", (long long)v);
    for (int64_t e = 0; e < E; ++e)
        if (nbrs[e] < 0 || nbrs[e] >= V)
            return fail(GI_ERR_RANGE, "neighbour id %d at entry %lld out of range", nbrs[e], (long long)e);
    gi_graph *g = new gi_graph();
    g->V = V;
    g->E = E;
    g->directed = directed;
    int rc = dalloc(&g->d_off, (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&g->d_nbr, (size_t)E + 1);
    if (rc == GI_OK && weights) rc = dalloc(&g->d_w, (size_t)E + 1);
    if (rc == GI_OK && hipMemcpy(g->d_off, offsets, sizeof(int32_t) * (V + 1), hipMemcpyHostToDevice) != hipSuccess) rc = fail(GI_ERR_HIP, "copy offsets");
    if (rc == GI_OK && E && hipMemcpy(g->d_nbr, nbrs, sizeof(int32_t) * E, hipMemcpyHostToDevice) != hipSuccess) rc = fail(GI_ERR_HIP, "copy nbrs");
    if (rc == GI_OK && weights && E && hipMemcpy(g->d_w, weights, sizeof(int32_t) * E, hipMemcpyHostToDevice) != hipSuccess) rc = fail(GI_ERR_HIP, "copy weights");
    if (rc != GI_OK) { gi_graph_destroy(g); return rc; }
    *out = g;
    return GI_OK;
}

int gi_graph_from_edges(int32_t V, int64_t m, const int32_t *src, const int32_t *dst,
                        int32_t symmetrize, gi_graph **out) {
    if (!out || (m > 0 && (!src || !dst))) return fail(GI_ERR_INVALID, "null argument");
    *out = nullptr;
    int64_t n = symmetrize ? 2 * m : m;
    GI_TRY(check_sizes(V, n));
    for (int64_t e = 0; e < m; ++e)
        if (src[e] < 0 || src[e] >= V || dst[e] < 0 || dst[e] >= V)
            return fail(GI_ERR_RANGE, "edge %lld has an endpoint out of range", (long long)e);
    int b = bits_for(V);
    int32_t *ds = nullptr, *dd = nullptr;
    uint64_t *keys = nullptr;
    GI_TRY(dalloc(&keys, (size_t)n + 1));
    int rc = dalloc(&ds, (size_t)m + 1);
    if (rc == GI_OK) rc = dalloc(&dd, (size_t)m + 1);
    if (rc == GI_OK && m && hipMemcpy(ds, src, sizeof(int32_t) * m, hipMemcpyHostToDevice) != hipSuccess) rc = fail(GI_ERR_HIP, "copy src");
    if (rc == GI_OK && m && hipMemcpy(dd, dst, sizeof(int32_t) * m, hipMemcpyHostToDevice) != hipSuccess) rc = fail(GI_ERR_HIP, "copy dst");
    if (rc == GI_OK && m) {
        pairs_keys_kernel<<<grid_for(m), 256>>>(ds, dd, m, symmetrize, b, keys);
        if (hipGetLastError() != hipSuccess) rc = fail(GI_ERR_HIP, "pairs kernel");
    }
    dfree(ds);
    dfree(dd);
    if (rc != GI_OK) { dfree(keys); return rc; }
    return finish_build(new gi_graph(), V, keys, n, b, !symmetrize, out);
}

int gi_graph_rmat(int32_t scale, int32_t edge_factor, uint64_t seed, int32_t symmetrize,
                  gi_graph **out) {
    if (!out) return fail(GI_ERR_INVALID, "null argument");
    *out = nullptr;
    if (scale < 1 || scale > 30 || edge_factor < 1) return fail(GI_ERR_RANGE, "bad scale/edge factor");
    int64_t V = 1ll << scale;
    int64_t m = V * edge_factor;
    int64_t n = symmetrize ? 2 * m : m;
    GI_TRY(check_sizes(V, n));
    int b = bits_for(V);
    uint64_t *keys = nullptr;
    GI_TRY(dalloc(&keys, (size_t)n + 1));
    IdPermutation perm;
    perm.init(scale, seed);
    rmat_keys_kernel<<<grid_for(m), 256>>>(scale, seed, m, 1, perm, symmetrize, b, keys);
    if (hipGetLastError() != hipSuccess) { dfree(keys); return fail(GI_ERR_HIP, "rmat kernel"); }
    return finish_build(new gi_graph(), (int32_t)V, keys, n, b, !symmetrize, out);
}

int gi_graph_uniform(int32_t V, int64_t m, uint64_t seed, int32_t symmetrize, gi_graph **out) {
    if (!out) return fail(GI_ERR_INVALID, "null argument");
    *out = nullptr;
    int64_t n = symmetrize ? 2 * m : m;
    GI_TRY(check_sizes(V, n));
    int b = bits_for(V);
    uint64_t *keys = nullptr;
    GI_TRY(dalloc(&keys, (size_t)n + 1));
    if (m) uniform_keys_kernel<<<grid_for(m), 256>>>((uint32_t)V, seed, m, symmetrize, b, keys);
    if (hipGetLastError() != hipSuccess) { dfree(keys); return fail(GI_ERR_HIP, "uniform kernel"); }
    return finish_build(new gi_graph(), V, keys, n, b, !symmetrize, out);
}

int gi_graph_grid(int32_t width, int64_t num_random, uint64_t seed, gi_graph **out) {
    if (!out) return fail(GI_ERR_INVALID, "null argument");
    *out = nullptr;
    if (width < 2 || num_random < 0) return fail(GI_ERR_RANGE, "bad grid size");
    int64_t V = (int64_t)width * width;
    int64_t ng = 2ll * width * (width - 1);
    int64_t m = ng + num_random;
    GI_TRY(check_sizes(V, 2 * m));
    int b = bits_for(V);
    uint64_t *keys = nullptr;
    GI_TRY(dalloc(&keys, (size_t)(2 * m) + 1));
    grid_keys_kernel<<<grid_for(m), 256>>>(width, ng, num_random, seed, b, keys);
    if (hipGetLastError() != hipSuccess) { dfree(keys); return fail(GI_ERR_HIP, "grid kernel"); }
    return finish_build(new gi_graph(), (int32_t)V, keys, 2 * m, b, 0, out);
}

int gi_graph_set_weights(gi_graph *g, uint64_t seed, int32_t wmin, int32_t wmax) {
    if (!g || wmax < wmin) return fail(GI_ERR_INVALID, "bad arguments");
    // a transposed or blocked view is owned by its base graph, whose weights
    // it copies: new weights go on the base graph, which rebuilds its views
    if (g->borrowed || !g->d_off) return fail(GI_ERR_UNSUPPORTED, "set the weights of a base graph");
    hipStream_t s = 0;
    GI_TRY(ensure_coo(g, s));
    if (!g->d_w) GI_TRY(dalloc(&g->d_w, (size_t)g->E + 1));
    if (g->E) weights_kernel<<<grid_for(g->E), 256, 0, s>>>(g->d_esrc, g->d_nbr, (int32_t)g->E, seed, wmin, wmax, g->d_w);
    GI_LAUNCH_CHECK();
    GI_HIP(hipStreamSynchronize(s));
    // cached derived graphs carry old weights, and the fused engines' state
    // points into them: drop both
    release_fused(g);
    release_pr(g);
    release_cc(g);
    release_sssp(g);
    if (g->transposed && g->transposed != g) { gi_graph *t = g->transposed; g->transposed = nullptr; t->borrowed = false; gi_graph_destroy(t); }
    g->transposed = nullptr;
    for (auto &kv : g->blocked) { kv.second->borrowed = false; gi_graph_destroy(kv.second); }
    g->blocked.clear();
    if (g->layout) {  // the degree-ordered copy carries the weights too
        release_layout(g);
        GI_TRY(build_layout(g));
    }
    return GI_OK;
}

int gi_graph_num_vertices(const gi_graph *g, int32_t *out) {
    if (!g || !out) return fail(GI_ERR_INVALID, "null argument");
    *out = g->V;
    return GI_OK;
}

int gi_graph_num_edges(const gi_graph *g, int64_t *out) {
    if (!g || !out) return fail(GI_ERR_INVALID, "null argument");
    *out = g->E;
    return GI_OK;
}

int gi_graph_read_csr(const gi_graph *g, int32_t *offsets, int32_t *nbrs, int32_t *weights) {
    if (!g) return fail(GI_ERR_INVALID, "null graph");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "blocked graphs have no CSR");
    if (offsets) GI_HIP(hipMemcpy(offsets, g->d_off, sizeof(int32_t) * (g->V + 1), hipMemcpyDeviceToHost));
    if (nbrs && g->E) GI_HIP(hipMemcpy(nbrs, g->d_nbr, sizeof(int32_t) * g->E, hipMemcpyDeviceToHost));
    if (weights) {
        if (!g->d_w) return fail(GI_ERR_INVALID, "graph has no weights");
        if (g->E) GI_HIP(hipMemcpy(weights, g->d_w, sizeof(int32_t) * g->E, hipMemcpyDeviceToHost));
    }
    return GI_OK;
}

int gi_graph_device_csr(const gi_graph *g, const int32_t **offsets, const int32_t **nbrs,
                        const int32_t **weights) {
    if (!g) return fail(GI_ERR_INVALID, "null graph");
    if (offsets) *offsets = g->d_off;
    if (nbrs) *nbrs = g->d_nbr;
    if (weights) *weights = g->d_w;
    return GI_OK;
}

int gi_graph_transpose(gi_graph *g, gi_graph **out) {
    if (!g || !out) return fail(GI_ERR_INVALID, "null argument");
    if (g->transposed) { *out = g->transposed; return GI_OK; }
    if (!g->directed) {  // symmetric CSR is its own transpose
        g->transposed = g;
        *out = g;
        return GI_OK;
    }
    hipStream_t s = 0;
    GI_TRY(ensure_coo(g, s));
    int64_t E = g->E;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    int32_t *vals = nullptr, *vals2 = nullptr;
    int rc = dalloc(&keys, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&keys2, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&vals, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&vals2, (size_t)E + 1);
    gi_graph *t = new gi_graph();
    t->V = g->V;
    t->E = E;
    t->directed = 1;
    int32_t *deg = nullptr;
    if (rc == GI_OK) rc = dalloc(&t->d_nbr, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&t->d_off, (size_t)g->V + 2);
    if (rc == GI_OK) rc = dalloc(&deg, (size_t)g->V + 2);
    if (rc == GI_OK && g->d_w) rc = dalloc(&t->d_w, (size_t)E + 1);
    if (rc == GI_OK && E) {
        transpose_keys_kernel<<<grid_for(E), 256, 0, s>>>(g->d_esrc, g->d_nbr, (int32_t)E, keys, vals);
        size_t bytes = 0;
        int b = 32 + bits_for(g->V);
        rocprim::double_buffer<uint64_t> kb(keys, keys2);
        rocprim::double_buffer<int32_t> vb(vals, vals2);
        (void)rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, (size_t)E, 0, b, s);
        void *tmp = nullptr;
        rc = dmalloc(&tmp, bytes);
        if (rc == GI_OK) {
            if (rocprim::radix_sort_pairs(tmp, bytes, kb, vb, (size_t)E, 0, b, s) != hipSuccess)
                rc = fail(GI_ERR_HIP, "transpose sort");
            dfree(tmp);
        }
        if (rc == GI_OK) {
            fill_i32(deg, g->V + 1, 0, s);
            transpose_scatter_kernel<<<grid_for(E), 256, 0, s>>>(kb.current(), vb.current(), g->d_w, (int32_t)E, t->d_nbr, t->d_w, deg);
            rc = exclusive_scan_i32(deg, t->d_off, g->V + 1, s);
        }
    } else if (rc == GI_OK) {
        rc = fill_i32(t->d_off, g->V + 1, 0, s);
    }
    if (rc == GI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(GI_ERR_HIP, "transpose sync");
    dfree(keys); dfree(keys2); dfree(vals); dfree(vals2); dfree(deg);
    if (rc != GI_OK) { gi_graph_destroy(t); return rc; }
    t->borrowed = true;
    t->transposed = g;
    t->transpose_owner = g;
    g->transposed = t;
    *out = t;
    return GI_OK;
}

int gi_graph_blocked(gi_graph *g, int32_t block_size, gi_graph **out) {
    if (!g || !out || block_size <= 0) return fail(GI_ERR_INVALID, "bad arguments");
    auto it = g->blocked.find(block_size);
    if (it != g->blocked.end()) { *out = it->second; return GI_OK; }
    hipStream_t s = 0;
    GI_TRY(ensure_coo(g, s));
    int64_t E = g->E;
    int32_t nb = (int32_t)ceil_div(g->V, block_size);
    gi_graph *bg = new gi_graph();
    bg->V = g->V;
    bg->E = E;
    bg->directed = g->directed;
    bg->block_size = block_size;
    bg->num_buckets = nb;
    int32_t *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *vals2 = nullptr;
    int rc = dalloc(&keys, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&keys2, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&vals, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&vals2, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&bg->d_esrc, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&bg->d_nbr, (size_t)E + 1);
    if (rc == GI_OK && g->d_w) rc = dalloc(&bg->d_w, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&bg->d_bucket_end, (size_t)nb + 1);
    if (rc == GI_OK && E) {
        block_keys_kernel<<<grid_for(E), 256, 0, s>>>(g->d_nbr, (int32_t)E, block_size, keys, vals);
        size_t bytes = 0;
        int b = bits_for(nb + 1);
        rocprim::double_buffer<int32_t> kb(keys, keys2);
        rocprim::double_buffer<int32_t> vb(vals, vals2);
        (void)rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, (size_t)E, 0, b, s);
        void *tmp = nullptr;
        rc = dmalloc(&tmp, bytes);
        if (rc == GI_OK) {
            if (rocprim::radix_sort_pairs(tmp, bytes, kb, vb, (size_t)E, 0, b, s) != hipSuccess)
                rc = fail(GI_ERR_HIP, "blocking sort");
            dfree(tmp);
        }
        if (rc == GI_OK) {
            block_gather_kernel<<<grid_for(E), 256, 0, s>>>(vb.current(), g->d_esrc, g->d_nbr, g->d_w, (int32_t)E, bg->d_esrc, bg->d_nbr, bg->d_w);
            bucket_end_kernel<<<grid_for(nb), 256, 0, s>>>(kb.current(), (int32_t)E, nb, bg->d_bucket_end);
        }
    } else if (rc == GI_OK) {
        rc = fill_i32(bg->d_bucket_end, nb, 0, s);
    }
    if (rc == GI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(GI_ERR_HIP, "blocking sync");
    dfree(keys); dfree(keys2); dfree(vals); dfree(vals2);
    if (rc != GI_OK) { gi_graph_destroy(bg); return rc; }
    bg->borrowed = true;
    g->blocked[block_size] = bg;
    *out = bg;
    return GI_OK;
}

int gi_graph_destroy(gi_graph *g) {
    if (!g) return GI_OK;
    if (g->borrowed) return fail(GI_ERR_INVALID, "graph is owned by another graph");
    for (auto &kv : g->blocked) {
        kv.second->borrowed = false;
        gi_graph_destroy(kv.second);
    }
    g->blocked.clear();
    if (g->transposed && g->transposed != g && g->transposed->transpose_owner == g) {
        gi_graph *t = g->transposed;
        t->borrowed = false;
        t->transposed = nullptr;
        gi_graph_destroy(t);
    }
    dfree(g->d_off);
    dfree(g->d_nbr);
    dfree(g->d_w);
    dfree(g->d_esrc);
    dfree(g->d_bucket_end);
    dfree(g->d_scratch);
    dfree(g->d_visited);
    if (g->work) gi_frontier_destroy(g->work);
    release_layout(g);
    release_fused(g);
    release_pr(g);
    release_cc(g);
    release_sssp(g);
    delete g;
    return GI_OK;
}

}  // extern "C"

// ====================================================================== memory
namespace gi {
static thread_local hipEvent_t g_timer_ev[2] = {nullptr, nullptr};
}

extern "C" {

int gi_malloc(void **ptr, uint64_t bytes) {
    if (!ptr) return fail(GI_ERR_INVALID, "null argument");
    return dmalloc(ptr, (size_t)bytes);
}

int gi_free(void *ptr) {
    if (ptr) GI_HIP(hipFree(ptr));
    return GI_OK;
}

int gi_memcpy_h2d(void *dst, const void *src, uint64_t bytes, void *stream) {
    if (!bytes) return GI_OK;
    if (!dst || !src) return fail(GI_ERR_INVALID, "null argument");
    GI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    GI_HIP(hipStreamSynchronize((hipStream_t)stream));  // host buffer may be freed on return
    return GI_OK;
}

int gi_memcpy_d2h(void *dst, const void *src, uint64_t bytes, void *stream) {
    if (!bytes) return GI_OK;
    if (!dst || !src) return fail(GI_ERR_INVALID, "null argument");
    GI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    GI_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GI_OK;
}

int gi_memcpy_d2d(void *dst, const void *src, uint64_t bytes, void *stream) {
    if (!bytes) return GI_OK;
    if (!dst || !src) return fail(GI_ERR_INVALID, "null argument");
    GI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return GI_OK;
}

int gi_fill_i32(int32_t *ptr, int64_t count, int32_t value, void *stream) {
    if (!ptr && count > 0) return fail(GI_ERR_INVALID, "null argument");
    return fill_i32(ptr, count, value, (hipStream_t)stream);
}

int gi_fill_f32(float *ptr, int64_t count, float value, void *stream) {
    if (!ptr && count > 0) return fail(GI_ERR_INVALID, "null argument");
    return fill_f32(ptr, count, value, (hipStream_t)stream);
}

int gi_fill_u8(uint8_t *ptr, int64_t count, uint8_t value, void *stream) {
    if (!ptr && count > 0) return fail(GI_ERR_INVALID, "null argument");
    return fill_u8(ptr, count, value, (hipStream_t)stream);
}

int gi_stream_create(void **stream) {
    if (!stream) return fail(GI_ERR_INVALID, "null argument");
    hipStream_t s = nullptr;
    GI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return GI_OK;
}

int gi_stream_destroy(void *stream) {
    if (stream) GI_HIP(hipStreamDestroy((hipStream_t)stream));
    return GI_OK;
}

int gi_stream_synchronize(void *stream) {
    GI_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GI_OK;
}

int32_t gi_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int gi_timer_start(void *stream) {
    for (int i = 0; i < 2; ++i)
        if (!g_timer_ev[i]) GI_HIP(hipEventCreate(&g_timer_ev[i]));
    GI_HIP(hipEventRecord(g_timer_ev[0], (hipStream_t)stream));
    return GI_OK;
}

int gi_timer_stop(void *stream, float *ms) {
    if (!ms) return fail(GI_ERR_INVALID, "null argument");
    if (!g_timer_ev[0]) return fail(GI_ERR_INVALID, "timer not started");
    GI_HIP(hipEventRecord(g_timer_ev[1], (hipStream_t)stream));
    GI_HIP(hipEventSynchronize(g_timer_ev[1]));
    GI_HIP(hipEventElapsedTime(ms, g_timer_ev[0], g_timer_ev[1]));
    return GI_OK;
}

}  // extern "C"

// ====================================================================== batch of small graphs
namespace gi {

// Vertex count of graph i: piecewise-linear log-uniform over the octaves of
// [vmin, vmax] from one 32-bit draw, integer-only so host and oracle agree.
static int32_t batch_graph_size(uint64_t seed, int32_t i, int32_t vmin, int32_t vmax) {
    int32_t octaves = 0;
    while (((int64_t)vmin << (octaves + 1)) <= vmax) ++octaves;
    if (octaves == 0) return vmin;
    uint32_t a = uniform_below(hash64(seed, 7, (uint64_t)i), (uint32_t)octaves * 65536u);
    int64_t lo = (int64_t)vmin << (a / 65536u);
    int64_t v = lo + (lo * (int64_t)(a % 65536u)) / 65536;
    return (int32_t)(v > vmax ? vmax : v);
}

__global__ void batch_keys_kernel(const int64_t *pair_start, const int32_t *vbase,
                                  const int32_t *vcount, int32_t G, int64_t M, uint64_t seed,
                                  int b, uint64_t *keys) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < M; e += stride) {
        int32_t lo = 0, hi = G;  // last graph whose first pair <= e
        while (hi - lo > 1) {
            int32_t mid = (lo + hi) / 2;
            if (pair_start[mid] <= e) lo = mid;
            else hi = mid;
        }
        uint32_t n = (uint32_t)vcount[lo];
        uint64_t u = (uint64_t)vbase[lo] + uniform_below(hash64(seed, 8, 2 * (uint64_t)e), n);
        uint64_t v = (uint64_t)vbase[lo] + uniform_below(hash64(seed, 8, 2 * (uint64_t)e + 1), n);
        keys[e] = (u << b) | v;
        keys[M + e] = (v << b) | u;
    }
}

}  // namespace gi

extern "C" int gi_graph_uniform_batch(int32_t num_graphs, int32_t vmin, int32_t vmax,
                                      int32_t avg_degree, uint64_t seed, int32_t *vertex_offsets,
                                      gi_graph **out) {
    if (!out || !vertex_offsets) return fail(GI_ERR_INVALID, "null argument");
    *out = nullptr;
    if (num_graphs < 1 || vmin < 1 || vmax < vmin || avg_degree < 0)
        return fail(GI_ERR_RANGE, "bad batch parameters");
    std::vector<int32_t> vbase(num_graphs), vcount(num_graphs);
    std::vector<int64_t> pstart(num_graphs);
    int64_t V = 0, M = 0;
    for (int32_t i = 0; i < num_graphs; ++i) {
        int32_t n = batch_graph_size(seed, i, vmin, vmax);
        vbase[i] = (int32_t)V;
        vcount[i] = n;
        vertex_offsets[i] = (int32_t)V;
        pstart[i] = M;
        V += n;
        M += (int64_t)n * avg_degree / 2;
        if (V >= (1ll << 31) - 1) return fail(GI_ERR_RANGE, "batch exceeds int32 vertex ids");
    }
    vertex_offsets[num_graphs] = (int32_t)V;
    GI_TRY(check_sizes(V, 2 * M));
    int b = bits_for(V);
    uint64_t *keys = nullptr;
    int64_t *d_ps = nullptr;
    int32_t *d_vb = nullptr, *d_vc = nullptr;
    GI_TRY(dalloc(&keys, (size_t)(2 * M) + 1));
    int rc = dalloc(&d_ps, (size_t)num_graphs);
    if (rc == GI_OK) rc = dalloc(&d_vb, (size_t)num_graphs);
    if (rc == GI_OK) rc = dalloc(&d_vc, (size_t)num_graphs);
    if (rc == GI_OK && (hipMemcpy(d_ps, pstart.data(), sizeof(int64_t) * num_graphs, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(d_vb, vbase.data(), sizeof(int32_t) * num_graphs, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(d_vc, vcount.data(), sizeof(int32_t) * num_graphs, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(GI_ERR_HIP, "copy batch layout");
    if (rc == GI_OK && M) {
        batch_keys_kernel<<<grid_for(M), 256>>>(d_ps, d_vb, d_vc, num_graphs, M, seed, b, keys);
        if (hipGetLastError() != hipSuccess) rc = fail(GI_ERR_HIP, "batch keys kernel");
    }
    dfree(d_ps);
    dfree(d_vb);
    dfree(d_vc);
    if (rc != GI_OK) { dfree(keys); return rc; }
    return finish_build(new gi_graph(), (int32_t)V, keys, 2 * M, b, 0, out);
}
