// Degree-ordered layout: an internal copy of a graph whose vertices are
// renumbered by decreasing degree (ties keep their original order). The
// vertices that most edges point at then occupy the first words of every
// per-vertex array and bitmap, so the random probes of frontier bits, labels
// and depths mostly hit a few cache lines. Results are mapped back to the
// original ids, so callers see no difference.
//
// Neighbour lists of the view are sorted by new id, so the widest
// neighbours come first: pull scans meet them early, and a wavefront walking
// one list claims runs of neighbours that share a bitmap word.
#include <climits>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "gi_internal.h"

namespace gi {

__global__ void layout_keys_kernel(const int32_t *off, int32_t V, uint32_t *keys, int32_t *vals) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        keys[v] = (uint32_t)INT_MAX - (uint32_t)(off[v + 1] - off[v]);  // ascending key = descending degree
        vals[v] = v;
    }
}

__global__ void layout_inverse_kernel(const int32_t *old_of_new, const int32_t *off, int32_t V,
                                      int32_t *new_of_old, int32_t *deg_new) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        const int32_t o = old_of_new[i];
        new_of_old[o] = i;
        deg_new[i] = off[o + 1] - off[o];
    }
}

// Every edge as a packed (new source, new target) key; sorted, the keys are
// the view's CSR with each neighbour list in increasing new id (hubs first).
__global__ void layout_edge_keys_kernel(const int32_t *esrc, const int32_t *nbr, const int32_t *new_of_old,
                                        int64_t E, uint64_t *keys) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x)
        keys[e] = ((uint64_t)(uint32_t)new_of_old[esrc[e]] << 32) | (uint32_t)new_of_old[nbr[e]];
}

__global__ void layout_unpack_kernel(const uint64_t *keys, int64_t E, int32_t *nbr_new) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x)
        nbr_new[e] = (int32_t)(uint32_t)keys[e];
}

static int bits_of(int32_t n) {
    int b = 1;
    while (b < 31 && (1ll << b) < (int64_t)n) ++b;
    return b;
}

__global__ void layout_gather_kernel(const int32_t *in_new, const int32_t *new_of_old, int32_t V,
                                     int32_t *out_old) {
    for (int32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < V; o += gridDim.x * blockDim.x)
        out_old[o] = in_new[new_of_old[o]];
}

// Parents are vertex ids too: map the value back as well.
__global__ void layout_gather_ids_kernel(const int32_t *in_new, const int32_t *new_of_old,
                                         const int32_t *old_of_new, int32_t V, int32_t *out_old) {
    for (int32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < V; o += gridDim.x * blockDim.x) {
        const int32_t p = in_new[new_of_old[o]];
        out_old[o] = p < 0 ? p : old_of_new[p];
    }
}

// BFS depths of a symmetric graph: a vertex past iso_begin in the view has
// no edges, so its depth is -1 unless it is the source (src_old); its random
// read of in_new is skipped.
__device__ __forceinline__ int32_t depth_back_one(const int32_t *in_new, int32_t nw, int32_t o, int32_t iso_begin,
                                                  int32_t src_old) {
    return nw < iso_begin ? in_new[nw] : (o == src_old ? 0 : -1);
}

__global__ void layout_gather_depth_kernel(const int32_t *in_new, const int32_t *new_of_old, int32_t V,
                                           int32_t iso_begin, int32_t src_old, int32_t *out_old) {
    for (int32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < V; o += gridDim.x * blockDim.x)
        out_old[o] = depth_back_one(in_new, new_of_old[o], o, iso_begin, src_old);
}

// The same, four consecutive vertices per thread (16-byte aligned out_old):
// one 16-byte load of new_of_old, four independent gathers in flight, one
// 16-byte store. The map-back is bound by the latency of its gathers, so the
// one-vertex form left most of the memory system idle.
__global__ void layout_gather_depth4_kernel(const int32_t *in_new, const int32_t *new_of_old, int32_t V,
                                            int32_t iso_begin, int32_t src_old, int32_t *out_old) {
    const int32_t n4 = V / 4;
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
        const int4 nw = reinterpret_cast<const int4 *>(new_of_old)[q];
        const int32_t o = 4 * q;
        int4 r;
        r.x = depth_back_one(in_new, nw.x, o, iso_begin, src_old);
        r.y = depth_back_one(in_new, nw.y, o + 1, iso_begin, src_old);
        r.z = depth_back_one(in_new, nw.z, o + 2, iso_begin, src_old);
        r.w = depth_back_one(in_new, nw.w, o + 3, iso_begin, src_old);
        reinterpret_cast<int4 *>(out_old)[q] = r;
    }
    for (int32_t o = 4 * n4 + blockIdx.x * blockDim.x + threadIdx.x; o < V; o += stride)
        out_old[o] = depth_back_one(in_new, new_of_old[o], o, iso_begin, src_old);
}

// First vertex of the view with no edges: degrees decrease along the view,
// so it is the first v with off[v] == E (V when every vertex has edges).
__global__ void layout_iso_kernel(const int32_t *off, int32_t V, int32_t *out) {
    const int64_t E = off[V];
    int32_t lo = 0, hi = V;  // answer in [lo, hi]
    while (lo < hi) {
        const int32_t mid = lo + (hi - lo) / 2;
        if ((int64_t)off[mid] >= E && off[mid + 1] == off[mid]) hi = mid;
        else lo = mid + 1;
    }
    *out = lo;
}

__global__ void layout_fill_kernel(int32_t *a, int32_t V, int32_t x) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) a[i] = x;
}

// Component minimum of the original ids: labels_new are vertex ids of the
// view. The vertices of the label `hint` (the giant component's, when the
// caller knows it; -1: none) fold into a per-thread minimum, reduced over
// the workgroup, one atomic per workgroup; other labels take one atomic per
// vertex, unless the current minimum is already lower.
__global__ void __launch_bounds__(256) layout_comp_min_kernel(const int32_t *labels_new, const int32_t *old_of_new,
                                                              int32_t V, const int32_t *hint, int32_t *comp_min) {
    __shared__ int32_t red[256 / 64];
    const int32_t h = hint ? *hint : -1;
    int32_t hm = INT_MAX;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        const int32_t r = labels_new[i], o = old_of_new[i];
        if (r == h) hm = o < hm ? o : hm;
        else if (o < comp_min[r]) atomicMin(&comp_min[r], o);
    }
    for (int k = 32; k > 0; k >>= 1) {
        const int32_t y = __shfl_xor(hm, k);
        hm = y < hm ? y : hm;
    }
    if (lane_id() == 0) red[threadIdx.x >> 6] = hm;
    __syncthreads();
    if (threadIdx.x == 0 && h >= 0) {
        for (int w = 1; w < 256 / 64; ++w) hm = red[w] < hm ? red[w] : hm;
        if (hm < INT_MAX) atomicMin(&comp_min[h], hm);
    }
}

// The most frequent label among 1024 spread-out vertices of [0, n) (one
// workgroup, one sample per thread counted against the others from LDS):
// the hint of layout_comp_min_kernel when the caller has none.
__global__ void __launch_bounds__(1024) layout_mode_kernel(const int32_t *labels, int32_t n, int32_t *out) {
    __shared__ int32_t lab[1024];
    __shared__ int32_t bn_s[1024 / 64], bl_s[1024 / 64];
    const int32_t mine = labels[(int32_t)(((uint64_t)threadIdx.x * 2654435761ull + 12345ull) % (uint64_t)n)];
    lab[threadIdx.x] = mine;
    __syncthreads();
    int32_t bn = 0;
#pragma unroll 16
    for (int32_t j = 0; j < 1024; ++j) bn += lab[j] == mine;
    int32_t bl = mine;
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t on = __shfl_xor(bn, o), ol = __shfl_xor(bl, o);
        if (on > bn || (on == bn && ol < bl)) {
            bn = on;
            bl = ol;
        }
    }
    if (lane_id() == 0) {
        bn_s[threadIdx.x >> 6] = bn;
        bl_s[threadIdx.x >> 6] = bl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 1024 / 64; ++w)
            if (bn_s[w] > bn || (bn_s[w] == bn && bl_s[w] < bl)) {
                bn = bn_s[w];
                bl = bl_s[w];
            }
        *out = bl;
    }
}

// Labels in original ids: the component minimum of the vertex's label. A
// vertex past iso_begin (of a symmetric graph: no edges) is a component of
// its own, labelled with its own id, with no gather (iso_begin = V: none).
// Four original ids per thread (16-byte index load and store, the gathers
// in flight together) when out_old is 16-byte aligned.
__global__ void layout_labels_back_kernel(const int32_t *labels_new, const int32_t *comp_min,
                                          const int32_t *new_of_old, int32_t V, int32_t iso_begin,
                                          int32_t *out_old) {
    const int32_t stride = gridDim.x * blockDim.x;
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const bool vec = (reinterpret_cast<uintptr_t>(out_old) & 15) == 0;  // (new_of_old: 256-byte aligned)
    const int32_t n4 = vec ? V / 4 : 0;
    for (int32_t q = tid; q < n4; q += stride) {
        const int4 nw = reinterpret_cast<const int4 *>(new_of_old)[q];
        const int32_t o = 4 * q;
        const int32_t l0 = nw.x < iso_begin ? labels_new[nw.x] : -1, l1 = nw.y < iso_begin ? labels_new[nw.y] : -1;
        const int32_t l2 = nw.z < iso_begin ? labels_new[nw.z] : -1, l3 = nw.w < iso_begin ? labels_new[nw.w] : -1;
        int4 r;
        r.x = l0 >= 0 ? comp_min[l0] : o;
        r.y = l1 >= 0 ? comp_min[l1] : o + 1;
        r.z = l2 >= 0 ? comp_min[l2] : o + 2;
        r.w = l3 >= 0 ? comp_min[l3] : o + 3;
        reinterpret_cast<int4 *>(out_old)[q] = r;
    }
    for (int32_t o = 4 * n4 + tid; o < V; o += stride) {
        const int32_t n = new_of_old[o];
        out_old[o] = n < iso_begin ? comp_min[labels_new[n]] : o;
    }
}

static int grid_of(int64_t n) {
    int64_t g = ceil_div(n, 256);
    return (int)(g < 1 ? 1 : (g < 8192 ? g : 8192));
}

// The per-graph layout state lives in the graph handle (gi_graph::layout).
struct GraphLayout {
    gi_graph *view = nullptr;
    int32_t *new_of_old = nullptr;
    int32_t *old_of_new = nullptr;
    std::vector<int32_t> h_new_of_old;  // for mapping host-side sources
    int32_t *tmp[2] = {nullptr, nullptr};  // results in view order
    int32_t iso_begin = 0;                 // first view vertex without edges (symmetric graphs)
};

void release_layout(gi_graph *g) {
    GraphLayout *L = g->layout;
    if (!L) return;
    if (L->view) {
        L->view->borrowed = false;
        gi_graph_destroy(L->view);
    }
    dfree(L->new_of_old);
    dfree(L->old_of_new);
    dfree(L->tmp[0]);
    dfree(L->tmp[1]);
    delete L;
    g->layout = nullptr;
}

int build_layout(gi_graph *g) {
    const int32_t V = g->V;
    const int64_t E = g->E;
    hipStream_t s = 0;
    GI_TRY(ensure_coo(g, s));
    GraphLayout *L = new GraphLayout();
    gi_graph *t = new gi_graph();
    t->V = V;
    t->E = E;
    t->directed = g->directed;
    t->degree_view = true;
    L->view = t;
    uint32_t *k0 = nullptr, *k1 = nullptr;
    int32_t *v0 = nullptr, *v1 = nullptr, *deg = nullptr;
    int rc = dalloc(&k0, (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&k1, (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&v0, (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&v1, (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&deg, (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&L->new_of_old, (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&t->d_off, (size_t)V + 2);
    if (rc == GI_OK) rc = dalloc(&t->d_nbr, (size_t)E + 1);
    if (rc == GI_OK && g->d_w) rc = dalloc(&t->d_w, (size_t)E + 1);
    if (rc == GI_OK) rc = dalloc(&L->tmp[0], (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&L->tmp[1], (size_t)V + 1);
    if (rc == GI_OK) {
        layout_keys_kernel<<<grid_of(V), 256, 0, s>>>(g->d_off, V, k0, v0);
        rocprim::double_buffer<uint32_t> kb(k0, k1);
        rocprim::double_buffer<int32_t> vb(v0, v1);
        size_t bytes = 0;
        (void)rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, (size_t)V, 0, 32, s);
        void *tmp = nullptr;
        rc = dmalloc(&tmp, bytes);
        if (rc == GI_OK && rocprim::radix_sort_pairs(tmp, bytes, kb, vb, (size_t)V, 0, 32, s) != hipSuccess)
            rc = fail(GI_ERR_HIP, "layout sort");
        dfree(tmp);
        if (rc == GI_OK) {
            // the sorted values are old_of_new; the other buffer is freed
            int32_t *res = vb.current();
            dfree(res == v0 ? v1 : v0);
            v0 = v1 = nullptr;
            L->old_of_new = res;
            layout_inverse_kernel<<<grid_of(V), 256, 0, s>>>(L->old_of_new, g->d_off, V, L->new_of_old, deg);
            rc = fill_i32(deg + V, 1, 0, s);  // the scan's last term: offsets[V] = E
            if (rc == GI_OK) rc = exclusive_scan_i32(deg, t->d_off, (int64_t)V + 1, s);
        }
        if (rc == GI_OK && E > 0) {
            uint64_t *ka = nullptr, *kb2 = nullptr;
            int32_t *wa = nullptr, *wb = nullptr;
            rc = dalloc(&ka, (size_t)E + 1);
            if (rc == GI_OK) rc = dalloc(&kb2, (size_t)E + 1);
            if (rc == GI_OK && g->d_w) rc = dalloc(&wb, (size_t)E + 1);
            if (rc == GI_OK) {
                layout_edge_keys_kernel<<<grid_of(E), 256, 0, s>>>(g->d_esrc, g->d_nbr, L->new_of_old, E, ka);
                const int end_bit = 32 + bits_of(V);
                rocprim::double_buffer<uint64_t> ek(ka, kb2);
                size_t eb = 0;
                void *etmp = nullptr;
                if (g->d_w) {  // weights travel with their edges
                    rc = dalloc(&wa, (size_t)E + 1);
                    if (rc == GI_OK && hipMemcpyAsync(wa, g->d_w, sizeof(int32_t) * E, hipMemcpyDeviceToDevice, s) !=
                                           hipSuccess)
                        rc = fail(GI_ERR_HIP, "layout weights");
                    rocprim::double_buffer<int32_t> ew(wa, wb);
                    (void)rocprim::radix_sort_pairs(nullptr, eb, ek, ew, (size_t)E, 0, end_bit, s);
                    if (rc == GI_OK) rc = dmalloc(&etmp, eb);
                    if (rc == GI_OK && rocprim::radix_sort_pairs(etmp, eb, ek, ew, (size_t)E, 0, end_bit, s) != hipSuccess)
                        rc = fail(GI_ERR_HIP, "layout edge sort");
                    if (rc == GI_OK &&
                        hipMemcpyAsync(t->d_w, ew.current(), sizeof(int32_t) * E, hipMemcpyDeviceToDevice, s) != hipSuccess)
                        rc = fail(GI_ERR_HIP, "layout weights");
                } else {
                    (void)rocprim::radix_sort_keys(nullptr, eb, ek, (size_t)E, 0, end_bit, s);
                    rc = dmalloc(&etmp, eb);
                    if (rc == GI_OK && rocprim::radix_sort_keys(etmp, eb, ek, (size_t)E, 0, end_bit, s) != hipSuccess)
                        rc = fail(GI_ERR_HIP, "layout edge sort");
                }
                if (rc == GI_OK) layout_unpack_kernel<<<grid_of(E), 256, 0, s>>>(ek.current(), E, t->d_nbr);
                if (rc == GI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(GI_ERR_HIP, "layout sync");
                dfree(etmp);
            }
            dfree(ka);
            dfree(kb2);
            dfree(wa);
            dfree(wb);
        }
        if (rc == GI_OK && hipGetLastError() != hipSuccess) rc = fail(GI_ERR_HIP, "layout kernels");
        if (rc == GI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(GI_ERR_HIP, "layout sync");
    }
    if (rc == GI_OK) {  // the edgeless tail of the view (only used for symmetric graphs)
        layout_iso_kernel<<<1, 1, 0, s>>>(t->d_off, V, L->tmp[1]);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(&L->iso_begin, L->tmp[1], sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(GI_ERR_HIP, "layout tail");
    }
    if (rc == GI_OK) {
        L->h_new_of_old.resize((size_t)V);
        if (hipMemcpy(L->h_new_of_old.data(), L->new_of_old, sizeof(int32_t) * V, hipMemcpyDeviceToHost) !=
            hipSuccess)
            rc = fail(GI_ERR_HIP, "layout copy");
    }
    dfree(k0);
    dfree(k1);
    dfree(v0);
    dfree(v1);
    dfree(deg);
    t->borrowed = true;
    g->layout = L;
    if (rc != GI_OK) {
        release_layout(g);
        return rc;
    }
    return GI_OK;
}

gi_graph *layout_view(const gi_graph *g) { return g->layout ? g->layout->view : nullptr; }

int32_t layout_iso_begin(const gi_graph *g) { return g->layout ? g->layout->iso_begin : g->V; }

int32_t layout_map_source(const gi_graph *g, int32_t v) { return g->layout->h_new_of_old[(size_t)v]; }

int32_t *layout_tmp(const gi_graph *g, int i) { return g->layout->tmp[i]; }

int layout_map_back(const gi_graph *g, const int32_t *in_new, int32_t *out_old, bool ids, hipStream_t s) {
    const GraphLayout *L = g->layout;
    if (ids)
        layout_gather_ids_kernel<<<grid_of(g->V), 256, 0, s>>>(in_new, L->new_of_old, L->old_of_new, g->V, out_old);
    else
        layout_gather_kernel<<<grid_of(g->V), 256, 0, s>>>(in_new, L->new_of_old, g->V, out_old);
    GI_LAUNCH_CHECK();
    return GI_OK;
}

int layout_depth_back(const gi_graph *g, const int32_t *depth_new, int32_t source_old, int32_t *out_old,
                      hipStream_t s) {
    const GraphLayout *L = g->layout;
    if (g->directed) return layout_map_back(g, depth_new, out_old, false, s);  // a sink may still be reached
    if (reinterpret_cast<uintptr_t>(out_old) % 16 == 0)  // new_of_old is a 256-byte aligned allocation
        layout_gather_depth4_kernel<<<grid_of(ceil_div(g->V, 4)), 256, 0, s>>>(depth_new, L->new_of_old, g->V,
                                                                              L->iso_begin, source_old, out_old);
    else
        layout_gather_depth_kernel<<<grid_of(g->V), 256, 0, s>>>(depth_new, L->new_of_old, g->V, L->iso_begin,
                                                                 source_old, out_old);
    GI_LAUNCH_CHECK();
    return GI_OK;
}

int layout_labels_back(const gi_graph *g, const int32_t *labels_new, int32_t *out_old, hipStream_t s,
                       const int32_t *hint) {
    const GraphLayout *L = g->layout;
    int32_t *comp_min = L->tmp[1];
    // on a symmetric graph the edgeless tail of the view is singletons: their
    // labels need no minimum (one atomic each would be millions of
    // device-scope atomics on RMAT); the components of the other vertices
    // lie below iso_begin, and so do their labels
    const int32_t n = g->directed ? g->V : L->iso_begin;
    if (n > 0) {
        if (!hint) {  // the most frequent label, from samples (in comp_min's spare last slot)
            layout_mode_kernel<<<1, 1024, 0, s>>>(labels_new, n, comp_min + g->V);
            hint = comp_min + g->V;
        }
        layout_fill_kernel<<<grid_of(n), 256, 0, s>>>(comp_min, n, INT_MAX);
        const int cg = grid_of(n) < 2048 ? grid_of(n) : 2048;  // (one atomic per workgroup on the hint's label)
        layout_comp_min_kernel<<<cg, 256, 0, s>>>(labels_new, L->old_of_new, n, hint, comp_min);
    }
    layout_labels_back_kernel<<<grid_of(ceil_div(g->V, 4)), 256, 0, s>>>(labels_new, comp_min, L->new_of_old, g->V, n,
                                                                         out_old);
    GI_LAUNCH_CHECK();
    return GI_OK;
}

}  // namespace gi

using namespace gi;

extern "C" {

int gi_graph_set_layout(gi_graph *g, int32_t layout) {
    if (!g) return fail(GI_ERR_INVALID, "null graph");
    if (g->borrowed || !g->d_off) return fail(GI_ERR_UNSUPPORTED, "set the layout of a base graph");
    if (layout == GI_LAYOUT_NATIVE) {
        release_layout(g);
        return GI_OK;
    }
    if (layout != GI_LAYOUT_DEGREE) return fail(GI_ERR_RANGE, "unknown layout %d", layout);
    if (g->layout) return GI_OK;
    return build_layout(g);
}

int32_t gi_graph_layout(const gi_graph *g) { return g && g->layout ? GI_LAYOUT_DEGREE : GI_LAYOUT_NATIVE; }

}  // extern "C"
