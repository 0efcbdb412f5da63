// Vertex frontiers, representation conversions, the level list, and the
// vertex-map operator.
//
// Replaces VertexFrontier (runtime/infra_gpu/vertex_frontier.h:26-432),
// vertex_representation.h:9-158, VertexFrontierList (list.h:18-116) and
// vertexset_apply (src/graphit/operators.cpp:10-65).
//
// Layout: sparse queue int32[V] x2, bitmap uint64[ceil(V/64)] x2 (one word
// per wavefront of 64 vertices), boolmap uint8[V] x2, counts int32[2] plus an
// overflow flag, per-vertex dedup epochs int32[V]. Every enqueue path
// deduplicates, so V entries suffice where the reference reserves 6V
// (vertex_frontier.h:131-132).
#include "gi_internal.h"

namespace gi {

// ------------------------------------------------------------ conversions
__global__ void sparse_to_bitmap_kernel(const int32_t *q, const int32_t *count, uint64_t *bm) {
    int32_t n = *count;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        bit_claim(bm, q[i]);
}

__global__ void sparse_to_boolmap_kernel(const int32_t *q, const int32_t *count, uint8_t *bmap) {
    int32_t n = *count;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        bmap[q[i]] = 1;
}

// Compaction: one lane per vertex, wave-aggregated append (ballot + mbcnt).
__global__ void bitmap_to_sparse_kernel(const uint64_t *bm, int32_t V, int32_t *q,
                                        int32_t *count) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t limit = ceil_div(V, 64) * 64;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < limit; v += stride) {
        bool in = v < V && bit_test(bm, (int32_t)v);
        int32_t pos = wave_atomic_add(count, in);
        if (in) q[pos] = (int32_t)v;
    }
}

__global__ void boolmap_to_sparse_kernel(const uint8_t *bmap, int32_t V, int32_t *q,
                                         int32_t *count) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t limit = ceil_div(V, 64) * 64;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < limit; v += stride) {
        bool in = v < V && bmap[v] != 0;
        int32_t pos = wave_atomic_add(count, in);
        if (in) q[pos] = (int32_t)v;
    }
}

__global__ void bitmap_to_boolmap_kernel(const uint64_t *bm, int32_t V, uint8_t *bmap) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride)
        bmap[v] = bit_test(bm, (int32_t)v) ? 1 : 0;
}

// Each lane builds one 64-bit word from 64 bytes.
__global__ void boolmap_to_bitmap_kernel(const uint8_t *bmap, int32_t V, uint64_t *bm) {
    int64_t words = ceil_div(V, 64);
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
        uint64_t x = 0;
        int64_t base = w * 64;
        for (int b = 0; b < 64; ++b)
            if (base + b < V && bmap[base + b]) x |= 1ull << b;
        bm[w] = x;
    }
}

static int grid_for(int64_t n, int block = 256) {
    int64_t want = ceil_div(n, block);
    int cap = persistent_grid(block);
    return (int)(want < 1 ? 1 : (want < cap ? want : cap));
}

__global__ void copy_count_kernel(const int32_t *from, int32_t *to) { *to = *from; }

// Convert the input side to fmt; the other side is scratch for the result
// and the sides are swapped, so counts carry over.
int frontier_convert(gi_frontier *f, int32_t fmt, hipStream_t s) {
    if (f->fmt == fmt) return GI_OK;
    int in = f->side, out = f->side ^ 1;
    int32_t V = f->V;
    if (fmt == GI_FMT_SPARSE) {
        GI_TRY(fill_i32(f->counts + out, 1, 0, s));
        if (f->fmt == GI_FMT_BITMAP)
            bitmap_to_sparse_kernel<<<grid_for(V), 256, 0, s>>>(f->bm[in], V, f->q[in], f->counts + out);
        else
            boolmap_to_sparse_kernel<<<grid_for(V), 256, 0, s>>>(f->bmap[in], V, f->q[in], f->counts + out);
        GI_LAUNCH_CHECK();
        copy_count_kernel<<<1, 1, 0, s>>>(f->counts + out, f->counts + in);
        GI_LAUNCH_CHECK();
    } else if (fmt == GI_FMT_BITMAP) {
        if (f->fmt == GI_FMT_SPARSE) {
            GI_TRY(fill_u64(f->bm[in], f->words, 0, s));
            sparse_to_bitmap_kernel<<<grid_for(V), 256, 0, s>>>(f->q[in], f->counts + in, f->bm[in]);
        } else {
            boolmap_to_bitmap_kernel<<<grid_for(f->words), 256, 0, s>>>(f->bmap[in], V, f->bm[in]);
        }
        GI_LAUNCH_CHECK();
    } else if (fmt == GI_FMT_BOOLMAP) {
        if (f->fmt == GI_FMT_SPARSE) {
            GI_TRY(fill_u8(f->bmap[in], V, 0, s));
            sparse_to_boolmap_kernel<<<grid_for(V), 256, 0, s>>>(f->q[in], f->counts + in, f->bmap[in]);
        } else {
            bitmap_to_boolmap_kernel<<<grid_for(V), 256, 0, s>>>(f->bm[in], V, f->bmap[in]);
        }
        GI_LAUNCH_CHECK();
    } else {
        return fail(GI_ERR_INVALID, "unknown frontier format %d", fmt);
    }
    (void)out;
    f->fmt = fmt;
    return GI_OK;
}

// Empty the output side before an operator writes into it in format fmt.
int frontier_clear_out(gi_frontier *f, int32_t fmt, hipStream_t s) {
    int out = f->side ^ 1;
    GI_TRY(fill_i32(f->counts + out, 1, 0, s));
    if (fmt == GI_FMT_BITMAP) GI_TRY(fill_u64(f->bm[out], f->words, 0, s));
    if (fmt == GI_FMT_BOOLMAP) GI_TRY(fill_u8(f->bmap[out], f->V, 0, s));
    return GI_OK;
}

Sink make_sink(gi_frontier *f, int32_t fmt, bool dedup) {
    Sink k{};
    if (!f) { k.mode = SINK_NONE; return k; }
    int out = f->side ^ 1;
    k.mode = fmt == GI_FMT_SPARSE ? SINK_SPARSE : (fmt == GI_FMT_BITMAP ? SINK_BITMAP : SINK_BOOLMAP);
    k.cap = f->cap;
    k.q = f->q[out];
    k.bm = f->bm[out];
    k.bmap = f->bmap[out];
    k.count = f->counts + out;
    k.overflow = f->overflow();
    if (dedup && fmt == GI_FMT_SPARSE) {
        f->epoch += 1;
        k.dedup = f->dedup;
        k.epoch = f->epoch;
    }
    return k;
}

Member make_member(const gi_frontier *f) {
    Member m{};
    m.mode = -1;
    if (!f) return m;
    m.mode = f->fmt;
    m.bm = f->bm[f->side];
    m.bmap = f->bmap[f->side];
    return m;
}

// ------------------------------------------------------------ vertex ops
__global__ void iota_kernel(int32_t *p, int32_t n) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

__global__ void set_one_kernel(int32_t *q, int32_t *count, int32_t v, int32_t cap,
                               int32_t *overflow) {
    int32_t n = *count;
    if (n >= cap) { *overflow = 1; return; }
    q[n] = v;
    *count = n + 1;
}

__global__ void dedup_kernel(const int32_t *q, const int32_t *count, int32_t *dedup, int32_t epoch,
                             int32_t *out_q, int32_t *out_count) {
    int32_t n = *count;
    int64_t limit = ceil_div(n, 64) * 64;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < limit; i += (int64_t)gridDim.x * blockDim.x) {
        bool keep = false;
        int32_t v = 0;
        if (i < n) {
            v = q[i];
            keep = atomicMax(&dedup[v], epoch) < epoch;
        }
        int32_t pos = wave_atomic_add(out_count, keep);
        if (keep) out_q[pos] = v;
    }
}

__global__ void vop_kernel(VEnum en, int32_t vop, gi_op_args a, int32_t V) {
    int32_t n = en.size();
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int32_t v = en.at(i);
        switch (vop) {
            case GI_VOP_FILL_I32: ((int32_t *)a.data0)[v] = a.i0; break;
            case GI_VOP_IOTA_I32: ((int32_t *)a.data0)[v] = v; break;
            case GI_VOP_CC_PJUMP: {  // apps/cc.cpp:43-50
                int32_t *ID = (int32_t *)a.data0;
                int32_t y = ID[v];
                int32_t x = ID[y];
                if (x != y) {
                    ID[v] = x;
                    int32_t *upd = (int32_t *)a.data1;
                    if (upd[1] == 0) upd[1] = 1;
                }
                break;
            }
            case GI_VOP_PR_CONTRIB: {  // apps/pr.cpp:28-30 (out-degree from CSR)
                const int32_t *off = (const int32_t *)a.data2;
                float *old_rank = (float *)a.data0;
                float *contrib = (float *)a.data1;
                contrib[v] = old_rank[v] / (float)(off[v + 1] - off[v]);
                break;
            }
            case GI_VOP_PR_UPDATE: {  // apps/pr.cpp:36-41
                float *old_rank = (float *)a.data0;
                float *new_rank = (float *)a.data1;
                float beta = (float)((1.0 - (double)a.f0) / (double)V);
                new_rank[v] = beta + a.f0 * new_rank[v];
                old_rank[v] = new_rank[v];
                new_rank[v] = 0.0f;
                break;
            }
            case GI_VOP_PR_RESET: {  // apps/pr.cpp:43-46
                ((float *)a.data0)[v] = (float)(1.0 / (double)V);
                ((float *)a.data1)[v] = 0.0f;
                break;
            }
            case GI_VOP_BC_RESET:  // apps/bc.cpp:28-32
                ((float *)a.data0)[v] = 0.0f;
                ((double *)a.data1)[v] = 0.0;
                ((uint8_t *)a.data2)[v] = 0;
                break;
            case GI_VOP_BC_MARK:  // apps/bc.cpp:42-47
                ((uint8_t *)a.data2)[v] = (uint8_t)a.i0;
                break;
            case GI_VOP_BC_BACKWARD_V: {  // apps/bc.cpp:49-52
                ((uint8_t *)a.data2)[v] = 1;
                float *dep = (float *)a.data0;
                dep[v] += (float)(1.0 / ((double *)a.data1)[v]);
                break;
            }
            case GI_VOP_BC_FINAL: {  // apps/bc.cpp:58-63
                float *dep = (float *)a.data0;
                double p = ((double *)a.data1)[v];
                if (p != 0.0) dep[v] = (float)(((double)dep[v] - 1.0 / p) * p);
                else dep[v] = 0.0f;
                break;
            }
            default: break;
        }
    }
}

// ------------------------------------------------------------ level list
__global__ void list_insert_kernel(int32_t *vertices, int32_t *level_idx, int32_t level,
                                   const int32_t *q, const int32_t *count, int32_t max_elems,
                                   int32_t *overflow) {
    int32_t n = *count;
    int32_t at = level_idx[level];
    if (at + n > max_elems) {  // keep the level empty and report
        if (blockIdx.x == 0 && threadIdx.x == 0) { level_idx[level + 1] = at; *overflow = 1; }
        return;
    }
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        vertices[at + i] = q[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) level_idx[level + 1] = at + n;
}

__global__ void list_retrieve_kernel(const int32_t *vertices, const int32_t *level_idx,
                                     int32_t level, int32_t *q, int32_t *count) {
    int32_t at = level_idx[level - 1];
    int32_t n = level_idx[level] - at;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        q[i] = vertices[at + i];
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = n;
}

}  // namespace gi

using namespace gi;

struct gi_frontier_list {
    int32_t max_elems = 0;
    int32_t levels = 0;
    int32_t *d_level_idx = nullptr;
    int32_t *d_vertices = nullptr;
    int32_t level_cap = 0;
};

extern "C" {

int gi_frontier_create(int32_t V, int32_t init_all, gi_frontier **out) {
    if (!out || V <= 0) return fail(GI_ERR_INVALID, "bad arguments");
    *out = nullptr;
    gi_frontier *f = new gi_frontier();
    f->V = V;
    f->cap = V;
    f->words = bitmap_words(V);
    int rc = GI_OK;
    for (int i = 0; i < 2 && rc == GI_OK; ++i) {
        rc = dalloc(&f->q[i], (size_t)V + 64);
        if (rc == GI_OK) rc = dalloc(&f->bm[i], (size_t)f->words + 1);
        if (rc == GI_OK) rc = dalloc(&f->bmap[i], (size_t)ceil_div(V, 4) * 4 + 64);
    }
    if (rc == GI_OK) rc = dalloc(&f->counts, 4);
    if (rc == GI_OK) rc = dalloc(&f->dedup, (size_t)V);
    hipStream_t s = 0;
    if (rc == GI_OK) rc = fill_i32(f->counts, 4, 0, s);
    if (rc == GI_OK) rc = fill_i32(f->dedup, V, 0, s);
    for (int i = 0; i < 2 && rc == GI_OK; ++i) {
        rc = fill_u64(f->bm[i], f->words, 0, s);
        if (rc == GI_OK) rc = fill_u8(f->bmap[i], V, 0, s);
    }
    if (rc == GI_OK && init_all) {
        iota_kernel<<<grid_for(V), 256, 0, s>>>(f->q[0], V);
        rc = fill_i32(f->counts, 1, V, s);
    }
    if (rc == GI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(GI_ERR_HIP, "frontier init");
    if (rc != GI_OK) { gi_frontier_destroy(f); return rc; }
    *out = f;
    return GI_OK;
}

int gi_frontier_destroy(gi_frontier *f) {
    if (!f) return GI_OK;
    for (int i = 0; i < 2; ++i) { dfree(f->q[i]); dfree(f->bm[i]); dfree(f->bmap[i]); }
    dfree(f->counts);
    dfree(f->dedup);
    delete f;
    return GI_OK;
}

int gi_frontier_add_vertex(gi_frontier *f, int32_t v, void *stream) {
    if (!f) return fail(GI_ERR_INVALID, "null frontier");
    if (v < 0 || v >= f->V) return fail(GI_ERR_RANGE, "vertex %d out of range [0,%d)", v, f->V);
    hipStream_t s = (hipStream_t)stream;
    GI_TRY(frontier_convert(f, GI_FMT_SPARSE, s));
    set_one_kernel<<<1, 1, 0, s>>>(f->q[f->side], f->count_in(), v, f->cap, f->overflow());
    GI_LAUNCH_CHECK();
    return GI_OK;
}

int gi_frontier_clear(gi_frontier *f, void *stream) {
    if (!f) return fail(GI_ERR_INVALID, "null frontier");
    hipStream_t s = (hipStream_t)stream;
    GI_TRY(fill_i32(f->count_in(), 1, 0, s));
    GI_TRY(fill_u64(f->bm[f->side], f->words, 0, s));
    GI_TRY(fill_u8(f->bmap[f->side], f->V, 0, s));
    f->fmt = GI_FMT_SPARSE;
    return GI_OK;
}

int gi_frontier_size(gi_frontier *f, int32_t *out, void *stream) {
    if (!f || !out) return fail(GI_ERR_INVALID, "null argument");
    return read_count(f->count_in(), out, (hipStream_t)stream);
}

int gi_frontier_convert(gi_frontier *f, int32_t format, void *stream) {
    if (!f) return fail(GI_ERR_INVALID, "null frontier");
    return frontier_convert(f, format, (hipStream_t)stream);
}

int gi_frontier_format(const gi_frontier *f, int32_t *format) {
    if (!f || !format) return fail(GI_ERR_INVALID, "null argument");
    *format = f->fmt;
    return GI_OK;
}

int gi_frontier_read(gi_frontier *f, int32_t *out, int32_t capacity, int32_t *count,
                     void *stream) {
    if (!f || !count) return fail(GI_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    GI_TRY(frontier_convert(f, GI_FMT_SPARSE, s));
    int32_t n = 0;
    GI_TRY(read_count(f->count_in(), &n, s));
    *count = n;
    if (n > capacity) return fail(GI_ERR_RANGE, "frontier has %d members, capacity %d", n, capacity);
    if (n && out) {
        GI_HIP(hipMemcpyAsync(out, f->q[f->side], sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        GI_HIP(hipStreamSynchronize(s));
    }
    return GI_OK;
}

int gi_frontier_swap(gi_frontier *f, int32_t format, void *stream) {
    if (!f) return fail(GI_ERR_INVALID, "null frontier");
    if (format < GI_FMT_SPARSE || format > GI_FMT_BOOLMAP) return fail(GI_ERR_INVALID, "bad format");
    f->side ^= 1;
    f->fmt = format;
    // clear the new output side's count (the reference memsets it,
    // vertex_frontier.h:226)
    return fill_i32(f->count_out(), 1, 0, (hipStream_t)stream);
}

int gi_frontier_dedup(gi_frontier *f, void *stream) {
    if (!f) return fail(GI_ERR_INVALID, "null frontier");
    hipStream_t s = (hipStream_t)stream;
    GI_TRY(frontier_convert(f, GI_FMT_SPARSE, s));
    if (f->fmt != GI_FMT_SPARSE) return GI_OK;
    f->epoch += 1;
    int out = f->side ^ 1;
    GI_TRY(fill_i32(f->counts + out, 1, 0, s));
    dedup_kernel<<<grid_for(f->V), 256, 0, s>>>(f->q[f->side], f->count_in(), f->dedup, f->epoch,
                                                 f->q[out], f->counts + out);
    GI_LAUNCH_CHECK();
    f->side = out;
    return fill_i32(f->count_out(), 1, 0, s);
}

int gi_vertexset_apply(gi_graph *g, gi_frontier *f, int32_t vop, const gi_op_args *args,
                       void *stream) {
    if (!g || !args) return fail(GI_ERR_INVALID, "null argument");
    if (vop < GI_VOP_FILL_I32 || vop > GI_VOP_BC_FINAL) return fail(GI_ERR_INVALID, "unknown vertex op %d", vop);
    hipStream_t s = (hipStream_t)stream;
    VEnum en{};
    en.V = g->V;
    if (f) {
        GI_TRY(frontier_convert(f, GI_FMT_SPARSE, s));
        en.q = f->q[f->side];
        en.count = f->count_in();
    }
    gi_op_args a = *args;
    if (vop == GI_VOP_PR_CONTRIB) a.data2 = g->d_off;
    vop_kernel<<<grid_for(g->V), 256, 0, s>>>(en, vop, a, g->V);
    GI_LAUNCH_CHECK();
    return GI_OK;
}

int gi_frontier_list_create(int32_t max_elems, gi_frontier_list **out) {
    if (!out || max_elems <= 0) return fail(GI_ERR_INVALID, "bad arguments");
    gi_frontier_list *l = new gi_frontier_list();
    l->max_elems = max_elems;
    l->level_cap = max_elems + 1;
    int rc = dalloc(&l->d_level_idx, (size_t)l->level_cap + 1);
    if (rc == GI_OK) rc = dalloc(&l->d_vertices, (size_t)max_elems);
    if (rc == GI_OK) rc = fill_i32(l->d_level_idx, 1, 0, 0);
    if (rc != GI_OK) { gi_frontier_list_destroy(l); return rc; }
    *out = l;
    return GI_OK;
}

int gi_frontier_list_insert(gi_frontier_list *l, gi_frontier *f, void *stream) {
    if (!l || !f) return fail(GI_ERR_INVALID, "null argument");
    if (l->levels + 1 >= l->level_cap) return fail(GI_ERR_RANGE, "frontier list is full");
    hipStream_t s = (hipStream_t)stream;
    GI_TRY(frontier_convert(f, GI_FMT_SPARSE, s));
    // capacity check on the host, as builtin_insert reads sizes (list.h:58-70)
    int32_t n = 0, at = 0;
    GI_TRY(read_count(f->count_in(), &n, s));
    GI_TRY(read_count(l->d_level_idx + l->levels, &at, s));
    if ((int64_t)at + n > l->max_elems)
        return fail(GI_ERR_RANGE, "frontier list holds %d of %d; cannot add %d", at, l->max_elems, n);
    list_insert_kernel<<<grid_for(f->V), 256, 0, s>>>(l->d_vertices, l->d_level_idx, l->levels,
                                                       f->q[f->side], f->count_in(), l->max_elems,
                                                       f->overflow());
    GI_LAUNCH_CHECK();
    l->levels++;
    return GI_OK;
}

int gi_frontier_list_retrieve(gi_frontier_list *l, gi_frontier *f, void *stream) {
    if (!l || !f) return fail(GI_ERR_INVALID, "null argument");
    if (l->levels == 0) return fail(GI_ERR_UNDERFLOW, "frontier list is empty");
    hipStream_t s = (hipStream_t)stream;
    list_retrieve_kernel<<<grid_for(f->V), 256, 0, s>>>(l->d_vertices, l->d_level_idx, l->levels,
                                                         f->q[f->side], f->count_in());
    GI_LAUNCH_CHECK();
    f->fmt = GI_FMT_SPARSE;
    l->levels--;
    return GI_OK;
}

int gi_frontier_list_levels(const gi_frontier_list *l, int32_t *out) {
    if (!l || !out) return fail(GI_ERR_INVALID, "null argument");
    *out = l->levels;
    return GI_OK;
}

int gi_frontier_list_destroy(gi_frontier_list *l) {
    if (!l) return GI_OK;
    dfree(l->d_level_idx);
    dfree(l->d_vertices);
    delete l;
    return GI_OK;
}

}  // extern "C"
