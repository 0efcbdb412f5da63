// Batched Brandes betweenness: up to 256 sources advance level-synchronously
// in one set of launches per level (apps/bc.cpp:150-209 computes one source
// per run, with a host round trip per level and per list operation).
//
// Layout (source-major, B sources, V vertices):
//   depth  int32 [B][V]   level of v from source b (-1 unreached)
//   sigma  f64   [B][V]   shortest-path counts (apps/bc.cpp:20 uses double)
//   dep    f32   [B][V]   D(v) of the backward recurrence (apps/bc.cpp:21 float)
//   F[l]   u64   [B][V/64] frontier bitmap of every level (kept for the
//                          backward pass, one 64-bit word per wave of vertices)
//   visited u64  [B][V/64]
// Forward level l: for each (b,u) in F[l]: for each neighbour v not seen by
// b: claim v in F[l+1][b] (first claimer records depth) and add sigma[b][u]
// into sigma[b][v]. Backward level l (deepest first): for each (b,w) in
// F[l+1]: add dep[b][w] into dep[b][v] for neighbours with 0 <= depth <= l
// (the reference's "not yet processed" filter, apps/bc.cpp:193-200), then
// dep += 1/sigma over F[l]. Score: sum_b (dep - 1/sigma) * sigma where
// sigma != 0 (apps/bc.cpp:58-63), the source keeping its own term.
// Push steps extract (b,u) pairs block-locally from 64-word chunks of the
// bitmaps through one atomic chunk cursor per step (chunks differ widely in
// work; dealing them out dynamically beats a fixed round-robin). On symmetric
// graphs a
// level can instead be pulled (the reference's hybrid BC schedule, push /
// pull, apps/bc.cpp:131-147): each wavefront owns one 64-vertex word of one
// source and reads its neighbours' values, so sigma, depth and D take no
// atomics. The host picks the direction per level from the level sizes it
// reads back anyway.
#include <vector>

#include "gi_lb.h"

namespace gi {

constexpr int kBcBlock = 256;
constexpr int kBcChunkWords = 64;
constexpr int kBcChunkVerts = kBcChunkWords * 64;
constexpr int kBcMaxBatch = 256;

struct BcArgs {
    const int32_t *off;
    const int32_t *nbr;
    int32_t V;
    int32_t words;       // bitmap words per source
    int32_t B;
    int32_t level;
    int32_t *depth;
    double *sigma;
    float *dep;
    uint64_t *visited;
    const uint64_t *fin;  // F[level]
    uint64_t *fout;       // F[level+1] (forward) / unused
    int32_t *cursor;      // chunk cursor (push steps)
    unsigned long long *count;  // next-frontier size (forward)
};

// Block-local extraction of one chunk (64 words of one source) into LDS.
// Returns the number of members; wave 0 does the work.
__device__ __forceinline__ int32_t bc_extract(const uint64_t *bits, int32_t w0, int32_t words,
                                              int32_t *list, int32_t *n_s, uint64_t *fold) {
    if (threadIdx.x < 64) {
        int32_t lane = lane_id();
        int32_t w = w0 + lane;
        uint64_t word = w < words ? bits[w] : 0ull;
        if (fold && word) fold[w] |= word;  // only this thread writes word w
        int32_t pc = __popcll(word);
        int32_t incl = pc;
        for (int o = 1; o < 64; o <<= 1) {
            int32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        int32_t pos = incl - pc;
        while (word) {
            int32_t b = __ffsll((unsigned long long)word) - 1;
            word &= word - 1;
            list[pos++] = w * 64 + b;
        }
        if (lane == 63) *n_s = incl;
    }
    __syncthreads();
    return *n_s;
}

__device__ __forceinline__ void bc_forward_edge(const BcArgs &a, int32_t b, double su,
                                                int32_t v, unsigned long long &cnt) {
    const uint64_t bit = 1ull << (v & 63);
    const int64_t wi = (int64_t)b * a.words + (v >> 6);
    if ((a.visited[wi] | a.fin[wi]) & bit) return;  // seen at an earlier level
    uint64_t cur = a.fout[wi];
    if (!(cur & bit)) {
        unsigned long long old = atomicOr((unsigned long long *)&a.fout[wi], bit);
        if (!(old & bit)) {
            a.depth[(int64_t)b * a.V + v] = a.level + 1;
            cnt += 1;
        }
    }
    atomicAdd(&a.sigma[(int64_t)b * a.V + v], su);
}

__global__ void __launch_bounds__(kBcBlock) bc_forward_kernel(BcArgs a) {
    __shared__ int32_t list[kBcChunkVerts];
    __shared__ int32_t chunk_s, n_s;
    const int32_t cps = (a.words + kBcChunkWords - 1) / kBcChunkWords;  // chunks per source
    const int32_t nchunks = cps * a.B;
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    unsigned long long cnt = 0;
    for (;;) {
        if (threadIdx.x == 0) chunk_s = atomicAdd(a.cursor, 1);
        __syncthreads();
        const int32_t chunk = chunk_s;
        if (chunk >= nchunks) break;
        const int32_t b = chunk / cps;
        const int32_t w0 = (chunk % cps) * kBcChunkWords;
        const int32_t n = bc_extract(a.fin + (int64_t)b * a.words, w0, a.words, list, &n_s,
                                     a.visited + (int64_t)b * a.words);
        const double *sig = a.sigma + (int64_t)b * a.V;
        // lanes take vertices of degree < 64, waves take the wider ones
        for (int32_t i0 = 0; i0 < n; i0 += kBcBlock) {
            int32_t i = i0 + (int32_t)threadIdx.x;
            bool has = i < n;
            int32_t u = has ? list[i] : 0;
            int32_t e0 = has ? a.off[u] : 0, e1 = has ? a.off[u + 1] : 0;
            bool wide = has && e1 - e0 >= 64;
            double su = has ? sig[u] : 0.0;
            uint64_t wm = __ballot(wide);
            while (wm) {
                int32_t l = __ffsll((unsigned long long)wm) - 1;
                wm &= wm - 1;
                int32_t w0e = __shfl(e0, l), w1e = __shfl(e1, l);
                double wsu = __shfl(su, l);
                for (int32_t e = w0e + lane; e < w1e; e += 64) bc_forward_edge(a, b, wsu, a.nbr[e], cnt);
            }
            if (has && !wide)
                for (int32_t e = e0; e < e1; ++e) bc_forward_edge(a, b, su, a.nbr[e], cnt);
        }
        __syncthreads();
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (lane == 0 && cnt) atomicAdd(a.count, cnt);
    (void)wid;
}

// Backward push from F[level+1] (passed as fin) into neighbours at depth
// 0..level of the same source.
__global__ void __launch_bounds__(kBcBlock) bc_backward_kernel(BcArgs a) {
    __shared__ int32_t list[kBcChunkVerts];
    __shared__ int32_t chunk_s, n_s;
    const int32_t cps = (a.words + kBcChunkWords - 1) / kBcChunkWords;
    const int32_t nchunks = cps * a.B;
    const int32_t lane = lane_id();
    for (;;) {
        if (threadIdx.x == 0) chunk_s = atomicAdd(a.cursor, 1);
        __syncthreads();
        const int32_t chunk = chunk_s;
        if (chunk >= nchunks) break;
        const int32_t b = chunk / cps;
        const int32_t w0 = (chunk % cps) * kBcChunkWords;
        const int32_t n = bc_extract(a.fin + (int64_t)b * a.words, w0, a.words, list, &n_s, nullptr);
        const int32_t *dp = a.depth + (int64_t)b * a.V;
        float *D = a.dep + (int64_t)b * a.V;
        for (int32_t i0 = 0; i0 < n; i0 += kBcBlock) {
            int32_t i = i0 + (int32_t)threadIdx.x;
            bool has = i < n;
            int32_t w = has ? list[i] : 0;
            int32_t e0 = has ? a.off[w] : 0, e1 = has ? a.off[w + 1] : 0;
            bool wide = has && e1 - e0 >= 64;
            float dw = has ? D[w] : 0.f;
            uint64_t wm = __ballot(wide);
            while (wm) {
                int32_t l = __ffsll((unsigned long long)wm) - 1;
                wm &= wm - 1;
                int32_t w0e = __shfl(e0, l), w1e = __shfl(e1, l);
                float wd = __shfl(dw, l);
                for (int32_t e = w0e + lane; e < w1e; e += 64) {
                    int32_t v = a.nbr[e];
                    int32_t d = dp[v];
                    if (d >= 0 && d <= a.level) atomicAdd(&D[v], wd);
                }
            }
            if (has && !wide)
                for (int32_t e = e0; e < e1; ++e) {
                    int32_t v = a.nbr[e];
                    int32_t d = dp[v];
                    if (d >= 0 && d <= a.level) atomicAdd(&D[v], dw);
                }
        }
        __syncthreads();
    }
}

// Sum over the neighbours u of v with bit u set in `bits` (one source's row)
// of val[u]: lane-serial for short lists, the whole wave for long ones.
// Every lane of the wave calls it; lanes with e0 == e1 contribute nothing.
template <class T>
__device__ __forceinline__ T bc_pull_sum(const BcArgs &a, const uint64_t *bits, const T *val, int32_t e0,
                                         int32_t e1) {
    const int32_t lane = lane_id();
    T s = 0;
    const bool wide = e1 - e0 >= 64;
    if (!wide)
        for (int32_t e = e0; e < e1; ++e) {
            const int32_t u = a.nbr[e];
            if ((bits[u >> 6] >> (u & 63)) & 1ull) s += val[u];
        }
    uint64_t wm = __ballot(wide);
    while (wm) {
        const int32_t l = __ffsll((unsigned long long)wm) - 1;
        wm &= wm - 1;
        const int32_t w0 = __shfl(e0, l), w1 = __shfl(e1, l);
        T t = 0;
        for (int32_t e = w0 + lane; e < w1; e += 64) {
            const int32_t u = a.nbr[e];
            if ((bits[u >> 6] >> (u & 63)) & 1ull) t += val[u];
        }
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if (lane == l) s = t;
    }
    return s;
}

// Pull forward level: every vertex not yet seen by source b takes the sum of
// sigma over its neighbours in F[level] of b (forward_update, apps/bc.cpp:33-35,
// pulled). sigma holds integer path counts, so the order of the sum does not
// change it. F[level] joins the seen set here.
__global__ void __launch_bounds__(kBcBlock) bc_forward_pull_kernel(BcArgs a) {
    const int32_t lane = lane_id();
    const int64_t units = (int64_t)a.B * a.words;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long cnt = 0;
    for (int64_t un = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; un < units; un += waves) {
        const int32_t b = (int32_t)(un / a.words), w = (int32_t)(un % a.words);
        const int64_t wi = (int64_t)b * a.words + w;
        const uint64_t fw = a.fin[wi];
        const uint64_t seen = a.visited[wi] | fw;
        if (lane == 0 && fw) a.visited[wi] = seen;  // this wave owns the word
        const int32_t v = w * 64 + lane;
        const bool open = v < a.V && !((seen >> lane) & 1ull);
        if (!__ballot(open)) continue;
        const int32_t e0 = open ? a.off[v] : 0, e1 = open ? a.off[v + 1] : 0;
        const double s = bc_pull_sum<double>(a, a.fin + (int64_t)b * a.words, a.sigma + (int64_t)b * a.V, e0, e1);
        const bool found = s > 0.0;
        const uint64_t fm = __ballot(found);
        if (found) {
            a.sigma[(int64_t)b * a.V + v] = s;
            a.depth[(int64_t)b * a.V + v] = a.level + 1;
        }
        if (lane == 0 && fm) {
            a.fout[wi] = fm;
            cnt += (unsigned long long)__popcll(fm);
        }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (lane == 0 && cnt) atomicAdd(a.count, cnt);
}

// Pull backward level: every vertex at depth 0..level of source b adds the
// dependencies of its neighbours in F[level+1] (backward_update,
// apps/bc.cpp:52-54, pulled; the same "not yet processed" set).
__global__ void __launch_bounds__(kBcBlock) bc_backward_pull_kernel(BcArgs a) {
    const int32_t lane = lane_id();
    const int64_t units = (int64_t)a.B * a.words;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t un = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; un < units; un += waves) {
        const int32_t b = (int32_t)(un / a.words), w = (int32_t)(un % a.words);
        const int32_t v = w * 64 + lane;
        const int32_t d = v < a.V ? a.depth[(int64_t)b * a.V + v] : -1;
        const bool open = d >= 0 && d <= a.level;
        if (!__ballot(open)) continue;
        const int32_t e0 = open ? a.off[v] : 0, e1 = open ? a.off[v + 1] : 0;
        float *D = a.dep + (int64_t)b * a.V;
        const float acc = bc_pull_sum<float>(a, a.fin + (int64_t)b * a.words, D, e0, e1);
        if (open && acc != 0.f) D[v] += acc;
    }
}

// dep += 1/sigma for every (b, v) of F[level] (apps/bc.cpp:49-52).
__global__ void bc_vertex_kernel(const uint64_t *f, int32_t words, int32_t B, int32_t V,
                                 const double *sigma, float *dep) {
    int64_t total = (int64_t)words * B;
    for (int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; wi < total;
         wi += (int64_t)gridDim.x * blockDim.x) {
        uint64_t word = f[wi];
        if (!word) continue;
        int32_t b = (int32_t)(wi / words);
        int32_t base = (int32_t)(wi % words) * 64;
        while (word) {
            int32_t bit = __ffsll((unsigned long long)word) - 1;
            word &= word - 1;
            int64_t k = (int64_t)b * V + base + bit;
            dep[k] = (float)((double)dep[k] + 1.0 / sigma[k]);
        }
    }
}

// scores[v] += sum_b (dep - 1/sigma) * sigma over sigma != 0 (apps/bc.cpp:58-63).
__global__ void bc_final_kernel(const double *sigma, const float *dep, int32_t B, int32_t V,
                                float *scores) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int32_t b = 0; b < B; ++b) {
            int64_t k = (int64_t)b * V + v;
            double s = sigma[k];
            if (s != 0.0) acc += ((double)dep[k] - 1.0 / s) * s;
        }
        scores[v] += (float)acc;
    }
}

__global__ void bc_seed_kernel_b(const int32_t *src, int32_t B, int32_t V, int32_t words,
                                 int32_t *depth, double *sigma, uint64_t *f0) {
    int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int32_t s = src[b];
    depth[(int64_t)b * V + s] = 0;
    sigma[(int64_t)b * V + s] = 1.0;
    f0[(int64_t)b * words + (s >> 6)] |= 1ull << (s & 63);  // one source per lane-private row
}

static int grid_for(int64_t n, int block = 256) {
    int64_t want = ceil_div(n, block);
    int cap = persistent_grid(block);
    return (int)(want < 1 ? 1 : (want < cap ? want : cap));
}

// One batch of B sources.
int bc_batch(gi_graph *g, const int32_t *h_src, int32_t B, float *d_scores, hipStream_t s,
             int32_t *levels_out) {
    const int32_t V = g->V;
    const int32_t words = (int32_t)bitmap_words(V);
    const int64_t BV = (int64_t)B * V, BW = (int64_t)B * words;
    int32_t *depth = nullptr, *d_src = nullptr, *cursor = nullptr;
    double *sigma = nullptr;
    float *dep = nullptr;
    uint64_t *visited = nullptr;
    unsigned long long *count = nullptr;
    std::vector<uint64_t *> F;
    int rc = dalloc(&depth, (size_t)BV);
    if (rc == GI_OK) rc = dalloc(&sigma, (size_t)BV);
    if (rc == GI_OK) rc = dalloc(&dep, (size_t)BV);
    if (rc == GI_OK) rc = dalloc(&visited, (size_t)BW);
    if (rc == GI_OK) rc = dalloc(&d_src, (size_t)B);
    if (rc == GI_OK) rc = dalloc(&cursor, 1);
    if (rc == GI_OK) rc = dalloc(&count, 1);
    auto new_level = [&](uint64_t **out) -> int {
        uint64_t *p = nullptr;
        GI_TRY(dalloc(&p, (size_t)BW));
        F.push_back(p);
        *out = p;
        return fill_u64(p, BW, 0, s);
    };
    uint64_t *f0 = nullptr;
    if (rc == GI_OK) rc = new_level(&f0);
    if (rc == GI_OK) rc = fill_i32(depth, BV, -1, s);
    if (rc == GI_OK) rc = fill_f64(sigma, BV, 0.0, s);
    if (rc == GI_OK) rc = fill_f32(dep, BV, 0.f, s);
    if (rc == GI_OK) rc = fill_u64(visited, BW, 0, s);
    if (rc == GI_OK && hipMemcpyAsync(d_src, h_src, sizeof(int32_t) * B, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = fail(GI_ERR_HIP, "copy sources");
    // one kernel per source would race on shared words only if two sources
    // shared a row; rows are per source, so the seed kernel is race free
    if (rc == GI_OK) bc_seed_kernel_b<<<(B + 63) / 64, 64, 0, s>>>(d_src, B, V, words, depth, sigma, f0);
    BcArgs a{};
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.V = V;
    a.words = words;
    a.B = B;
    a.depth = depth;
    a.sigma = sigma;
    a.dep = dep;
    a.visited = visited;
    a.cursor = cursor;
    a.count = count;
    int grid = persistent_grid(kBcBlock, 8);
    // level sizes summed over the batch decide each level's direction: pull a
    // forward level whose frontier holds more than V/16 vertices per source,
    // and a backward level whose receiving levels are under 1/64 of the level
    // that sends; push otherwise, and on directed graphs (measured on RMAT-20,
    // 256 sources: 285 ms all pushed, 218 ms with these rules)
    const bool can_pull = !g->directed;
    std::vector<int64_t> sizes(1, B);
    int64_t seen = B;
    int32_t level = 0;
    while (rc == GI_OK) {  // forward
        uint64_t *fn = nullptr;
        rc = new_level(&fn);
        if (rc != GI_OK) break;
        a.level = level;
        a.fin = F[level];
        a.fout = fn;
        rc = fill_i32(cursor, 1, 0, s);
        if (rc == GI_OK) rc = fill_u64((uint64_t *)count, 1, 0, s);
        if (rc != GI_OK) break;
        if (can_pull && 16 * sizes[level] > (int64_t)B * V) bc_forward_pull_kernel<<<grid, kBcBlock, 0, s>>>(a);
        else bc_forward_kernel<<<grid, kBcBlock, 0, s>>>(a);
        unsigned long long n = 0;
        if (hipMemcpyAsync(&n, count, sizeof(n), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(GI_ERR_HIP, "bc forward");
            break;
        }
        ++level;
        sizes.push_back((int64_t)n);
        seen += (int64_t)n;
        if (n == 0) break;
    }
    // F[0..level-1] hold the levels; F[level] is empty
    const int32_t L = level;
    if (rc == GI_OK && L > 0)
        bc_vertex_kernel<<<grid_for(BW), 256, 0, s>>>(F[L - 1], words, B, V, sigma, dep);
    std::vector<int64_t> upto(sizes.size() + 1, 0);  // upto[l + 1] = sizes of levels 0..l
    for (size_t i = 0; i < sizes.size(); ++i) upto[i + 1] = upto[i] + sizes[i];
    for (int32_t l = L - 2; l >= 0 && rc == GI_OK; --l) {  // backward
        a.level = l;
        a.fin = F[l + 1];
        rc = fill_i32(cursor, 1, 0, s);
        if (rc != GI_OK) break;
        if (can_pull && 64 * upto[l + 1] < sizes[l + 1]) bc_backward_pull_kernel<<<grid, kBcBlock, 0, s>>>(a);
        else bc_backward_kernel<<<grid, kBcBlock, 0, s>>>(a);
        bc_vertex_kernel<<<grid_for(BW), 256, 0, s>>>(F[l], words, B, V, sigma, dep);
        if (hipGetLastError() != hipSuccess) rc = fail(GI_ERR_HIP, "bc backward launch");
    }
    if (rc == GI_OK) bc_final_kernel<<<grid_for(V), 256, 0, s>>>(sigma, dep, B, V, d_scores);
    if (rc == GI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(GI_ERR_HIP, "bc sync");
    for (uint64_t *p : F) dfree(p);
    dfree(depth);
    dfree(sigma);
    dfree(dep);
    dfree(visited);
    dfree(d_src);
    dfree(cursor);
    dfree(count);
    if (levels_out) *levels_out = L;
    return rc;
}

}  // namespace gi
