// Batched Brandes betweenness: up to 256 sources advance level-synchronously
// in one set of launches per level (apps/bc.cpp:64-125 computes one source
// per run, with a host round trip per level and per list operation). With
// one source per batch this is the fused single-source mode (fuse_kernel,
// apps/bc.cpp:92-113): the level loops run on the device.
//
// Layout (source-major, B sources, V vertices):
//   depth  int32 [B][V]   level of v from source b (-1 unreached)
//   sigma  f64   [B][V]   shortest-path counts (apps/bc.cpp:20 uses double)
//   dep    f32   [B][V]   D(v) of the backward recurrence (apps/bc.cpp:21 float)
//   bm[3]  u64   [B][V/64] rotating frontier bitmaps: F[l] = bm[l % 3]
//   snd    u64   [B][V/64] senders of a backward step (level l + 1)
//   visited u64  [B][V/64] levels before the current one
// Forward level l: for each (b,u) in F[l]: for each neighbour v not seen by
// b: claim v in F[l+1][b] (first claimer records depth) and add sigma[b][u]
// into sigma[b][v]; F[l-1] is cleared meanwhile for reuse as F[l+2].
// Backward level l (deepest first): dep += 1/sigma over level l + 1, whose
// bitmap is rebuilt from depth (no bitmap per level is kept, so memory does
// not grow with the number of levels); then for each (b,w) at level l + 1:
// add dep[b][w] into dep[b][v] for neighbours with 0 <= depth <= l (the
// reference's "not yet processed" filter, apps/bc.cpp:108-114); a last pass
// adds 1/sigma over level 0. Score: sum_b (dep - 1/sigma) * sigma where
// sigma != 0 (apps/bc.cpp:58-63), the source keeping its own term.
// Push steps extract (b,u) pairs block-locally from 64-word chunks of the
// bitmaps through one atomic chunk cursor per step (chunks differ widely in
// work; dealing them out dynamically beats a fixed round-robin). On symmetric
// graphs a level can instead be pulled (the reference's hybrid BC schedule,
// push / pull, apps/bc.cpp:131-147): each wavefront owns one 64-vertex word
// of one source and reads its neighbours' values, so sigma, depth and D take
// no atomics.
// Control: the state (level, direction, counts) lives on the device; a
// forward slot is a step and a decide launch, a backward slot a level, a step
// and a decide launch; slots are recorded as hipGraphs and replayed in
// blocks, the host reading the state once per block. Direction rules, from
// level sizes summed over the batch: pull a forward level whose frontier
// holds more than V/16 vertices per source, and a backward level whose
// lower levels together hold under 1/64 of the vertices of the level that
// sends (fewer than it, for a single source); push otherwise, and on
// directed graphs (RMAT-20, 256 sources: 285 ms all pushed, 218 ms with
// these rules).
#include <map>
#include <tuple>
#include <vector>

#include "gi_lb.h"

namespace gi {

constexpr int kBcBlock = 256;
constexpr int kBcChunkWords = 64;
constexpr int kBcChunkVerts = kBcChunkWords * 64;
constexpr int kBcMaxBatch = 256;

enum BcPhase : int32_t { BC_FWD = 0, BC_BWD = 1, BC_DONE = 2 };

struct BcState {
    int32_t phase;
    int32_t level;       // forward: the level expanded; backward: l (senders at l + 1)
    int32_t levels;      // L, set when the forward pass ends
    int32_t pull;        // forward: the next step pulls
    int32_t cursor;      // chunk cursor of push steps
    int32_t pad;
    unsigned long long count;  // forward: |F[l+1]|; backward: |level l+1|
    long long remaining;       // vertices at levels not yet sent from (backward)
};

struct BcArgs {
    const int32_t *off;
    const int32_t *nbr;
    int32_t V;
    int32_t words;       // bitmap words per source
    int32_t B;
    int32_t can_pull;    // symmetric graph: in-neighbours are the neighbours
    BcState *st;
    int32_t *depth;
    double *sigma;
    float *dep;
    uint64_t *visited;
    uint64_t *bm[3];
    uint64_t *snd;
    int32_t *cursor;            // &st->cursor
    unsigned long long *count;  // &st->count
    // bound by each step from the state
    int32_t level;
    const uint64_t *fin;  // forward: F[level]; backward: the senders
    uint64_t *fout;       // forward: F[level+1]
};

// Pointers passed by value: selecting through a reference to the kernel's
// argument copy puts the whole copy in scratch memory.
__device__ __forceinline__ uint64_t *bc_pick3(uint64_t *b0, uint64_t *b1, uint64_t *b2, int32_t i) {
    return i == 0 ? b0 : (i == 1 ? b1 : b2);
}

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

__device__ __forceinline__ void bc_forward_push(const BcArgs &a, int32_t *list) {
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

// Backward push from the senders (fin, level + 1) into neighbours at depth
// 0..level of the same source.
__device__ __forceinline__ void bc_backward_push(const BcArgs &a, int32_t *list) {
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
__device__ __forceinline__ void bc_forward_pull(const BcArgs &a) {
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

// Pull backward level (symmetric graphs only): every vertex at depth `level`
// of source b adds the dependencies of its neighbours among the senders
// (backward_update, apps/bc.cpp:52-54, pulled). The reference's receivers
// are all vertices at depth 0..level ("not yet processed"), but on a
// symmetric graph a neighbour of a vertex at level + 1 lies at level,
// level + 1 or level + 2, so only depth == level can receive.
__device__ __forceinline__ void bc_backward_pull(const BcArgs &a) {
    const int32_t lane = lane_id();
    const int64_t units = (int64_t)a.B * a.words;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t un = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; un < units; un += waves) {
        const int32_t b = (int32_t)(un / a.words), w = (int32_t)(un % a.words);
        const int32_t v = w * 64 + lane;
        const int32_t d = v < a.V ? a.depth[(int64_t)b * a.V + v] : -1;
        const bool open = d == a.level;
        if (!__ballot(open)) continue;
        const int32_t e0 = open ? a.off[v] : 0, e1 = open ? a.off[v + 1] : 0;
        float *D = a.dep + (int64_t)b * a.V;
        const float acc = bc_pull_sum<float>(a, a.fin + (int64_t)b * a.words, D, e0, e1);
        if (open && acc != 0.f) D[v] += acc;
    }
}

// Forward step of the state's level l: push from F[l] into F[l+1], or pull
// (st->pull, set by the decide launch); F[l-1], already folded into visited,
// is cleared for reuse as F[l+2].
__global__ void __launch_bounds__(kBcBlock) bc_fwd_kernel(BcArgs a) {
    BcState *st = a.st;
    if (st->phase != BC_FWD) return;  // uniform
    __shared__ int32_t list[kBcChunkVerts];
    const int32_t l = st->level;
    a.level = l;
    a.fin = bc_pick3(a.bm[0], a.bm[1], a.bm[2], l % 3);
    a.fout = bc_pick3(a.bm[0], a.bm[1], a.bm[2], (l + 1) % 3);
    uint64_t *clr = bc_pick3(a.bm[0], a.bm[1], a.bm[2], (l + 2) % 3);
    const int64_t n = (int64_t)a.B * a.words;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        clr[i] = 0;
    if (st->pull)
        bc_forward_pull(a);
    else
        bc_forward_push(a, list);
}

// After a forward step: an empty F[l+1] ends the pass with L = l + 1 levels
// and starts the backward pass at l = L - 2; otherwise the next level, and
// its direction.
__global__ void bc_fwd_decide_kernel(BcArgs a) {
    BcState *st = a.st;
    if (st->phase != BC_FWD || threadIdx.x != 0) return;
    const unsigned long long n = st->count;
    st->count = 0;
    st->cursor = 0;
    if (n == 0) {
        st->levels = st->level + 1;
        st->level = st->levels - 2;
        st->phase = BC_BWD;
    } else {
        st->remaining += (long long)n;
        st->level += 1;
        st->pull = a.can_pull && 16ull * n > (unsigned long long)a.B * (unsigned long long)a.V;
    }
}

// Backward bookkeeping for the state's l: dep += 1/sigma over level l + 1
// (backward_vertex_f, apps/bc.cpp:49-52), and for l >= 0 the bitmap of level
// l + 1 (the senders of the step) with its size. One wave per word of one
// source; every sender word is written, so snd needs no clearing.
__global__ void __launch_bounds__(kBcBlock) bc_level_kernel(BcArgs a) {
    BcState *st = a.st;
    if (st->phase != BC_BWD) return;  // uniform
    const int32_t l = st->level;
    const int32_t lane = lane_id();
    const int64_t units = (int64_t)a.B * a.words;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long cnt = 0;
    for (int64_t un = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; un < units; un += waves) {
        const int32_t b = (int32_t)(un / a.words), w = (int32_t)(un % a.words);
        const int32_t v = w * 64 + lane;
        const int64_t k = (int64_t)b * a.V + v;
        const bool in = v < a.V && a.depth[k] == l + 1;
        if (in) a.dep[k] = (float)((double)a.dep[k] + 1.0 / a.sigma[k]);
        const uint64_t m = __ballot(in);
        if (l >= 0 && lane == 0) {
            a.snd[un] = m;
            cnt += (unsigned long long)__popcll(m);
        }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (lane == 0 && cnt) atomicAdd(a.count, cnt);
}

// Backward step of the state's l >= 0 from the senders of level l + 1.
__global__ void __launch_bounds__(kBcBlock) bc_bwd_kernel(BcArgs a) {
    BcState *st = a.st;
    if (st->phase != BC_BWD || st->level < 0) return;  // uniform
    __shared__ int32_t list[kBcChunkVerts];
    a.level = st->level;
    a.fin = a.snd;
    const long long c = (long long)st->count;  // senders
    const long long below = st->remaining - c;   // vertices at levels 0..l (>= the receivers)
    // one source: its sender bitmap (V/8 bytes) stays in L2 and pulling pays
    // once the lower levels are smaller than the senders; a batch's bitmaps
    // and values leave L2, and pulling pays only for far smaller receivers
    // (RMAT-20: one source 859 -> 820 ms per 256; a batch of 256 223 ms with
    // the 1/64 rule, 389 ms with the other)
    if (a.can_pull && (a.B == 1 ? below < c : 64 * below < c))
        bc_backward_pull(a);
    else
        bc_backward_push(a, list);
}

__global__ void bc_bwd_decide_kernel(BcArgs a) {
    BcState *st = a.st;
    if (st->phase != BC_BWD || threadIdx.x != 0) return;
    st->remaining -= (long long)st->count;
    st->count = 0;
    st->cursor = 0;
    st->level -= 1;
    if (st->level < -1) st->phase = BC_DONE;
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

// Fresh arrays and state for a batch of B sources (reset, apps/bc.cpp:24-27).
__global__ void bc_reset_kernel(BcArgs a) {
    const int64_t BV = (int64_t)a.B * a.V, BW = (int64_t)a.B * a.words;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < BV; i += (int64_t)gridDim.x * blockDim.x) {
        a.depth[i] = -1;
        a.sigma[i] = 0.0;
        a.dep[i] = 0.f;
        if (i < BW) {
            a.visited[i] = 0;
            a.bm[0][i] = 0;
            a.bm[1][i] = 0;
            a.bm[2][i] = 0;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        BcState *st = a.st;
        st->phase = BC_FWD;
        st->level = 0;
        st->levels = 0;
        st->pull = 0;
        st->cursor = 0;
        st->count = 0;
        st->remaining = a.B;  // the sources, level 0
    }
}

// Level 0: each source in its own row (one thread per source).
__global__ void bc_seed_kernel_b(const int32_t *src, int32_t B, int32_t V, int32_t words,
                                 int32_t *depth, double *sigma, uint64_t *f0) {
    int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    int32_t s = src[b];
    depth[(int64_t)b * V + s] = 0;
    sigma[(int64_t)b * V + s] = 1.0;
    f0[(int64_t)b * words + (s >> 6)] |= 1ull << (s & 63);
}

static int grid_for(int64_t n, int block = 256) {
    int64_t want = ceil_div(n, block);
    int cap = persistent_grid(block);
    return (int)(want < 1 ? 1 : (want < cap ? want : cap));
}

// Arrays, state and recorded slot blocks of one BC call (batches of up to
// `cap` sources); freed when the call ends.
struct BcRun {
    int32_t cap = 0;
    int32_t *depth = nullptr, *d_src = nullptr;
    double *sigma = nullptr;
    float *dep = nullptr;
    uint64_t *visited = nullptr, *snd = nullptr;
    uint64_t *bm[3] = {nullptr, nullptr, nullptr};
    BcState *st = nullptr;
    BcState *host = nullptr;  // pinned
    std::map<std::tuple<int32_t, int32_t, int32_t>, hipGraphExec_t> execs;  // (phase, slots, B)
    int32_t recent[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // levels of the last batches (ring)
    int32_t recent_n = 0;
    int32_t hint() const {  // the deepest of them (8 before the first)
        int32_t m = 0;
        for (int32_t l : recent) m = l > m ? l : m;
        return m > 0 ? m : 8;
    }
    ~BcRun() {
        for (auto &kv : execs) (void)hipGraphExecDestroy(kv.second);
        dfree(depth);
        dfree(d_src);
        dfree(sigma);
        dfree(dep);
        dfree(visited);
        dfree(snd);
        for (auto p : bm) dfree(p);
        dfree(st);
        if (host) (void)hipHostFree(host);
    }
    int init(int32_t V, int32_t B) {
        cap = B;
        const int64_t BV = (int64_t)B * V, BW = (int64_t)B * bitmap_words(V);
        GI_TRY(dalloc(&depth, (size_t)BV));
        GI_TRY(dalloc(&sigma, (size_t)BV));
        GI_TRY(dalloc(&dep, (size_t)BV));
        GI_TRY(dalloc(&visited, (size_t)BW));
        GI_TRY(dalloc(&snd, (size_t)BW));
        for (auto &p : bm) GI_TRY(dalloc(&p, (size_t)BW));
        GI_TRY(dalloc(&d_src, (size_t)B));
        GI_TRY(dalloc(&st, 1));
        if (hipHostMalloc((void **)&host, sizeof(BcState)) != hipSuccess) return fail(GI_ERR_OOM, "pinned allocation failed");
        return GI_OK;
    }
};

static void bc_launch_slot(const BcArgs &a, int32_t phase, Launcher &L) {
    const int grid = persistent_grid(kBcBlock, 8);
    if (phase == BC_FWD) {
        L(bc_fwd_kernel, grid, kBcBlock, a);
        L(bc_fwd_decide_kernel, 1, 64, a);
    } else {
        L(bc_level_kernel, grid, kBcBlock, a);
        L(bc_bwd_kernel, grid, kBcBlock, a);
        L(bc_bwd_decide_kernel, 1, 64, a);
    }
}

// Replays blocks of recorded slots of one pass until the state leaves it:
// first `first` slots (rounded up to 4), then 4 at a time (a slot past the
// end costs its no-op launches, a block too short a host round trip).
static int bc_pass(BcRun &r, const BcArgs &a, int32_t phase, int32_t first, hipStream_t s) {
    int32_t block = first < 4 ? 4 : (first + 3) / 4 * 4;
    for (int64_t used = 0; used < (1ll << 24);) {
        if (block > 512) block = 512;
        const auto key = std::make_tuple(phase, block, a.B);
        auto it = r.execs.find(key);
        if (it == r.execs.end()) {
            hipGraphExec_t exec = nullptr;
            GI_TRY(build_graph(
                [&](Launcher &L) {
                    for (int32_t i = 0; i < block; ++i) bc_launch_slot(a, phase, L);
                },
                &exec));
            it = r.execs.emplace(key, exec).first;
        }
        GI_HIP(hipGraphLaunch(it->second, s));
        GI_HIP(hipMemcpyAsync(r.host, r.st, sizeof(BcState), hipMemcpyDeviceToHost, s));
        GI_HIP(hipStreamSynchronize(s));
        used += block;
        if (r.host->phase != phase) return GI_OK;
        block = 4;
    }
    return fail(GI_ERR_RANGE, "BC did not finish");
}

// One batch of B <= r.cap sources: scores += their BC terms; *levels = L.
static int bc_batch_run(gi_graph *g, BcRun &r, const int32_t *h_src, int32_t B, float *d_scores, hipStream_t s,
                        int32_t *levels) {
    BcArgs a{};
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.V = g->V;
    a.words = (int32_t)bitmap_words(g->V);
    a.B = B;
    a.can_pull = !g->directed;
    a.st = r.st;
    a.depth = r.depth;
    a.sigma = r.sigma;
    a.dep = r.dep;
    a.visited = r.visited;
    for (int i = 0; i < 3; ++i) a.bm[i] = r.bm[i];
    a.snd = r.snd;
    a.cursor = &r.st->cursor;
    a.count = &r.st->count;
    GI_HIP(hipMemcpyAsync(r.d_src, h_src, sizeof(int32_t) * B, hipMemcpyHostToDevice, s));
    bc_reset_kernel<<<grid_for((int64_t)B * g->V), 256, 0, s>>>(a);
    bc_seed_kernel_b<<<(B + 63) / 64, 64, 0, s>>>(r.d_src, B, g->V, a.words, r.depth, r.sigma, r.bm[0]);
    GI_LAUNCH_CHECK();
    GI_TRY(bc_pass(r, a, BC_FWD, r.hint(), s));
    const int32_t L = r.host->levels;
    GI_TRY(bc_pass(r, a, BC_BWD, L, s));  // l = L-2 .. -1: L slots
    bc_final_kernel<<<grid_for(g->V), 256, 0, s>>>(r.sigma, r.dep, B, g->V, d_scores);
    GI_LAUNCH_CHECK();
    r.recent[r.recent_n++ % 8] = L;
    if (levels) *levels = L;
    return GI_OK;
}

// Sources in batches of `batch` (1: the fused single-source mode); scores are
// added into d_scores. *levels: the levels summed over the batches.
int bc_batched(gi_graph *g, const int32_t *h_src, int32_t num_sources, int32_t batch, float *d_scores,
               hipStream_t s, int64_t *levels) {
    if (batch > 1) return bc_vertex_major(g, h_src, num_sources, batch, d_scores, s, levels);
    if (batch < 1) batch = 1;
    if (batch > kBcMaxBatch) batch = kBcMaxBatch;
    if (batch > num_sources) batch = num_sources;
    if (levels) *levels = 0;
    if (num_sources <= 0) return GI_OK;
    BcRun r;
    GI_TRY(r.init(g->V, batch));
    for (int32_t i = 0; i < num_sources; i += batch) {
        const int32_t nb = num_sources - i < batch ? num_sources - i : batch;
        int32_t lv = 0;
        GI_TRY(bc_batch_run(g, r, h_src + i, nb, d_scores, s, &lv));
        if (levels) *levels += lv;
    }
    GI_HIP(hipStreamSynchronize(s));
    return GI_OK;
}

}  // namespace gi
