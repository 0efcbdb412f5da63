// Batched delta-stepping SSSP: the distances from each of many sources
// (dist[source][vertex]) with one set of launches per step for a batch of up
// to 256 sources. The SSSP app (apps/sssp.cpp:34-57) runs one source per
// call: dequeue the window [lo, lo + delta) of the priority queue, relax its
// out-edges with writeMin, re-enqueue what improved
// (gpu_priority_queue.h:115-185). Here every source keeps its own window and
// queue state; a step serves all of them:
//
//   relax   for each (b, u) in F[b] (source-major bitmaps, listed per chunk
//           of 64 words through a chunk cursor): writeMin dist[b][v] to
//           dist[b][u] + w; an improvement below hi[b] joins F'[b], one at or
//           beyond it joins the pending set P[b]
//   decide  a source whose F' is empty advances: its window is settled
//   pmin    for advancing sources, the least pending distance (P members
//           below the settled window end are stale and dropped)
//   window  the next window starts at that distance (the reference's queue
//           starts a window at the least remaining priority too,
//           gpu_priority_queue.h:126-146); none pending: the source is done
//   move    P members inside the new window join F
//
// Shortest-path distances are unique, so the order of relaxations does not
// change them. A step is five launches recorded in hipGraph blocks (every
// kernel reads the sources' modes from device state); the host reads the
// batch's done flag once per replayed block.
#include <climits>
#include <vector>

#include "gi_internal.h"

namespace gi {

constexpr int kSbBlock = 256;
constexpr int kSbChunkWords = 64;
constexpr int kSbMaxBatch = 256;
constexpr int kSbBlockSteps = 16;

enum SbMode : int32_t { SB_RELAX = 0, SB_ADVANCE = 1, SB_MOVE = 2, SB_DONE = 3 };

struct SbState {
    int32_t done;     // every source done
    int32_t step;     // F = bm[step & 1]
    int32_t cursor;   // chunk cursor of the relax step
    int32_t B;
    int32_t delta;
    int32_t pad;
    long long steps;  // relax steps summed over the sources
    int32_t mode[kSbMaxBatch];
    int32_t lo[kSbMaxBatch];
    int32_t hi[kSbMaxBatch];
    int32_t count[kSbMaxBatch];  // |F'| of the step (or members moved)
    int32_t pmin[kSbMaxBatch];   // least pending distance (advance)
    int32_t *dist;               // [B][V] of the batch
};

struct SbArgs {
    SbState *st;
    const int32_t *off, *nbr, *wt;
    int32_t V;
    int32_t words;  // bitmap words per source
    uint64_t *bm[2];
    uint64_t *pend;
};

__device__ __forceinline__ uint64_t *sb_pick(uint64_t *a, uint64_t *b, int32_t i) { return (i & 1) ? b : a; }

// writeMin of dist[b][v] to nd; an improvement joins F' (below hi) or P
__device__ __forceinline__ void sb_relax(int32_t *dist, uint64_t *fo, uint64_t *po, int32_t hi,
                                         int32_t v, int32_t nd, int32_t &cnt) {
    if (nd >= dist[v]) return;  // test before the atomic
    const int32_t old = atomicMin(&dist[v], nd);
    if (nd >= old) return;
    const uint64_t bit = 1ull << (v & 63);
    if (nd < hi) {
        const unsigned long long o = atomicOr((unsigned long long *)&fo[v >> 6], bit);
        if (!(o & bit)) cnt += 1;
    } else if (!(po[v >> 6] & bit)) {
        atomicOr((unsigned long long *)&po[v >> 6], bit);
    }
}

__global__ void __launch_bounds__(kSbBlock) sb_relax_kernel(SbArgs a) {
    SbState *st = a.st;
    if (st->done) return;  // uniform
    const int32_t step = st->step, B = st->B;
    uint64_t *fin = sb_pick(a.bm[0], a.bm[1], step);
    uint64_t *fout = sb_pick(a.bm[0], a.bm[1], step + 1);
    int32_t *dist_all = st->dist;
    const int32_t cps = (a.words + kSbChunkWords - 1) / kSbChunkWords;
    const int32_t nchunks = cps * B;
    const int32_t lane = lane_id();
    __shared__ int32_t chunk_s, n_s;
    __shared__ int32_t list[kSbChunkWords * 64];
    for (;;) {
        if (threadIdx.x == 0) chunk_s = atomicAdd(&st->cursor, 1);
        __syncthreads();
        const int32_t chunk = chunk_s;
        if (chunk >= nchunks) break;
        const int32_t b = chunk / cps;
        const int32_t w0 = (chunk % cps) * kSbChunkWords;
        if (st->mode[b] != SB_RELAX) {  // block-uniform
            __syncthreads();
            continue;
        }
        uint64_t *fb = fin + (int64_t)b * a.words;
        if (threadIdx.x < 64) {  // wave 0 lists the chunk's members and clears them
            const int32_t w = w0 + lane;
            uint64_t word = w < a.words ? fb[w] : 0ull;
            if (word) fb[w] = 0;  // only this lane reads word w of F in this step
            const int32_t pc = __popcll(word);
            int32_t incl = pc;
            for (int o = 1; o < 64; o <<= 1) {
                const int32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            int32_t pos = incl - pc;
            while (word) {
                const int32_t bit = __ffsll((unsigned long long)word) - 1;
                word &= word - 1;
                list[pos++] = w * 64 + bit;
            }
            if (lane == 63) n_s = incl;
        }
        __syncthreads();
        const int32_t n = n_s;
        int32_t *dist = dist_all + (int64_t)b * a.V;
        uint64_t *fo = fout + (int64_t)b * a.words;
        uint64_t *po = a.pend + (int64_t)b * a.words;
        const int32_t hi = st->hi[b];
        int32_t bcnt = 0;
        for (int32_t i0 = 0; i0 < n; i0 += kSbBlock) {
            const int32_t i = i0 + (int32_t)threadIdx.x;
            const bool has = i < n;
            const int32_t u = has ? list[i] : 0;
            const int32_t e0 = has ? a.off[u] : 0, e1 = has ? a.off[u + 1] : 0;
            const int32_t du = has ? dist[u] : 0;
            const bool wide = has && e1 - e0 >= 64;
            uint64_t wm = __ballot(wide);
            while (wm) {  // wide lists: the whole wave
                const int32_t j = __ffsll((unsigned long long)wm) - 1;
                wm &= wm - 1;
                const int32_t we0 = __shfl(e0, j), we1 = __shfl(e1, j), wdu = __shfl(du, j);
                for (int32_t e = we0 + lane; e < we1; e += 64) sb_relax(dist, fo, po, hi, a.nbr[e], wdu + a.wt[e], bcnt);
            }
            if (has && !wide)
                for (int32_t e = e0; e < e1; ++e) sb_relax(dist, fo, po, hi, a.nbr[e], du + a.wt[e], bcnt);
        }
        for (int o = 32; o > 0; o >>= 1) bcnt += __shfl_xor(bcnt, o);
        if (lane == 0 && bcnt) atomicAdd(&st->count[b], bcnt);
        __syncthreads();
    }
}

// One thread per source: settle, advance, or keep relaxing.
__global__ void sb_decide_kernel(SbState *st) {
    const int32_t b = threadIdx.x;
    if (st->done) return;
    if (b < st->B && st->mode[b] == SB_RELAX) {
        if (st->count[b] == 0) {
            st->mode[b] = SB_ADVANCE;
            st->pmin[b] = INT_MAX;
        }
        st->count[b] = 0;
    }
    __syncthreads();
    if (b == 0) {
        int32_t relaxing = 0;
        for (int32_t i = 0; i < st->B; ++i) relaxing += st->mode[i] == SB_RELAX;
        st->steps += relaxing;
        st->step += 1;
        st->cursor = 0;
    }
}

// Least pending distance of each advancing source; stale members (settled
// below the window just finished) leave P.
__global__ void __launch_bounds__(kSbBlock) sb_pmin_kernel(SbArgs a) {
    SbState *st = a.st;
    if (st->done) return;
    const int32_t B = st->B, lane = lane_id();
    const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
    // a wave covers 64 consecutive words of one source (words per source is
    // padded to a multiple of 64 in the grid stride below)
    const int64_t padded = (int64_t)B * ((a.words + 63) / 64 * 64);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < padded; t += nthreads) {
        const int32_t wp = (a.words + 63) / 64 * 64;
        const int32_t b = (int32_t)(t / wp), w = (int32_t)(t % wp);
        int32_t m = INT_MAX;
        if (st->mode[b] == SB_ADVANCE && w < a.words) {  // uniform over the wave
            uint64_t *pw = a.pend + (int64_t)b * a.words + w;
            uint64_t word = *pw, keep = word;
            const int32_t hi = st->hi[b];
            const int32_t *dist = st->dist + (int64_t)b * a.V;
            while (word) {
                const int32_t bit = __ffsll((unsigned long long)word) - 1;
                word &= word - 1;
                const int32_t d = dist[w * 64 + bit];
                if (d < hi) keep &= ~(1ull << bit);  // settled
                else m = d < m ? d : m;
            }
            if (keep != *pw) *pw = keep;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const int32_t y = __shfl_xor(m, o);
            m = y < m ? y : m;
        }
        if (lane == 0 && m != INT_MAX) atomicMin(&st->pmin[b], m);
    }
}

__global__ void sb_window_kernel(SbState *st) {
    const int32_t b = threadIdx.x;
    if (st->done) return;
    if (b < st->B && st->mode[b] == SB_ADVANCE) {
        const int32_t m = st->pmin[b];
        if (m == INT_MAX) {
            st->mode[b] = SB_DONE;
        } else {
            st->lo[b] = m;
            st->hi[b] = m > INT_MAX - st->delta ? INT_MAX : m + st->delta;
            st->mode[b] = SB_MOVE;
        }
    }
}

// P members inside the new window of a moving source join F.
__global__ void __launch_bounds__(kSbBlock) sb_move_kernel(SbArgs a) {
    SbState *st = a.st;
    if (st->done) return;
    const int32_t B = st->B;
    uint64_t *fin = sb_pick(a.bm[0], a.bm[1], st->step);
    const int64_t total = (int64_t)B * a.words;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = (int32_t)(t / a.words), w = (int32_t)(t % a.words);
        if (st->mode[b] != SB_MOVE) continue;
        uint64_t *pw = a.pend + t;
        uint64_t word = *pw;
        if (!word) continue;
        const int32_t hi = st->hi[b];
        const int32_t *dist = st->dist + (int64_t)b * a.V;
        uint64_t in = 0, x = word;
        while (x) {
            const int32_t bit = __ffsll((unsigned long long)x) - 1;
            x &= x - 1;
            if (dist[w * 64 + bit] < hi) in |= 1ull << bit;
        }
        if (in) {
            pw[0] = word & ~in;
            fin[t] |= in;  // this thread owns word w of source b
        }
    }
}

__global__ void sb_resume_kernel(SbState *st) {
    const int32_t b = threadIdx.x;
    if (st->done) return;
    if (b < st->B && st->mode[b] == SB_MOVE) st->mode[b] = SB_RELAX;
    __syncthreads();
    if (b == 0) {
        int32_t live = 0;
        for (int32_t i = 0; i < st->B; ++i) live += st->mode[i] != SB_DONE;
        if (!live) st->done = 1;
    }
}

struct SbSeed {
    int32_t *dist;
    int32_t B;
    int32_t delta;
};

__global__ void sb_reset_kernel(SbArgs a, SbSeed sd) {
    const int64_t nd = (int64_t)sd.B * a.V, nw = (int64_t)sd.B * a.words;
    const int64_t n = nd > nw ? nd : nw;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nd) sd.dist[i] = INT_MAX;
        if (i < nw) {
            a.bm[0][i] = 0;
            a.bm[1][i] = 0;
            a.pend[i] = 0;
        }
    }
}

__global__ void sb_seed_kernel(SbArgs a, SbSeed sd, const int32_t *src) {
    const int32_t b = threadIdx.x;
    SbState *st = a.st;
    if (b < sd.B) {
        const int32_t s = src[b];
        sd.dist[(int64_t)b * a.V + s] = 0;
        a.bm[0][(int64_t)b * a.words + (s >> 6)] |= 1ull << (s & 63);  // one source per row
        st->mode[b] = SB_RELAX;
        st->lo[b] = 0;
        st->hi[b] = sd.delta;
        st->count[b] = 0;
        st->pmin[b] = INT_MAX;
    }
    if (b == 0) {
        st->done = 0;
        st->step = 0;
        st->cursor = 0;
        st->B = sd.B;
        st->delta = sd.delta;
        st->steps = 0;
        st->dist = sd.dist;
    }
}

static void sb_launch_step(const SbArgs &a, Launcher &L) {
    const int grid = persistent_grid(kSbBlock, 8);
    L(sb_relax_kernel, grid, kSbBlock, a);
    L(sb_decide_kernel, 1, kSbMaxBatch, a.st);
    L(sb_pmin_kernel, grid, kSbBlock, a);
    L(sb_window_kernel, 1, kSbMaxBatch, a.st);
    L(sb_move_kernel, grid, kSbBlock, a);
    L(sb_resume_kernel, 1, kSbMaxBatch, a.st);
}

struct SbRun {
    uint64_t *bm[2] = {nullptr, nullptr};
    uint64_t *pend = nullptr;
    SbState *st = nullptr;
    SbState *host = nullptr;  // pinned
    int32_t *d_src = nullptr;
    hipGraphExec_t exec = nullptr;
    ~SbRun() {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (auto p : bm) dfree(p);
        dfree(pend);
        dfree(st);
        dfree(d_src);
        if (host) (void)hipHostFree(host);
    }
};

static int sssp_batched(gi_graph *g, const int32_t *h_src, int32_t n, int32_t delta, int32_t batch,
                        int32_t *d_dist, gi_run_stats *stats, hipStream_t s) {
    SbRun r;
    const int32_t words = (int32_t)bitmap_words(g->V);
    const size_t BW = (size_t)batch * words;
    GI_TRY(dalloc(&r.bm[0], BW + 1));
    GI_TRY(dalloc(&r.bm[1], BW + 1));
    GI_TRY(dalloc(&r.pend, BW + 1));
    GI_TRY(dalloc(&r.st, 1));
    GI_TRY(dalloc(&r.d_src, (size_t)batch));
    if (hipHostMalloc((void **)&r.host, sizeof(SbState)) != hipSuccess) return fail(GI_ERR_OOM, "pinned allocation failed");
    SbArgs a{};
    a.st = r.st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.wt = g->d_w;
    a.V = g->V;
    a.words = words;
    a.bm[0] = r.bm[0];
    a.bm[1] = r.bm[1];
    a.pend = r.pend;
    GI_TRY(build_graph(
        [&](Launcher &L) {
            for (int32_t i = 0; i < kSbBlockSteps; ++i) sb_launch_step(a, L);
        },
        &r.exec));
    EventTimer timer(s);
    long long steps = 0;
    for (int32_t g0 = 0; g0 < n; g0 += batch) {
        const int32_t B = n - g0 < batch ? n - g0 : batch;
        SbSeed sd{d_dist + (int64_t)g0 * g->V, B, delta};
        GI_HIP(hipMemcpyAsync(r.d_src, h_src + g0, sizeof(int32_t) * B, hipMemcpyHostToDevice, s));
        const int64_t work = (int64_t)B * g->V;
        int grid = (int)ceil_div(work, 256);
        const int cap = persistent_grid(256);
        sb_reset_kernel<<<grid < cap ? grid : cap, 256, 0, s>>>(a, sd);
        sb_seed_kernel<<<1, kSbMaxBatch, 0, s>>>(a, sd, r.d_src);
        GI_LAUNCH_CHECK();
        for (int64_t used = 0;; used += kSbBlockSteps) {
            // every step settles a window of at least one source or relaxes: a
            // bound the state cannot exceed unless it never reports done
            if (used > 4 * ((int64_t)g->V + 64) * B) return fail(GI_ERR_HIP, "batched SSSP never reported done");
            GI_HIP(hipGraphLaunch(r.exec, s));
            GI_HIP(hipMemcpyAsync(r.host, r.st, offsetof(SbState, mode), hipMemcpyDeviceToHost, s));
            GI_HIP(hipStreamSynchronize(s));
            if (r.host->done) break;
        }
        steps += r.host->steps;
    }
    const double ms = timer.stop();
    if (stats) {
        *stats = gi_run_stats{};
        stats->iterations = (int32_t)(steps < INT32_MAX ? steps : INT32_MAX);
        stats->push_iterations = stats->iterations;
        stats->ms = ms;
    }
    return GI_OK;
}

}  // namespace gi

using namespace gi;

extern "C" int gi_sssp_batched(gi_graph *g, const int32_t *sources, int32_t num_sources, int32_t delta,
                               int32_t batch, int32_t *d_dist, gi_run_stats *stats, void *stream) {
    if (!g || (num_sources > 0 && (!sources || !d_dist))) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (!g->d_w) return fail(GI_ERR_INVALID, "SSSP needs edge weights");
    if (delta < 1) return fail(GI_ERR_RANGE, "delta must be positive");
    if (num_sources < 0) return fail(GI_ERR_RANGE, "negative source count");
    for (int32_t i = 0; i < num_sources; ++i)
        if (sources[i] < 0 || sources[i] >= g->V)
            return fail(GI_ERR_RANGE, "source %d out of range [0,%d)", sources[i], g->V);
    if (num_sources == 0) {
        if (stats) *stats = gi_run_stats{};
        return GI_OK;
    }
    if (batch < 1 || batch > kSbMaxBatch) batch = kSbMaxBatch;
    if (batch > num_sources) batch = num_sources;
    return sssp_batched(g, sources, num_sources, delta, batch, d_dist, stats, (hipStream_t)stream);
}
