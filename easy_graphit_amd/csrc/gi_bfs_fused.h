// Shared definitions of the fused BFS engine and the schedule selector.
#pragma once

#include <functional>
#include <vector>

#include "gi_internal.h"

namespace gi {

// Step modes of the fused engine (the selector's output classes).
enum BfsMode : int32_t {
    BFS_PUSH_THREAD = 0,  // VERTEX_BASED push
    BFS_PUSH_WAVE = 1,    // WM push
    BFS_PUSH_TWCE = 2,    // TWCE push (lane / wave granules inside a workgroup)
    BFS_PUSH_CM = 3,      // CM push (workgroup merge-path)
    BFS_PUSH_TWC = 4,     // TWC push with a grid-wide bin for hubs
    BFS_PULL_THREAD = 5,  // pull, lane per vertex, bitmap frontier
    BFS_PULL_GROUP = 6,   // pull, wave-cooperative over the unresolved vertices of a word
    BFS_NUM_MODES = 7
};

constexpr int kBfsMaxLevels = 4096;
constexpr int kSelInputs = 16;
constexpr int kSelHidden = 64;
constexpr int kSelMaxOut = 8;
// Push vertices at least this wide go to the grid-wide bin (TWC): they are
// listed when discovered, and the next step spreads their edges over its
// whole grid before it walks its chunks. Low enough that a small frontier of
// long lists (the source's neighbours) is spread over the whole GPU instead
// of one wavefront per list.
constexpr int32_t kHubDegree = 512;
constexpr int32_t kMaxHubs = 1024;
// Giant hubs (degree >= kGiantFactor * the hub threshold) are listed in a
// list of their own, kept in every step: they are few, and one wavefront
// walking a list of 10^5..10^6 edges (where the hub list was not kept or
// overflowed) would take milliseconds.
constexpr int32_t kGiantFactor = 64;
constexpr int32_t kMaxGiants = 512;
constexpr int32_t kMaxGrid = 8192;          // workgroups of a step launch at most
constexpr int32_t kMaxSumWords = 256;       // summary words (V up to 2^26)

// Device view of the selector network (BF16 weights, FP32 accumulation),
// kept in the run's device state. The BF16 weights are stored one row
// per hidden unit, so that a lane (one hidden unit) reads its own weights
// with 16-byte loads: w1 [hid][in], w2 [hid][hid] (row = the unit fed), w3
// [hid][kSelMaxOut] (row = the unit feeding, zero-padded), biases b1 [hid],
// b2 [hid], b3 [out].
// The blob holds them in that order from w1, at these uint16 offsets.
constexpr int kSelOffW1 = 0;
constexpr int kSelOffW2 = kSelOffW1 + kSelInputs * kSelHidden;
constexpr int kSelOffW3 = kSelOffW2 + kSelHidden * kSelHidden;
constexpr int kSelOffB1 = kSelOffW3 + kSelHidden * kSelMaxOut;
constexpr int kSelOffB2 = kSelOffB1 + kSelHidden;
constexpr int kSelOffB3 = kSelOffB2 + kSelHidden;
constexpr int kSelBlob = kSelOffB3 + kSelMaxOut;  // uint16 entries
struct SelectorView {
    const uint16_t *w1, *b1, *w2, *b2, *w3, *b3;
    int32_t n_out;
    int32_t allowed;     // bit mask of modes the selector may pick
    int32_t mode;        // gi_selector_mode
    int32_t mode_s1;     // heuristic's s1 / s2 modes
    int32_t mode_s2;
    float threshold;
    float feat_mean[kSelInputs];   // feature normalisation
    float feat_scale[kSelInputs];
    unsigned long long version;    // the installed model's (an engine copies the blob when it changes)
};

// One level's bookkeeping, written by the first workgroup of the level's
// step launch and read by the next launch.
struct BfsLevel {
    int32_t level;
    int32_t mode;                   // mode the level's step ran
    int32_t count;                  // |F_level|
    int32_t pull_levels;            // pull steps up to and including this one
    unsigned long long fdeg;        // sum of the degrees of F_level
    long long unvisited;            // vertices not reached once F_level is
    long long unvisited_edges;      // the sum of their degrees
};

struct BfsState {
    int32_t done;
    int32_t levels;                 // BFS depth + 1 when done
    int32_t pull_levels;            // levels expanded by a pull step, when done
    int32_t force_mode;             // >= 0: the mode of the next step (traced levels, sweeps)
    int32_t base;                   // level of the first launch of a replayed block of levels
    int32_t lane_max;               // push: degree < lane_max walked by one lane
    int32_t hub_min;                // push TWC: degree >= hub_min goes to the grid bin
    int32_t want_fdeg;              // accumulate frontier degrees (learned selector, sweeps)
    int32_t want_hubs;              // list hubs at discovery (the selector may pick TWC)
    int32_t hub_list_count;         // ... during push steps of at most this many vertices
    long long hub_list_fdeg;        // ... and at most this many edges (when degrees are summed)
    int32_t *depth;                 // output arrays of the current run
    int32_t *parent;
    SelectorView sel;               // the run's selector (the reset copies it here)
    BfsLevel lv[2];                 // by level parity; lv[1] holds level -1 after the reset
    int32_t hub_n[3];               // listed hubs of F_l in hub_n[l % 3]; more than kMaxHubs: void
    int32_t hubs[3][kMaxHubs];      // (below the giant threshold)
    int32_t giant_n[3];             // listed giant hubs of F_l; more than kMaxGiants: void
    int32_t giants[3][kMaxGiants];
    // per-level history of the first kBfsMaxLevels levels (traces and
    // statistics); deeper runs go on, unrecorded
    int32_t hist_mode[kBfsMaxLevels];
    int32_t hist_next[kBfsMaxLevels];
};

// A level's selector inputs as the step launch sees them (sweeps).
struct BfsPeek {
    int32_t n;
    int32_t prev_mode;
    int32_t prev_n;
    int32_t level;
    unsigned long long fdeg;
    long long unvisited;
    long long unvisited_edges;
};

// The selector's 16 input features from the run's counters, in double: the
// dataset recorder's definition (training inputs). The decide
// evaluates the same formulas in FP32 with the hardware log2
// (bfs_features_f32), equal to these within FP32 rounding.
__host__ __device__ inline float bfs_raw_feature(int k, int32_t n, unsigned long long fdeg,
                                                 long long unvisited, long long unvisited_edges,
                                                 int32_t level, int32_t prev_mode,
                                                 int32_t prev_n, int32_t V, int64_t E) {
    const double dV = V > 0 ? (double)V : 1.0, dE = E > 0 ? (double)E : 1.0;
    const double nn = n, fd = (double)fdeg;
    switch (k) {
        case 0: return (float)(nn / dV);                                  // frontier share of V
        case 1: return (float)(fd / dE);                                  // frontier share of E
        case 2: return (float)((double)unvisited / dV);                   // unvisited share of V
        case 3: return (float)((double)unvisited_edges / dE);             // unvisited edges share
        case 4: return (float)(log2(1.0 + nn) / 32.0);
        case 5: return (float)(log2(1.0 + fd) / 40.0);
        case 6: return (float)(log2(1.0 + fd / (nn > 0 ? nn : 1.0)) / 20.0);  // mean frontier degree
        case 7: return (float)(log2(1.0 + dE / dV) / 20.0);              // mean degree
        case 8: return (float)(log2(dV) / 32.0);
        case 9: return (float)(level < 64 ? level / 64.0 : 1.0);
        case 10: return prev_mode >= BFS_PULL_THREAD ? 1.f : 0.f;         // previous step pulled
        case 11: return (float)(log2(1.0 + nn) - log2(1.0 + (double)prev_n)) / 16.f;  // growth
        case 12: return (float)(fd / (1.0 + (double)unvisited_edges));    // Beamer's m_f / m_u
        case 13: return (float)(nn / (1.0 + (double)unvisited));
        case 14: return (float)(log2(1.0 + (double)unvisited) / 32.0);
        default: return 1.f;                                              // bias feature
    }
}

__host__ __device__ inline void bfs_raw_features(float *f, int32_t n, unsigned long long fdeg,
                                                 long long unvisited, long long unvisited_edges,
                                                 int32_t level, int32_t prev_mode,
                                                 int32_t prev_n, int32_t V, int64_t E) {
    for (int k = 0; k < kSelInputs; ++k)
        f[k] = bfs_raw_feature(k, n, fdeg, unvisited, unvisited_edges, level, prev_mode, prev_n, V, E);
}

struct FusedTrace {
    std::vector<hipEvent_t> ev;   // two per level: around the step kernel
    std::vector<int32_t> modes;   // mode of each level (the first kBfsMaxLevels)
    std::vector<int32_t> sizes;   // frontier size of each level (the first kBfsMaxLevels)
    int32_t pull_levels = 0;      // levels expanded by pull steps (all of them)
};

int32_t fused_mode_of(const gi_schedule &s);
// Work queued behind the level loop (e.g. mapping results back to original
// ids). It is enqueued before the host waits for the loop's done flag, so it
// overlaps that wait; if the loop needs another block of levels it is
// enqueued again after it.
using FusedEpilogue = std::function<int(hipStream_t)>;
int bfs_fused_run(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, SelectorView sel,
                  int32_t *d_depth, int32_t *d_parent, hipStream_t s, FusedTrace *trace,
                  bool timed, int32_t *levels_out, const FusedEpilogue *epilogue = nullptr,
                  int32_t depth_len = -1);
void release_fused(const gi_graph *g);
int fused_reached(gi_graph *g, const int32_t *depth, int32_t *out, hipStream_t s);
int fused_set_bins(gi_graph *g, int32_t lane_max, int32_t hub_min, int32_t blocks_per_cu);
int bfs_fused_run_multi(gi_graph *g, const int32_t *h_sources, int32_t n_sources, int32_t single,
                        const gi_hybrid_schedule *h, SelectorView sel, int32_t *d_depth,
                        int32_t *d_parent, hipStream_t s, FusedTrace *trace, bool timed,
                        int32_t *levels_out, const FusedEpilogue *epilogue = nullptr,
                        int32_t depth_len = -1);
// The process-wide selector model (null weights when none is loaded).
SelectorView selector_view();

}  // namespace gi
