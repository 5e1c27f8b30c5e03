// Many scenes rendered in one launch sequence on one device (the MIS integrator of gz_mis.hip
// over several scene handles, each with its own camera, film, parameters and BVH).
//
// The paths of all scenes are cut into rounds of at most `roundPaths` path slots. A round holds
// items: whole pixels of one scene, in scene order, so a round has at most one item more than
// there are scenes. A round is one fixed sequence of launches that reads everything it needs
// from device tables and a device round cursor:
//   camera  one lane per slot: the item's camera sample and primary ray (film_device.h);
//   clear   the per-level, per-item queue counters (a kernel, no memset: graph-safe);
//   per level 0 .. maxDepth:
//     plan    one lane: each item's chunk count and their running sum;
//     bounce  persistent waves take one chunk per atomic; a chunk belongs to one item, so the
//             wave's scene is uniform and its MisArgs come from the scene table by scalar loads;
//             survivors go to the item's queue of the next level by ballot + prefix popcount;
//   film    one lane per pixel: the pixel's samples into its scene's film, in sample order;
//   next    the round cursor moves on.
// Captured once as a HIP graph, the sequence is replayed once per round with no host work
// between rounds; the eager mode launches the same sequence directly.
//
// A scene with more than `soloPaths` paths fills the chip by itself: it stays out of the rounds
// and renders alone, with the launch sequence of the single-scene path (camera kernel,
// gz_gpu_mis_trace_device's kernels under the scene's own launch shape, film), in pieces of whole
// pixels, in four lanes: lane 0 on the caller's stream after the rounds, lanes 1-3 on side streams
// from the start of the render, each with its own path slots and queue storage, so that the level
// tails of one scene (the chip draining a level's last waves) overlap the other lanes' work. A
// scene's pieces stay in one lane, and scenes that share a film buffer share a lane. The rounds' kernels pay a
// per-chunk item lookup and MisArgs read from memory that the single-scene kernels (MisArgs as a
// kernel argument) do not; that only pays for scenes too small to fill the chip alone.
#include <algorithm>
#include <cstring>
#include <vector>

#include "film_device.h"
#include "mis_device.h"

#pragma clang fp contract(off)

namespace gz {

struct BatchItem {
        int32_t scene;
        int32_t count;      // path slots (whole pixels of the scene's spp)
        int64_t first;      // first path of the item in the scene's crop numbering
        int32_t offset;     // first slot in the round's path buffer
        int32_t pixOffset;  // first pixel of the item among the round's pixels
        int32_t wave;       // wavefront rounds: 1 = the wavefront kernels bounce it, 0 = the bounce kernel
        int32_t pad;
};

struct BatchRound {
        int32_t firstItem;
        int32_t nItems;
        int32_t total;   // path slots
        int32_t pixels;
};

struct BatchFilm {
        float *light, *weight, *albedo, *normal;
        float iso, maxComponent;
        int32_t resX, resY, spp, pad;
};

// The round's item containing position `pos` of a running sum `key` (largest i with key(i) <= pos).
template <typename Key>
__device__ __forceinline__ int find_item(int n, int pos, Key &&key) {
        int lo = 0, hi = n;  // answer in [lo, hi)
        while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (key(mid) <= pos)
                        lo = mid;
                else
                        hi = mid;
        }
        return lo;
}

__global__ void k_batch_begin(int *cursor) {
        if (threadIdx.x == 0 && blockIdx.x == 0) cursor[0] = 0;
}

__global__ void k_batch_next(int *cursor) {
        if (threadIdx.x == 0 && blockIdx.x == 0) cursor[0] = cursor[0] + 1;
}

constexpr int kBatchBlock = 256;

__global__ __launch_bounds__(kBatchBlock) void k_batch_camera(const int *__restrict__ cursor,
                                                              const BatchRound *__restrict__ rounds,
                                                              const BatchItem *__restrict__ items,
                                                              const GzCamera *__restrict__ cams,
                                                              const uint32_t *__restrict__ table,
                                                              GzPath *__restrict__ paths) {
        const BatchRound R = rounds[cursor[0]];
        const int t = (int)(blockIdx.x * kBatchBlock + threadIdx.x);
        if (t >= R.total) return;
        const BatchItem *its = items + R.firstItem;
        const int k = find_item(R.nItems, t, [&](int i) { return its[i].offset; });
        const BatchItem it = its[k];
        paths[t] = camera_path(cams[it.scene], table, it.first + (t - it.offset));
}

// Per level: n[level][item] queue entries, head[level] chunks handed out, base[level][item]
// the chunk running sum (base[level][nItems] = the level's chunks).
struct BatchCounters {
        unsigned *n;
        unsigned *head;
        unsigned *base;
        int maxItems;
        int levels;
};

__global__ void k_batch_clear(BatchCounters c) {
        const int nn = (c.levels + 1) * c.maxItems;
        for (int k = threadIdx.x; k < nn; k += blockDim.x) c.n[k] = 0u;
        for (int k = threadIdx.x; k < c.levels; k += blockDim.x) c.head[k] = 0u;
}

// One lane: each item's chunk count at `level` and their running sum, into `base`. select -1
// counts every item, 0 and 1 only the items whose `wave` flag equals it (the others get no chunks).
__global__ void k_batch_plan(const int *__restrict__ cursor, const BatchRound *__restrict__ rounds,
                             const BatchItem *__restrict__ items, BatchCounters c, int level, unsigned chunk,
                             unsigned *base, int select) {
        if (threadIdx.x != 0 || blockIdx.x != 0) return;
        const BatchRound R = rounds[cursor[0]];
        unsigned sum = 0;
        for (int i = 0; i < R.nItems; ++i) {
                base[i] = sum;
                const BatchItem &it = items[R.firstItem + i];
                if (select >= 0 && it.wave != select) continue;
                const unsigned n = level == 0 ? (unsigned)it.count : c.n[level * c.maxItems + i];
                sum += (n + chunk - 1) / chunk;
        }
        base[R.nItems] = sum;
}

// One bounce level of a round: waves take chunks of the level's work items through one global
// counter; survivors go to the item's next-level queue (appends held over four runs of 64 and
// written with one atomic). kDrain: a lane runs each path's remaining bounces instead, and no
// queue is written (the last levels of a round, in one launch).
template <int kAll, bool kDrain>
__global__ __launch_bounds__(kMisBlock, mis_min_waves(kAll)) void k_batch_bounce(const int *__restrict__ cursor,
                                                            const BatchRound *__restrict__ rounds,
                                                            const BatchItem *__restrict__ items,
                                                            const MisArgs *__restrict__ args, GzPath *__restrict__ paths,
                                                            const int *__restrict__ qin, int *__restrict__ qout,
                                                            BatchCounters c, int level, unsigned chunk) {
        const BatchRound R = rounds[cursor[0]];
        const BatchItem *its = items + R.firstItem;
        const unsigned *base = c.base + (size_t)level * (c.maxItems + 1);
        const unsigned total = base[R.nItems];
        unsigned *nOut = c.n + (size_t)(level + 1) * c.maxItems;
        const unsigned *nIn = c.n + (size_t)level * c.maxItems;
        const int lane = (int)(threadIdx.x & 63);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        for (;;) {
                unsigned g = 0;
                if (lane == 0) g = atomicAdd(&c.head[level], 1u);
                g = (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl(g, 0));
                if (g >= total) break;
                const int k = __builtin_amdgcn_readfirstlane(
                        find_item(R.nItems, (int)g, [&](int i) { return (int)base[i]; }));
                const BatchItem it = its[k];
                const MisArgs &a = args[__builtin_amdgcn_readfirstlane(it.scene)];
                const unsigned n = level == 0 ? (unsigned)it.count : nIn[k];
                const unsigned e0 = (g - base[k]) * chunk;
                const unsigned e1 = n - e0 < chunk ? n : e0 + chunk;
                if (kDrain) {
                        for (unsigned j = e0 + (unsigned)lane; j < e1; j += 64) {
                                const int idx = level == 0 ? it.offset + (int)j : qin[it.offset + (int)j];
                                for (int b = 0; b <= a.maxDepth; ++b)
                                        if (!bounce_one<false, kAll, false>(a, paths, idx, nullptr, AdjArgs{nullptr, nullptr},
                                                                           nullptr))
                                                break;
                        }
                        continue;
                }
                int *const qo[1] = {qout + it.offset};
                unsigned *const on[1] = {&nOut[k]};
                Keep4 keep;
                unsigned bits[1] = {0u};
                int runs = 0;
                for (unsigned j0 = e0; j0 < e1; j0 += 64) {
                        const unsigned j = j0 + (unsigned)lane;
                        int idx = -1;
                        if (j < e1) idx = level == 0 ? it.offset + (int)j : qin[it.offset + (int)j];
                        const bool go = idx >= 0 && bounce_one<false, kAll, false>(a, paths, idx, nullptr,
                                                                                   AdjArgs{nullptr, nullptr}, nullptr);
                        keep.set(runs, idx);
                        bits[0] |= (go ? 1u : 0u) << runs;
                        if (++runs == 4) {
                                flush_appends<1>(keep, bits, runs, qo, on, lane, below);
                                runs = 0;
                                bits[0] = 0u;
                        }
                }
                if (runs > 0) flush_appends<1>(keep, bits, runs, qo, on, lane, below);
        }
}

__global__ __launch_bounds__(kBatchBlock) void k_batch_film(const int *__restrict__ cursor,
                                                            const BatchRound *__restrict__ rounds,
                                                            const BatchItem *__restrict__ items,
                                                            const BatchFilm *__restrict__ films,
                                                            const GzPath *__restrict__ paths) {
        const BatchRound R = rounds[cursor[0]];
        const int p = (int)(blockIdx.x * kBatchBlock + threadIdx.x);
        if (p >= R.pixels) return;
        const BatchItem *its = items + R.firstItem;
        const int k = find_item(R.nItems, p, [&](int i) { return its[i].pixOffset; });
        const BatchItem it = its[k];
        const BatchFilm f = films[it.scene];
        film_add_pixel(f.resX, f.resY, f.spp, paths + it.offset + (int64_t)(p - it.pixOffset) * f.spp, f.iso,
                       f.maxComponent, f.light, f.weight, f.albedo, f.normal);
}

// ---- Wavefront rounds (launch queued = 3) ------------------------------------------------------
//
// Each bounce level of a round as the single-scene full wavefront does it (gz_mis.hip): a
// closest-hit kernel that also sorts the hits by BSDF class, one shade kernel per class that
// queues each path's shadow ray and BSDF ray, an any-hit and a closest-hit kernel over those ray
// queues and one finish kernel per class. Every queue is kept per item (at the item's slot
// offset), so every chunk a wave takes belongs to one scene and its MisArgs stay uniform. The
// level queues and their counters are the megakernel rounds' own (BatchCounters), so the drain
// launch serves both.
// kWLevel, kWLevelSplit: only the chunk sums of the wavefront items' level queues (their counts
// are the BatchCounters'), in the traversal kernel's chunks and in the split kernel's.
enum : int { kWDiff = 0, kWAll = 1, kWShadow = 2, kWBsdf = 3, kWLevel = 4, kWLevelSplit = 5, kWKinds = 6 };

// n[level][kind][item] entries, base[level][kind][item] the chunk running sum
// (base[...][nItems] = the chunks of that kind).
struct WaveCounters {
        unsigned *n;
        unsigned *base;
        int maxItems;
        int levels;
        __device__ __forceinline__ unsigned *count(int level, int kind) const {
                return n + ((size_t)level * kWKinds + kind) * maxItems;
        }
        __device__ __forceinline__ unsigned *bases(int level, int kind) const {
                return base + ((size_t)level * kWKinds + kind) * (maxItems + 1);
        }
        unsigned *baseOf(int level, int kind) const { return base + ((size_t)level * kWKinds + kind) * (maxItems + 1); }
        unsigned *countOf(int level, int kind) const { return n + ((size_t)level * kWKinds + kind) * maxItems; }
};

// Per-slot state of the wavefront rounds.
struct WaveSlots {
        GzHit *hits;
        Pending *pend;
        Ray_C *srays, *mrays;
        uint8_t *occ;
        GzHit *mhits;
        int *q[4];  // kWDiff .. kWBsdf
};

__global__ void k_bw_clear(WaveCounters w) {
        const size_t nn = (size_t)w.levels * kWKinds * w.maxItems;
        for (size_t k = threadIdx.x; k < nn; k += blockDim.x) w.n[k] = 0u;
}

// One lane: the chunk running sums of the kinds in `mask` at `level`.
__global__ void k_bw_plan(const int *__restrict__ cursor, const BatchRound *__restrict__ rounds, WaveCounters w,
                          int level, int mask, unsigned chunk) {
        if (threadIdx.x != 0 || blockIdx.x != 0) return;
        const BatchRound R = rounds[cursor[0]];
        for (int kind = 0; kind < kWLevel; ++kind) {
                if (!((mask >> kind) & 1)) continue;
                const unsigned *n = w.count(level, kind);
                unsigned *base = w.bases(level, kind);
                unsigned sum = 0;
                for (int i = 0; i < R.nItems; ++i) {
                        base[i] = sum;
                        sum += (n[i] + chunk - 1) / chunk;
                }
                base[R.nItems] = sum;
        }
}

// The next chunk of a queue kind for the wave: its item k and entries [e0, e1) of the item's
// queue; false once the kind's chunks are done. Wave w takes the chunks w, w + waves, ... (no
// atomic: a counter that every wave hits serialises). Wave-uniform.
struct Chunk {
        int k;  // the item of the current chunk (-1 before the first)
        unsigned e0, e1;
        unsigned g;  // the wave's next chunk
};
__device__ __forceinline__ void first_chunk(Chunk &c) {
        c.g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        c.k = -1;
}
// The item holding chunk g: the wave's previous item or the next one when g lies in it (a
// round's items are few and long, and a wave's chunks come in order), else a binary search.
__device__ __forceinline__ int chunk_item(const BatchRound &R, const unsigned *base, unsigned g, int prev) {
        if (prev >= 0 && prev < R.nItems && g >= base[prev]) {
                if (g < base[prev + 1]) return prev;
                if (prev + 1 < R.nItems && g < base[prev + 2]) return prev + 1;
        }
        return find_item(R.nItems, (int)g, [&](int i) { return (int)base[i]; });
}
__device__ __forceinline__ bool next_chunk(const BatchRound &R, const unsigned *base, const unsigned *n, unsigned chunk,
                                           Chunk &c) {
        const unsigned g = (unsigned)__builtin_amdgcn_readfirstlane((int)c.g);
        c.g = g + gridDim.x * (blockDim.x >> 6);
        if (g >= base[R.nItems]) return false;
        c.k = __builtin_amdgcn_readfirstlane(chunk_item(R, base, g, c.k));
        const unsigned cnt = n[c.k];
        c.e0 = (g - base[c.k]) * chunk;
        c.e1 = cnt - c.e0 < chunk ? cnt : c.e0 + chunk;
        return true;
}

// The rays of one queue kind of the round's items. A wave takes 64-entry chunks by static
// striding, one entry per lane; a chunk lies in one item, so the BVH pointers stay uniform
// (scalar registers), and each lane traces its ray to the end. The chunks are 64 entries whatever
// the launch's chunk: with a loop over a longer chunk's entries around the traversal loop the
// kernel ran 9% slower on a 64-scene batch (2.22 against 2.01 s). (Lane refill across chunks, which lets lanes of one wave hold
// different scenes and so keeps their BVH pointers per lane, measured slower: path rays 958
// against 752 ms, BSDF rays 597 against 505 ms per batch of 64 scenes.)
// kMode 0: the path rays of the level queue (closest hit into hits, paths that stop get a miss),
// 1: the shadow rays (any hit below 1 - 1e-5 into occ), 2: the BSDF rays (closest hit into mhits).
template <int kMode>
__global__ __launch_bounds__(kMisBlock) void k_bw_rays(const int *__restrict__ cursor,
                                                       const BatchRound *__restrict__ rounds,
                                                       const BatchItem *__restrict__ items,
                                                       const MisArgs *__restrict__ args,
                                                       const GzPath *__restrict__ paths, const int *__restrict__ qin,
                                                       const unsigned *__restrict__ base, const unsigned *__restrict__ nIn,
                                                       WaveSlots ws, int level) {
        constexpr bool kAny = kMode == 1;
        Stack<kLdsMax> S;
        const BatchRound R = rounds[cursor[0]];
        const BatchItem *its = items + R.firstItem;
        const int lane = (int)(threadIdx.x & 63);
        Chunk ch;
        first_chunk(ch);
        for (;;) {
                const unsigned g = (unsigned)__builtin_amdgcn_readfirstlane((int)ch.g);
                ch.g = g + gridDim.x * (blockDim.x >> 6);
                if (g >= base[R.nItems]) break;
                const int k = __builtin_amdgcn_readfirstlane(chunk_item(R, base, g, ch.k));
                ch.k = k;
                const BatchItem it = its[k];
                const unsigned n = (kMode == 0 && level == 0) ? (unsigned)it.count : nIn[k];
                const unsigned j = (g - base[k]) * 64u + (unsigned)lane;
                const MisArgs &a = args[__builtin_amdgcn_readfirstlane(it.scene)];
                if (j < n) {
                        const int idx = (kMode == 0 && level == 0) ? it.offset + (int)j : qin[it.offset + (int)j];
                        Lane L;
                        L.hitPrim = -1;
                        bool trace = true;
                        if (kMode == 0) {
                                const GzPath &p = paths[idx];
                                trace = path_live(a, p);
                                if (trace) lane_init(L, p.ray, INFINITY, a.rootDesc, a.sceneSafe);
                        } else {
                                lane_init(L, kAny ? ws.srays[idx] : ws.mrays[idx], kAny ? 1.0f - 0.00001f : INFINITY,
                                          a.rootDesc, a.sceneSafe);
                        }
                        if (trace && !a.empty) {
                                int st;
                                do {
                                        st = step<kAny, false, kLdsMax>(L, S, a.nodes, a.tris);
                                } while (st == kStepGo);
                                if (st == kStepOverflow) {
                                        atomicOr(a.err, 1);
                                        L.hitPrim = -1;
                                }
                        }
                        if (kAny) {
                                ws.occ[idx] = L.hitPrim >= 0 ? 1 : 0;
                        } else {
                                const GzHit h = L.hitPrim >= 0 ? GzHit{L.hitPrim, L.tBest, L.bu, L.bv}
                                                               : GzHit{-1, INFINITY, 0.0f, 0.0f};
                                if (kMode == 0)
                                        ws.hits[idx] = h;
                                else
                                        ws.mhits[idx] = h;
                        }
                }
        }
}

// Level `level`: the wavefront items' level queues sorted by the BSDF class of their hits into
// the items' class queues (misses and stopped paths to the diffuse one, whose shade kernel
// clears them). A block takes tiles of kSplitTile entries of one item (static striding) and
// appends each class with one atomic per tile: the waves of the chip all work on the same item
// at a time, so per-wave atomics on its counters would serialise.
constexpr int kSplitRuns = 8;  // runs of 64 per wave and tile
constexpr unsigned kSplitTile = (unsigned)kMisBlock * kSplitRuns;

__global__ __launch_bounds__(kMisBlock) void k_bw_split(const int *__restrict__ cursor,
                                                        const BatchRound *__restrict__ rounds,
                                                        const BatchItem *__restrict__ items,
                                                        const MisArgs *__restrict__ args, const int *__restrict__ qin,
                                                        BatchCounters c, WaveCounters w, WaveSlots ws, int level) {
        constexpr int kWaves = kMisBlock / 64;
        __shared__ unsigned sCnt[kWaves][2];
        __shared__ unsigned sBase[2];
        const BatchRound R = rounds[cursor[0]];
        const BatchItem *its = items + R.firstItem;
        const unsigned *base = w.bases(level, kWLevelSplit);
        const unsigned *nIn = c.n + (size_t)level * c.maxItems;
        const int lane = (int)(threadIdx.x & 63);
        const int wave = (int)(threadIdx.x >> 6);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const unsigned tiles = base[R.nItems];
        int kPrev = -1;
        for (unsigned g = blockIdx.x; g < tiles; g += gridDim.x) {
                const int k = __builtin_amdgcn_readfirstlane(chunk_item(R, base, g, kPrev));
                kPrev = k;
                const BatchItem it = its[k];
                const unsigned n = level == 0 ? (unsigned)it.count : nIn[k];
                const unsigned t0 = (g - base[k]) * kSplitTile;
                const MisArgs &a = args[__builtin_amdgcn_readfirstlane(it.scene)];
                const unsigned w0 = t0 + (unsigned)wave * 64u * kSplitRuns;
                int idx[kSplitRuns];
                unsigned bitsD = 0, bitsA = 0, cd = 0, ca = 0;
#pragma unroll
                for (int r = 0; r < kSplitRuns; ++r) {
                        const unsigned j = w0 + (unsigned)r * 64u + (unsigned)lane;
                        int i = -1;
                        if (j < n) i = level == 0 ? it.offset + (int)j : qin[it.offset + (int)j];
                        idx[r] = i;
                        const bool toA = i >= 0 && hit_needs_all_bsdfs(a, ws.hits[i].prim);
                        const bool toD = i >= 0 && !toA;
                        bitsA |= (unsigned)toA << r;
                        bitsD |= (unsigned)toD << r;
                        cd += (unsigned)__popcll(__ballot(toD));
                        ca += (unsigned)__popcll(__ballot(toA));
                }
                if (lane == 0) {
                        sCnt[wave][0] = cd;
                        sCnt[wave][1] = ca;
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                        unsigned td = 0, ta = 0;
                        for (int v = 0; v < kWaves; ++v) {
                                td += sCnt[v][0];
                                ta += sCnt[v][1];
                        }
                        sBase[0] = td ? atomicAdd(w.count(level, kWDiff) + k, td) : 0u;
                        sBase[1] = ta ? atomicAdd(w.count(level, kWAll) + k, ta) : 0u;
                }
                __syncthreads();
                unsigned od = sBase[0], oa = sBase[1];
                for (int v = 0; v < wave; ++v) {
                        od += sCnt[v][0];
                        oa += sCnt[v][1];
                }
                int *qd = ws.q[kWDiff] + it.offset, *qa = ws.q[kWAll] + it.offset;
#pragma unroll
                for (int r = 0; r < kSplitRuns; ++r) {
                        const bool toD = (bitsD >> r) & 1u, toA = (bitsA >> r) & 1u;
                        const uint64_t md = __ballot(toD), ma = __ballot(toA);
                        if (toD) qd[od + (unsigned)__popcll(md & below)] = idx[r];
                        if (toA) qa[oa + (unsigned)__popcll(ma & below)] = idx[r];
                        od += (unsigned)__popcll(md);
                        oa += (unsigned)__popcll(ma);
                }
                __syncthreads();  // sCnt and sBase are the next tile's
        }
}

// shade_phase over one class queue of the items; survivors to the items' next-level queues,
// shadow and BSDF rays to the items' ray queues.
template <int kAll>
__global__ __launch_bounds__(kMisBlock, mis_min_waves(kAll)) void k_bw_phase(const int *__restrict__ cursor,
                                                                      const BatchRound *__restrict__ rounds,
                                                                      const BatchItem *__restrict__ items,
                                                                      const MisArgs *__restrict__ args,
                                                                      GzPath *__restrict__ paths, int *__restrict__ qout,
                                                                      BatchCounters c, WaveCounters w, WaveSlots ws,
                                                                      int level, unsigned chunk) {
        const int kind = kAll ? kWAll : kWDiff;
        const BatchRound R = rounds[cursor[0]];
        const BatchItem *its = items + R.firstItem;
        const int lane = (int)(threadIdx.x & 63);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const int *qin = ws.q[kind];
        unsigned *nNext = c.n + (size_t)(level + 1) * c.maxItems;
        // Held runs carry over from chunk to chunk while the item stays the same, so that small
        // chunks cost no more queue atomics than large ones.
        Keep4 keep;
        unsigned bits[3] = {0u, 0u, 0u};
        int runs = 0;
        int heldK = -1, heldOff = 0;
        auto flush = [&]() {
                if (runs > 0) {
                        int *const qo[3] = {qout + heldOff, ws.q[kWShadow] + heldOff, ws.q[kWBsdf] + heldOff};
                        unsigned *const on[3] = {nNext + heldK, w.count(level, kWShadow) + heldK,
                                                 w.count(level, kWBsdf) + heldK};
                        flush_appends<3>(keep, bits, runs, qo, on, lane, below);
                }
                runs = 0;
                bits[0] = bits[1] = bits[2] = 0u;
        };
        Chunk ch;
        first_chunk(ch);
        while (next_chunk(R, w.bases(level, kind), w.count(level, kind), chunk, ch)) {
                const BatchItem it = its[ch.k];
                const MisArgs &a = args[__builtin_amdgcn_readfirstlane(it.scene)];
                if (ch.k != heldK) {
                        flush();
                        heldK = ch.k;
                        heldOff = it.offset;
                }
                for (unsigned j0 = ch.e0; j0 < ch.e1; j0 += 64) {
                        const unsigned j = j0 + (unsigned)lane;
                        int idx = -1;
                        if (j < ch.e1) idx = qin[it.offset + (int)j];
                        unsigned mask = 0u;
                        if (idx >= 0) {
                                const GzPath p = paths[idx];
                                Pending pd;
                                pd.flags = 0;
                                if (!(p.flags & GZ_PATH_ACTIVE)) {
                                } else if (p.bounce > a.maxDepth) {
                                        paths[idx].flags = p.flags & ~GZ_PATH_ACTIVE;
                                } else {
                                        const GzHit h = ws.hits[idx];
                                        Ray_C sr, mr;
                                        const bool go = shade_phase<false, kAll>(a, paths, idx, p, h.prim >= 0, h.prim, h.u,
                                                                                 h.v, pd, sr, mr, nullptr);
                                        if (pd.flags & kPendShadow) ws.srays[idx] = sr;
                                        if (pd.flags & kPendMis) ws.mrays[idx] = mr;
                                        mask = (go ? 1u : 0u) | ((pd.flags & kPendShadow) ? 2u : 0u) |
                                               ((pd.flags & kPendMis) ? 4u : 0u);
                                }
                                ws.pend[idx] = pd;
                        }
                        keep.set(runs, idx);
#pragma unroll
                        for (int q = 0; q < 3; ++q) bits[q] |= ((mask >> q) & 1u) << runs;
                        if (++runs == 4) flush();
                }
        }
        flush();
}

// finish_phase over one class queue of the items.
__global__ __launch_bounds__(kMisBlock) void k_bw_finish(const int *__restrict__ cursor,
                                                         const BatchRound *__restrict__ rounds,
                                                         const BatchItem *__restrict__ items,
                                                         const MisArgs *__restrict__ args, GzPath *__restrict__ paths,
                                                         WaveCounters w, WaveSlots ws, int level, int kind,
                                                         unsigned chunk) {
        const BatchRound R = rounds[cursor[0]];
        const BatchItem *its = items + R.firstItem;
        const int lane = (int)(threadIdx.x & 63);
        const int *qin = ws.q[kind];
        Chunk ch;
        first_chunk(ch);
        while (next_chunk(R, w.bases(level, kind), w.count(level, kind), chunk, ch)) {
                const BatchItem it = its[ch.k];
                const MisArgs &a = args[__builtin_amdgcn_readfirstlane(it.scene)];
                for (unsigned j = ch.e0 + (unsigned)lane; j < ch.e1; j += 64) {
                        const int idx = qin[it.offset + (int)j];
                        const Pending pd = ws.pend[idx];
                        if (!(pd.flags & kPendWork)) continue;
                        const bool o = (pd.flags & kPendShadow) ? ws.occ[idx] != 0 : false;
                        const Ray_C mr = (pd.flags & kPendMis) ? ws.mrays[idx] : Ray_C{0, 0, 0, 0, 0, 0};
                        const GzHit mh = (pd.flags & kPendMis) ? ws.mhits[idx] : GzHit{-1, 0.0f, 0.0f, 0.0f};
                        finish_phase(a, paths, idx, pd, o, mr, mh);
                }
        }
}

// Lanes of the scenes that render alone: the render's stream and kSoloLanes - 1 side streams.
constexpr int kSoloLanes = 4;  // 64-scene batch: 1598 ms with 2, 1612 with 3, 1550 with 4

struct Batch {
        int device = 0;
        int nScenes = 0;
        std::vector<Scene *> scenes;
        bool all = false;  // some scene has a non-diffuse material
        int bsdfSet = kBsdfDiffuse;  // the BSDF set of the bounce kernels: the widest of the scenes' sets
        bool anyWave = false, anyMega = false;  // wavefront rounds: items of each kind exist
        int levels = 0;    // maxDepth + 1 over the scenes
        int64_t totalPaths = 0;
        int roundPaths = 0;
        int maxRoundPixels = 0;
        int maxItems = 0;
        std::vector<BatchRound> hRounds;
        std::vector<BatchItem> hItems;
        GzMisLaunch launch{1, 256, 1, 256, 3, {0, 0, 0}};
        int cuCount = 256;

        int *dCursor = nullptr;
        BatchRound *dRounds = nullptr;
        BatchItem *dItems = nullptr;
        MisArgs *dArgs = nullptr;
        GzCamera *dCams = nullptr;
        BatchFilm *dFilms = nullptr;
        GzPath *dPaths = nullptr;
        int *dQ[2] = {nullptr, nullptr};
        BatchCounters ctr{nullptr, nullptr, nullptr, 0, 0};
        const uint32_t *sobol = nullptr;
        // wavefront rounds (allocated by the first launch shape that asks for them)
        WaveCounters wctr{nullptr, nullptr, 0, 0};
        WaveSlots ws{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, {nullptr, nullptr, nullptr, nullptr}};

        hipStream_t captureStream = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;

        // The batch's own error flag (the MisArgs copies raise it; gz_batch_check reads it) and
        // each scene's MIS-table generation at create: dArgs holds pointers into those tables,
        // which gz_gpu_set_mis_scene frees and replaces.
        int *dErr = nullptr;
        std::vector<uint64_t> misGeneration;

        // Scenes that render alone (more than soloPaths paths), in scene order, and their lanes:
        // lane 0 uses dPaths (pathSlots >= soloSlots) on the render's stream, lane l > 0 uses
        // dPathsSide[l - 1] on side stream side[l - 1] (created for the lanes that have scenes);
        // each lane has one queue storage shared by its scenes.
        std::vector<int> solo;
        std::vector<int> soloLane;
        GzPath *dPathsSide[kSoloLanes - 1] = {};
        hipStream_t side[kSoloLanes - 1] = {};
        hipEvent_t join[kSoloLanes - 1] = {};
        hipEvent_t fork = nullptr;
        std::vector<GzMisParams> params;
        std::vector<GzCamera> hCams;
        std::vector<BatchFilm> hFilms;
        int64_t soloSlots = 0;
        int64_t pathSlots = 0;
        MisQueues *soloQueues[kSoloLanes] = {};
};

static void release_graph(Batch *b) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
}

static void free_batch(Batch *b) {
        if (!b) return;
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != b->device) (void)hipSetDevice(b->device);
        release_graph(b);
        if (b->captureStream) (void)hipStreamDestroy(b->captureStream);
        for (int l = 0; l < kSoloLanes - 1; ++l) {
                if (b->side[l]) (void)hipStreamDestroy(b->side[l]);
                if (b->join[l]) (void)hipEventDestroy(b->join[l]);
                if (b->dPathsSide[l]) (void)hipFree(b->dPathsSide[l]);
        }
        if (b->fork) (void)hipEventDestroy(b->fork);
        void *ptrs[] = {b->dErr,      b->dCursor,   b->dRounds,  b->dItems,   b->dArgs,     b->dCams,     b->dFilms,
                        b->dPaths,    b->dQ[0],    b->dQ[1],    b->ctr.n,     b->ctr.head,  b->ctr.base,
                        b->wctr.n,    b->wctr.base, b->ws.hits,  b->ws.pend,   b->ws.srays,
                        b->ws.mrays,  b->ws.occ,   b->ws.mhits, b->ws.q[0],  b->ws.q[1],   b->ws.q[2],
                        b->ws.q[3]};
        for (void *p : ptrs)
                if (p) (void)hipFree(p);
        for (MisQueues *q : b->soloQueues) release_mis_queues(q);
        delete b;
}

// The wavefront rounds' counters and per-slot state (once per batch).
static bool ensure_wave(Batch *b) {
        if (b->wctr.n) return true;
        const size_t slots = (size_t)b->roundPaths;
        WaveCounters w{nullptr, nullptr, b->ctr.maxItems, b->levels};
        WaveSlots ws{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, {nullptr, nullptr, nullptr, nullptr}};
        bool ok = hip_ok(hipMalloc(&w.n, (size_t)w.levels * kWKinds * w.maxItems * sizeof(unsigned)), "batch wave counters") &&
                  hip_ok(hipMalloc(&w.base, (size_t)w.levels * kWKinds * (w.maxItems + 1) * sizeof(unsigned)),
                         "batch wave counters") &&
                  hip_ok(hipMalloc(&ws.hits, slots * sizeof(GzHit)), "batch hits") &&
                  hip_ok(hipMalloc(&ws.pend, slots * sizeof(Pending)), "batch pending terms") &&
                  hip_ok(hipMalloc(&ws.srays, slots * sizeof(Ray_C)), "batch shadow rays") &&
                  hip_ok(hipMalloc(&ws.mrays, slots * sizeof(Ray_C)), "batch BSDF rays") &&
                  hip_ok(hipMalloc(&ws.occ, slots), "batch shadow results") &&
                  hip_ok(hipMalloc(&ws.mhits, slots * sizeof(GzHit)), "batch BSDF ray hits");
        for (int k = 0; ok && k < 4; ++k) ok = hip_ok(hipMalloc(&ws.q[k], slots * sizeof(int)), "batch wave queues");
        b->wctr = w;  // freed with the batch, also after a failure
        b->ws = ws;
        if (!ok) {
                void *ptrs[] = {w.n, w.base, ws.hits, ws.pend, ws.srays, ws.mrays, ws.occ, ws.mhits,
                                ws.q[0], ws.q[1], ws.q[2], ws.q[3]};
                for (void *p : ptrs)
                        if (p) (void)hipFree(p);
                b->wctr = WaveCounters{nullptr, nullptr, 0, 0};
                b->ws = WaveSlots{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, {nullptr, nullptr, nullptr, nullptr}};
        }
        return ok;
}

// Persistent grid of a batch kernel that takes `chunk`-entry chunks, one per wave:
// gridScale x its resident blocks per CU x CUs, or fewer when a round has fewer chunks.
template <typename K>
static unsigned batch_grid(const Batch *b, K kern, int block, unsigned chunk) {
        int perCU = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, block, 0) != hipSuccess || perCU < 1) perCU = 1;
        const int64_t want = (int64_t)b->cuCount * perCU * b->launch.gridScale;
        const int64_t need = ((int64_t)b->roundPaths + chunk - 1) / chunk * 64 / block + 1;
        return (unsigned)std::max<int64_t>(1, std::min(want, need));
}

template <class T>
static bool upload(T **dst, const T *src, size_t count, const char *what) {
        if (!hip_ok(hipMalloc(dst, std::max<size_t>(count, 1) * sizeof(T)), what)) return false;
        return count == 0 || hip_ok(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice), what);
}

// Enqueues one round's launch sequence on `st`.
static bool enqueue_round(Batch *b, hipStream_t st) {
        const unsigned chunk = (unsigned)b->launch.chunk;
        const int block = b->launch.block;
        auto kern = pick_set(b->bsdfSet, k_batch_bounce<0, false>, k_batch_bounce<1, false>, k_batch_bounce<2, false>);
        auto kdrain = pick_set(b->bsdfSet, k_batch_bounce<0, true>, k_batch_bounce<1, true>, k_batch_bounce<2, true>);
        const int drainLevel = b->launch.drainLevel;
        int perCU = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, block, 0) != hipSuccess || perCU < 1) perCU = 1;
        const int64_t want = (int64_t)b->cuCount * perCU * b->launch.gridScale;
        const int64_t need = ((int64_t)b->roundPaths + chunk - 1) / chunk * 64 / block + 1;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min(want, need));
        const unsigned camBlocks = (unsigned)((b->roundPaths + kBatchBlock - 1) / kBatchBlock);
        const unsigned filmBlocks = (unsigned)((b->maxRoundPixels + kBatchBlock - 1) / kBatchBlock);
        hipLaunchKernelGGL(k_batch_camera, dim3(camBlocks), dim3(kBatchBlock), 0, st, b->dCursor, b->dRounds, b->dItems,
                           b->dCams, b->sobol, b->dPaths);
        hipLaunchKernelGGL(k_batch_clear, dim3(1), dim3(256), 0, st, b->ctr);
        if (b->launch.queued == 3) {
                hipLaunchKernelGGL(k_bw_clear, dim3(1), dim3(256), 0, st, b->wctr);
                auto kd = k_bw_phase<kBsdfDiffuse>;
                auto ka = b->bsdfSet == kBsdfNoLayer ? k_bw_phase<kBsdfNoLayer> : k_bw_phase<kBsdfAll>;
                // the ray kernels take 64-entry chunks; the shade and finish kernels chunks of at
                // least 256 entries (their work per entry is short next to the per-chunk item lookup
                // and queue appends)
                const unsigned pchunk = std::max(256u, chunk);
                const unsigned gt = batch_grid(b, k_bw_rays<0>, block, 64u),
                               gsplit = (unsigned)std::max<int64_t>(
                                       1, std::min<int64_t>((int64_t)b->cuCount * 8,
                                                            ((int64_t)b->roundPaths + kSplitTile - 1) / kSplitTile + 64)),
                               gd = batch_grid(b, kd, block, pchunk), ga = batch_grid(b, ka, block, pchunk),
                               gany = batch_grid(b, k_bw_rays<1>, block, 64u),
                               gclose = batch_grid(b, k_bw_rays<2>, block, 64u),
                               gfin = batch_grid(b, k_bw_finish, block, pchunk);
                const int *c = b->dCursor;
                for (int level = 0; level < b->levels; ++level) {
                        unsigned *lbase = b->ctr.base + (size_t)level * (b->ctr.maxItems + 1);
                        if (drainLevel > 0 && level >= drainLevel) {
                                hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, st, c, b->dRounds, b->dItems, b->ctr,
                                                   level, chunk, lbase, -1);
                                hipLaunchKernelGGL(kdrain, dim3(grid), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs,
                                                   b->dPaths, b->dQ[level & 1], b->dQ[(level + 1) & 1], b->ctr, level,
                                                   chunk);
                                break;
                        }
                        // the small scenes' items: the bounce kernel; the others: the wavefront kernels
                        if (b->anyMega) {
                                hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, st, c, b->dRounds, b->dItems, b->ctr,
                                                   level, chunk, lbase, 0);
                                hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs,
                                                   b->dPaths, b->dQ[level & 1], b->dQ[(level + 1) & 1], b->ctr, level,
                                                   chunk);
                        }
                        if (!b->anyWave) continue;
                        hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, st, c, b->dRounds, b->dItems, b->ctr, level,
                                           64u, b->wctr.baseOf(level, kWLevel), 1);
                        hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, st, c, b->dRounds, b->dItems, b->ctr, level,
                                           kSplitTile, b->wctr.baseOf(level, kWLevelSplit), 1);
                        const unsigned *nLevel = b->ctr.n + (size_t)level * b->ctr.maxItems;
                        hipLaunchKernelGGL(k_bw_rays<0>, dim3(gt), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs,
                                           b->dPaths, b->dQ[level & 1], b->wctr.baseOf(level, kWLevel), nLevel, b->ws,
                                           level);
                        hipLaunchKernelGGL(k_bw_split, dim3(gsplit), dim3(kMisBlock), 0, st, c, b->dRounds, b->dItems,
                                           b->dArgs, b->dQ[level & 1], b->ctr, b->wctr, b->ws, level);
                        hipLaunchKernelGGL(k_bw_plan, dim3(1), dim3(64), 0, st, c, b->dRounds, b->wctr, level,
                                           (1 << kWDiff) | (1 << kWAll), pchunk);
                        hipLaunchKernelGGL(kd, dim3(gd), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs, b->dPaths,
                                           b->dQ[(level + 1) & 1], b->ctr, b->wctr, b->ws, level, pchunk);
                        if (b->all)
                                hipLaunchKernelGGL(ka, dim3(ga), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs,
                                                   b->dPaths, b->dQ[(level + 1) & 1], b->ctr, b->wctr, b->ws, level, pchunk);
                        hipLaunchKernelGGL(k_bw_plan, dim3(1), dim3(64), 0, st, c, b->dRounds, b->wctr, level,
                                           (1 << kWShadow) | (1 << kWBsdf), 64u);
                        hipLaunchKernelGGL(k_bw_rays<1>, dim3(gany), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs,
                                           b->dPaths, b->ws.q[kWShadow], b->wctr.baseOf(level, kWShadow),
                                           b->wctr.countOf(level, kWShadow), b->ws, level);
                        hipLaunchKernelGGL(k_bw_rays<2>, dim3(gclose), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs,
                                           b->dPaths, b->ws.q[kWBsdf], b->wctr.baseOf(level, kWBsdf),
                                           b->wctr.countOf(level, kWBsdf), b->ws, level);
                        hipLaunchKernelGGL(k_bw_finish, dim3(gfin), dim3(block), 0, st, c, b->dRounds, b->dItems, b->dArgs,
                                           b->dPaths, b->wctr, b->ws, level, (int)kWDiff, pchunk);
                        if (b->all)
                                hipLaunchKernelGGL(k_bw_finish, dim3(gfin), dim3(block), 0, st, c, b->dRounds, b->dItems,
                                                   b->dArgs, b->dPaths, b->wctr, b->ws, level, (int)kWAll, pchunk);
                }
                hipLaunchKernelGGL(k_batch_film, dim3(filmBlocks), dim3(kBatchBlock), 0, st, b->dCursor, b->dRounds,
                                   b->dItems, b->dFilms, b->dPaths);
                hipLaunchKernelGGL(k_batch_next, dim3(1), dim3(64), 0, st, b->dCursor);
                return hip_ok(hipGetLastError(), "batch round launch");
        }
        for (int level = 0; level < b->levels; ++level) {
                hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, st, b->dCursor, b->dRounds, b->dItems, b->ctr,
                                   level, chunk, b->ctr.base + (size_t)level * (b->ctr.maxItems + 1), -1);
                const bool drain = drainLevel > 0 && level >= drainLevel;
                hipLaunchKernelGGL(drain ? kdrain : kern, dim3(grid), dim3(block), 0, st, b->dCursor, b->dRounds,
                                   b->dItems, b->dArgs, b->dPaths, b->dQ[level & 1], b->dQ[(level + 1) & 1], b->ctr,
                                   level, chunk);
                if (drain) break;
        }
        hipLaunchKernelGGL(k_batch_film, dim3(filmBlocks), dim3(kBatchBlock), 0, st, b->dCursor, b->dRounds, b->dItems,
                           b->dFilms, b->dPaths);
        hipLaunchKernelGGL(k_batch_next, dim3(1), dim3(64), 0, st, b->dCursor);
        return hip_ok(hipGetLastError(), "batch round launch");
}

static bool capture(Batch *b) {
        release_graph(b);
        if (!b->captureStream &&
            !hip_ok(hipStreamCreateWithFlags(&b->captureStream, hipStreamNonBlocking), "batch capture stream"))
                return false;
        if (!hip_ok(hipStreamBeginCapture(b->captureStream, hipStreamCaptureModeThreadLocal), "begin capture"))
                return false;
        const bool ok = enqueue_round(b, b->captureStream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->captureStream, &g);
        if (!ok || !hip_ok(e, "end capture")) {
                if (g) (void)hipGraphDestroy(g);
                return false;
        }
        b->graph = g;
        if (!hip_ok(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate")) {
                release_graph(b);
                return false;
        }
        return true;
}

}  // namespace gz

using namespace gz;

extern "C" void *gz_batch_create(const GzBatchScene *scenes, int32_t sceneCount, int64_t roundPaths,
                                 int64_t soloPaths) {
        clear_error();
        if (!scenes || sceneCount <= 0 || roundPaths <= 0 || roundPaths > 0x7fffffffLL || soloPaths < 0) {
                set_error("gz_batch_create: need scenes, 0 < roundPaths < 2^31 and soloPaths >= 0");
                return nullptr;
        }
        Batch *b = new Batch;
        b->nScenes = sceneCount;
        b->roundPaths = (int)roundPaths;
        std::vector<MisArgs> args((size_t)sceneCount);
        std::vector<GzCamera> cams((size_t)sceneCount);
        std::vector<BatchFilm> films((size_t)sceneCount);
        std::vector<char> waveScene((size_t)sceneCount, 0), soloScene((size_t)sceneCount, 0);
        int maxDepth = 0;
        b->misGeneration.assign((size_t)sceneCount, 0);
        for (int32_t s = 0; s < sceneCount; ++s) {
                const GzBatchScene &bs = scenes[s];
                Scene *sc = static_cast<Scene *>(bs.handle);
                if (!sc) {
                        set_error("gz_batch_create: scene %d has no handle", s);
                        free_batch(b);
                        return nullptr;
                }
                b->scenes.push_back(sc);
                if (!camera_ok(&bs.camera, "gz_batch_create")) {
                        free_batch(b);
                        return nullptr;
                }
                if (!bs.d_light || !bs.d_weight || !bs.d_albedo || !bs.d_normal) {
                        set_error("gz_batch_create: scene %d lacks a film image", s);
                        free_batch(b);
                        return nullptr;
                }
                if (bs.camera.spp > b->roundPaths) {
                        set_error("gz_batch_create: roundPaths %d below scene %d's spp %d", b->roundPaths, s,
                                  bs.camera.spp);
                        free_batch(b);
                        return nullptr;
                }
                std::lock_guard<std::mutex> lock(sc->mu);
                if (s == 0) {
                        b->device = sc->device;
                        b->cuCount = sc->cuCount;
                        if (!enter_device(sc) || !hip_ok(hipMalloc((void **)&b->dErr, sizeof(int)), "batch error flag") ||
                            !hip_ok(hipMemset(b->dErr, 0, sizeof(int)), "batch error flag")) {
                                free_batch(b);
                                return nullptr;
                        }
                } else if (sc->device != b->device) {
                        set_error("gz_batch_create: scene %d is on device %d, scene 0 on %d", s, sc->device, b->device);
                        free_batch(b);
                        return nullptr;
                }
                b->misGeneration[(size_t)s] = sc->misGeneration;
                if (!enter_device(sc) || !mis_args(sc, &bs.params, b->dErr, args[(size_t)s])) {
                        free_batch(b);
                        return nullptr;
                }
                cams[(size_t)s] = bs.camera;
                films[(size_t)s] = BatchFilm{bs.d_light,       bs.d_weight,         bs.d_albedo,  bs.d_normal,
                                             bs.iso,           bs.maxComponentValue, bs.camera.resX, bs.camera.resY,
                                             bs.camera.spp,    0};
                b->params.push_back(bs.params);
                const int64_t paths = crop_paths(&bs.camera);
                if (soloPaths > 0 && paths > soloPaths) {
                        soloScene[(size_t)s] = 1;
                        b->solo.push_back(s);
                        b->soloSlots = std::max(b->soloSlots, std::min<int64_t>(paths, roundPaths));
                        continue;
                }
                if (!mis_all_diffuse(sc)) b->all = true;
                const int set = mis_bsdf_set(sc);
                if (set == kBsdfAll || (set == kBsdfNoLayer && b->bsdfSet == kBsdfDiffuse)) b->bsdfSet = set;
                // in wavefront rounds, the scenes whose single-scene default is the wavefront take
                // the wavefront kernels, the small all-diffuse ones the bounce kernel
                waveScene[(size_t)s] = default_wavefront(sc);
                if (waveScene[(size_t)s])
                        b->anyWave = true;
                else
                        b->anyMega = true;
                maxDepth = std::max(maxDepth, (int)bs.params.maxDepth);
        }
        b->levels = maxDepth + 1;
        b->hCams = cams;
        b->hFilms = films;
        // Rounds of whole pixels, scenes in order.
        int used = 0, pixels = 0, firstItem = 0;
        auto close_round = [&]() {
                if (used == 0) return;
                b->hRounds.push_back(BatchRound{firstItem, (int)b->hItems.size() - firstItem, used, pixels});
                b->maxRoundPixels = std::max(b->maxRoundPixels, pixels);
                b->maxItems = std::max(b->maxItems, (int)b->hItems.size() - firstItem);
                firstItem = (int)b->hItems.size();
                used = 0;
                pixels = 0;
        };
        for (int32_t s = 0; s < sceneCount; ++s) {
                const GzCamera &c = cams[(size_t)s];
                b->totalPaths += crop_paths(&c);
                if (soloScene[(size_t)s]) continue;
                const int64_t cropPixels = crop_paths(&c) / c.spp;
                int64_t done = 0;
                while (done < cropPixels) {
                        const int64_t room = (b->roundPaths - used) / c.spp;
                        if (room == 0) {
                                close_round();
                                continue;
                        }
                        const int64_t take = std::min(room, cropPixels - done);
                        b->hItems.push_back(BatchItem{s, (int)(take * c.spp), done * c.spp, used, pixels,
                                                      waveScene[(size_t)s] ? 1 : 0, 0});
                        used += (int)(take * c.spp);
                        pixels += (int)take;
                        done += take;
                }
        }
        close_round();
        // slots for the largest round only: a batch smaller than roundPaths allocates what it uses
        int maxRound = 1;
        for (const BatchRound &r : b->hRounds) maxRound = std::max(maxRound, (int)r.total);
        b->roundPaths = std::min(b->roundPaths, maxRound);
        b->pathSlots = std::max<int64_t>(b->roundPaths, b->soloSlots);
        b->ctr.maxItems = std::max(b->maxItems, 1);
        b->ctr.levels = b->levels;
        const size_t slots = (size_t)b->roundPaths;
        int cur = -1;
        if (!hip_ok(hipGetDevice(&cur), "get device") || (cur != b->device && !hip_ok(hipSetDevice(b->device), "set device")) ||
            !(b->sobol = device_sobol_table()) || !upload(&b->dRounds, b->hRounds.data(), b->hRounds.size(), "batch rounds") ||
            !upload(&b->dItems, b->hItems.data(), b->hItems.size(), "batch items") ||
            !upload(&b->dArgs, args.data(), args.size(), "batch scene arguments") ||
            !upload(&b->dCams, cams.data(), cams.size(), "batch cameras") ||
            !upload(&b->dFilms, films.data(), films.size(), "batch films") ||
            !hip_ok(hipMalloc(&b->dCursor, sizeof(int)), "batch cursor") ||
            !hip_ok(hipMalloc(&b->dPaths, (size_t)b->pathSlots * sizeof(GzPath)), "batch paths") ||
            !hip_ok(hipMalloc(&b->dQ[0], slots * sizeof(int)), "batch queue") ||
            !hip_ok(hipMalloc(&b->dQ[1], slots * sizeof(int)), "batch queue") ||
            !hip_ok(hipMalloc(&b->ctr.n, (size_t)(b->levels + 1) * b->ctr.maxItems * sizeof(unsigned)), "batch counters") ||
            !hip_ok(hipMalloc(&b->ctr.head, (size_t)b->levels * sizeof(unsigned)), "batch counters") ||
            !hip_ok(hipMalloc(&b->ctr.base, (size_t)b->levels * (b->ctr.maxItems + 1) * sizeof(unsigned)),
                    "batch counters")) {
                free_batch(b);
                return nullptr;
        }
        // Default launch: wavefront rounds when some scene is large enough for the wavefront kernels.
        if (b->anyWave) {
                if (!ensure_wave(b)) {
                        free_batch(b);
                        return nullptr;
                }
                b->launch = GzMisLaunch{3, 256, 1, 64, 3, {0, 0, 0}};
        }
        // Lanes of the scenes that render alone: each scene to the lane with the fewest paths so
        // far (lane 0 starts with the rounds' paths, which it runs first; ties to the lower lane),
        // except that a scene sharing a film buffer with an earlier solo scene joins that scene's
        // lane, so that no two lanes add to one film.
        int64_t load[kSoloLanes] = {b->totalPaths};
        for (const int k : b->solo) load[0] -= crop_paths(&b->hCams[(size_t)k]);
        for (size_t j = 0; j < b->solo.size(); ++j) {
                const BatchFilm &f = b->hFilms[(size_t)b->solo[j]];
                int lane = 0;
                for (int l = 1; l < kSoloLanes; ++l)
                        if (load[l] < load[lane]) lane = l;
                for (size_t i = 0; i < j; ++i) {
                        const BatchFilm &g = b->hFilms[(size_t)b->solo[i]];
                        const float *fp[4] = {f.light, f.weight, f.albedo, f.normal};
                        const float *gp[4] = {g.light, g.weight, g.albedo, g.normal};
                        bool shared = false;
                        for (const float *x : fp)
                                for (const float *y : gp) shared = shared || x == y;
                        if (shared) lane = b->soloLane[i];
                }
                b->soloLane.push_back(lane);
                load[lane] += crop_paths(&b->hCams[(size_t)b->solo[j]]);
        }
        for (int l = 1; l < kSoloLanes; ++l) {
                if (std::find(b->soloLane.begin(), b->soloLane.end(), l) == b->soloLane.end()) continue;
                if ((!b->fork && !hip_ok(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "batch fork event")) ||
                    !hip_ok(hipStreamCreateWithFlags(&b->side[l - 1], hipStreamNonBlocking), "batch side stream") ||
                    !hip_ok(hipEventCreateWithFlags(&b->join[l - 1], hipEventDisableTiming), "batch join event") ||
                    !hip_ok(hipMalloc(&b->dPathsSide[l - 1], (size_t)b->soloSlots * sizeof(GzPath)), "batch side paths")) {
                        free_batch(b);
                        return nullptr;
                }
        }
        return b;
}

extern "C" int gz_batch_info(void *batch, GzBatchInfo *info) {
        clear_error();
        if (!batch || !info) {
                set_error("gz_batch_info: bad arguments");
                return GZ_E_INVALID;
        }
        const Batch *b = static_cast<Batch *>(batch);
        info->totalPaths = b->totalPaths;
        info->rounds = (int32_t)b->hRounds.size();
        info->items = (int32_t)b->hItems.size();
        info->roundPaths = b->roundPaths;
        info->maxItemsPerRound = b->maxItems;
        info->levels = b->levels;
        info->allBsdfs = b->all ? 1 : 0;
        info->soloScenes = (int32_t)b->solo.size();
        info->sideScenes = (int32_t)(b->soloLane.size() - std::count(b->soloLane.begin(), b->soloLane.end(), 0));
        return GZ_OK;
}

extern "C" int gz_batch_set_launch(void *batch, const GzMisLaunch *launch) {
        clear_error();
        if (!batch || !launch) {
                set_error("gz_batch_set_launch: bad arguments");
                return GZ_E_INVALID;
        }
        const GzMisLaunch &c = *launch;
        if ((c.queued != 1 && c.queued != 3) || (c.block != 64 && c.block != 128 && c.block != 256) || c.gridScale < 1 ||
            c.gridScale > 16 || c.chunk < 64 || c.chunk > 4096 || c.chunk % 64 != 0 || c.drainLevel < 0 ||
            c.drainLevel > 64 || c.reserved[0] != 0 || c.reserved[1] != 0 || c.reserved[2] != 0) {
                set_error("gz_batch_set_launch: queued %d block %d gridScale %d chunk %d drainLevel %d out of range "
                          "(queued must be 1 or 3)",
                          c.queued, c.block, c.gridScale, c.chunk, c.drainLevel);
                return GZ_E_INVALID;
        }
        Batch *b = static_cast<Batch *>(batch);
        if (c.queued == 3) {
                int cur = -1;
                if (!hip_ok(hipGetDevice(&cur), "get device") ||
                    (cur != b->device && !hip_ok(hipSetDevice(b->device), "set device")) || !ensure_wave(b))
                        return GZ_E_ALLOC;
        }
        b->launch = c;
        release_graph(b);  // captured with the old shape
        return GZ_OK;
}

extern "C" int gz_batch_render_device(void *batch, int32_t useGraph, void *stream) {
        clear_error();
        if (!batch) {
                set_error("gz_batch_render_device: null batch");
                return GZ_E_INVALID;
        }
        Batch *b = static_cast<Batch *>(batch);
        for (int k = 0; k < b->nScenes; ++k) {
                Scene *sc = b->scenes[(size_t)k];
                std::lock_guard<std::mutex> lock(sc->mu);
                if (sc->misGeneration != b->misGeneration[(size_t)k]) {
                        set_error("gz_batch_render_device: scene %d's MIS tables were replaced after gz_batch_create "
                                  "(gz_gpu_set_mis_scene); create the batch again", k);
                        return GZ_E_INVALID;
                }
        }
        int cur = -1;
        if (!hip_ok(hipGetDevice(&cur), "get device") || (cur != b->device && !hip_ok(hipSetDevice(b->device), "set device")))
                return GZ_E_HIP;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        // The side streams (lanes 1-3) start after the work queued on `stream` before this call, and
        // `stream` waits for them at the end, also after a failure. Then the rounds on `stream`,
        // and the scenes that render alone, each with the single-scene launch sequence on its
        // lane's stream, in pieces of whole pixels.
        if (b->fork) {
                if (!hip_ok(hipEventRecord(b->fork, st), "batch fork")) return GZ_E_HIP;
                for (hipStream_t sd : b->side)
                        if (sd && !hip_ok(hipStreamWaitEvent(sd, b->fork, 0), "batch fork")) return GZ_E_HIP;
        }
        int rc = GZ_OK;
        if (!b->hRounds.empty()) {
                if (useGraph && !b->exec && !capture(b)) rc = GZ_E_HIP;
                if (rc == GZ_OK) {
                        hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, b->dCursor);
                        if (!hip_ok(hipGetLastError(), "batch begin")) rc = GZ_E_HIP;
                }
                for (size_t r = 0; rc == GZ_OK && r < b->hRounds.size(); ++r) {
                        if (useGraph) {
                                if (!hip_ok(hipGraphLaunch(b->exec, st), "graph launch")) rc = GZ_E_HIP;
                        } else if (!enqueue_round(b, st)) {
                                rc = GZ_E_HIP;
                        }
                }
        }
        for (size_t j = 0; rc == GZ_OK && j < b->solo.size(); ++j) {
                const int k = b->solo[j], lane = b->soloLane[j];
                const hipStream_t ls = lane ? b->side[lane - 1] : st;
                GzPath *paths = lane ? b->dPathsSide[lane - 1] : b->dPaths;
                const GzCamera &c = b->hCams[(size_t)k];
                const BatchFilm &f = b->hFilms[(size_t)k];
                const int64_t n = crop_paths(&c), piece = b->soloSlots / c.spp * c.spp;
                for (int64_t first = 0; rc == GZ_OK && first < n; first += piece) {
                        const int64_t count = std::min(piece, n - first);
                        rc = gz_gpu_camera_paths_device(&c, first, count, paths, ls);
                        if (rc == GZ_OK)
                                rc = mis_trace_shared(b->scenes[(size_t)k], paths, count, &b->params[(size_t)k], ls,
                                                      &b->soloQueues[lane], b->dErr);
                        if (rc == GZ_OK)
                                rc = gz_gpu_film_add_device(&c, paths, count, f.iso, f.maxComponent, f.light, f.weight,
                                                            f.albedo, f.normal, ls);
                }
        }
        for (int l = 0; l < kSoloLanes - 1; ++l) {
                if (!b->side[l]) continue;
                if (!hip_ok(hipEventRecord(b->join[l], b->side[l]), "batch join") ||
                    !hip_ok(hipStreamWaitEvent(st, b->join[l], 0), "batch join"))
                        rc = rc == GZ_OK ? GZ_E_HIP : rc;
        }
        return rc;
}

extern "C" int gz_batch_check(void *batch, void *stream) {
        clear_error();
        if (!batch) {
                set_error("gz_batch_check: null batch");
                return GZ_E_INVALID;
        }
        Batch *b = static_cast<Batch *>(batch);
        int cur = -1;
        if (!hip_ok(hipGetDevice(&cur), "get device") || (cur != b->device && !hip_ok(hipSetDevice(b->device), "set device")))
                return GZ_E_HIP;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        int flag = 0;
        if (!hip_ok(hipMemcpyAsync(&flag, b->dErr, sizeof(int), hipMemcpyDeviceToHost, st), "read batch error flag") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        if (!flag) return GZ_OK;
        if (!hip_ok(hipMemsetAsync(b->dErr, 0, sizeof(int), st), "clear batch error flag") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        if (flag & 1) {
                set_error("batch traversal stack overflow: a ray needed more than %d entries", GZ_STACK_CAPACITY);
                return GZ_E_OVERFLOW;
        }
        set_error("batch: a shading kernel met a primitive, mesh or vertex index out of range");
        return GZ_E_INVALID;
}

extern "C" void gz_batch_free(void *batch) { free_batch(static_cast<Batch *>(batch)); }
