// Wavefront MIS path tracing on gfx950: the kernels and entry points of one scene's trace,
// bounce and path-replay adjoint. The bounce itself is in mis_device.h.
#include <algorithm>
#include <cmath>
#include <climits>
#include <cstring>
#include <functional>
#include <map>
#include <vector>

#include "mis_device.h"

#pragma clang fp contract(off)

namespace gz {

template <bool kStats, int kAll>
__global__ __launch_bounds__(kMisBlock) void k_mis_bounce(MisArgs a, GzPath *__restrict__ paths, int64_t n,
                                                          uint32_t *__restrict__ stats) {
        const int64_t i = (int64_t)blockIdx.x * kMisBlock + threadIdx.x;
        if (i < n) bounce_one<kStats, kAll, false>(a, paths, i, stats, AdjArgs{nullptr, nullptr}, nullptr);
}

// The replayed bounce with gradients; per-block LDS accumulators when the material table fits
// (one global atomic per touched entry per block instead of one per vertex).
template <int kAll>
__global__ __launch_bounds__(kMisBlock) void k_mis_adjoint(MisArgs a, GzPath *__restrict__ paths, int64_t n,
                                                           AdjArgs adj) {
        __shared__ double sg[kAdjLdsMats * 4];
        const bool useLds = a.nMats <= kAdjLdsMats;
        if (useLds) {
                for (int k = threadIdx.x; k < a.nMats * 4; k += kMisBlock) sg[k] = 0.0;
                __syncthreads();
        }
        const int64_t i = (int64_t)blockIdx.x * kMisBlock + threadIdx.x;
        if (i < n) bounce_one<false, kAll, true>(a, paths, i, nullptr, adj, useLds ? sg : nullptr);
        if (useLds) {
                __syncthreads();
                for (int k = threadIdx.x; k < a.nMats * 4; k += kMisBlock) {
                        const double v = sg[k];
                        if (v != 0.0) atomicAdd(&adj.grad[k], v);
                }
        }
}

// ---- Device path queues --------------------------------------------------------------------
//
// Bounce b of a trace runs over the queue of the paths still active after bounce b - 1 (bounce 0
// over all paths). A persistent wave takes `chunk` queue entries per atomic on the level's head,
// bounces them 64 at a time, and appends the survivors to the next level's queue: a ballot of
// the lanes that go on, a prefix popcount for each lane's slot and one atomic per wave for the
// base. A wave's survivors keep their relative order, so neighbouring pixels stay together.
// The levels' counters live in device memory and are cleared by a kernel, so a whole trace is
// a fixed sequence of launches with no host round trip (capturable in a graph).
struct QueueLevel {
        unsigned n;     // entries appended to this level's queue
        unsigned head;  // entries of this level handed out to waves
};

__global__ void k_queue_clear(QueueLevel *ctl, int levels) {
        for (int k = threadIdx.x; k < levels; k += blockDim.x) {
                ctl[k].n = 0u;
                ctl[k].head = 0u;
        }
}

// Runs body(i) for every entry i of a queue of `nin` entries (the identity 0 .. nin-1 when qin is
// null), taken `chunk` (a multiple of 64) at a time through the atomic `head` (null: wave w takes
// chunks w, w + waves, w + 2 waves, ... with no atomic); body returns a mask
// of the kQ output queues i goes to, appended through their atomic counters outN after every four
// runs of 64. Every lane of a wave reaches each ballot: the loop bounds are wave-uniform.
template <int kQ, typename Body>
__device__ __forceinline__ void queue_run_n(unsigned nin, const int *__restrict__ qin, unsigned *head, unsigned chunk,
                                            int *const (&qout)[kQ], unsigned *const (&outN)[kQ], Body &&body) {
        const int lane = (int)(threadIdx.x & 63);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        Keep4 keep;
        unsigned bits[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) bits[q] = 0u;
        int runs = 0;
        unsigned c = 0, end = 0;
        const uint64_t nWaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
        uint64_t nextChunk = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        for (;;) {
                if (c >= end) {
                        uint64_t t = 0;
                        if (head) {
                                unsigned h = 0;
                                if (lane == 0) h = atomicAdd(head, chunk);
                                t = __shfl(h, 0);
                        } else {
                                t = nextChunk * chunk;
                                nextChunk += nWaves;
                        }
                        if (t >= nin) break;
                        c = (unsigned)t;
                        end = nin - c < chunk ? nin : c + chunk;
                }
                const unsigned j = c + (unsigned)lane;
                int i = -1;
                if (j < end) i = qin ? qin[j] : (int)j;
                const unsigned mask = i >= 0 ? (unsigned)body(i) : 0u;
                keep.set(runs, i);
#pragma unroll
                for (int q = 0; q < kQ; ++q) bits[q] |= ((mask >> q) & 1u) << runs;
                ++runs;
                c += 64;
                if (runs == 4) {
                        flush_appends<kQ>(keep, bits, runs, qout, outN, lane, below);
                        runs = 0;
#pragma unroll
                        for (int q = 0; q < kQ; ++q) bits[q] = 0u;
                }
        }
        if (runs > 0) flush_appends<kQ>(keep, bits, runs, qout, outN, lane, below);
}

// One output queue: body returns whether i goes on.
template <typename Body>
__device__ __forceinline__ void queue_run(unsigned nin, const int *__restrict__ qin, unsigned *head,
                                          int *__restrict__ qout, unsigned *outN, unsigned chunk, Body &&body) {
        int *const qo[1] = {qout};
        unsigned *const on[1] = {outN};
        queue_run_n<1>(nin, qin, head, chunk, qo, on, [&](int i) { return body(i) ? 1u : 0u; });
}

// The level form: entries of level `level` (identity at level 0), survivors to level + 1.
template <typename Body>
__device__ __forceinline__ void queue_loop(unsigned count, const int *__restrict__ qin, int *__restrict__ qout,
                                           QueueLevel *ctl, int level, unsigned chunk, Body &&body) {
        const unsigned nin = qin ? ctl[level].n : count;
        queue_run(nin, qin, &ctl[level].head, qout, &ctl[level + 1].n, chunk, body);
}

// ---- Split wavefront: one trace kernel and two shade kernels per bounce level ----------------
//
// k_wf_trace: persistent lanes with refill trace the closest hit of every path of the level's
// queue (idle lanes take new paths from the queue when `refill` of them are idle) and store it;
// k_wf_split sorts the paths by the BSDF their hit shades with into the diffuse queue (misses and
// stopped paths too) or the queue of the other BSDFs. k_wf_shade<kAll> then runs the rest of the
// bounce (bounce_rest) over one class queue: the diffuse kernel is the small one (4 waves/SIMD),
// only the other class pays for the 256-register kernel. Survivors go to the next level's queue.
struct WfLevel {
        unsigned qn, qhead;    // the level's queue (appended by the previous level's shade kernels)
        unsigned dn, dhead;    // diffuse class
        unsigned an, ahead;    // other BSDFs
        unsigned sn, mn;       // shadow rays and BSDF rays (full wavefront; taken by static striding)
        unsigned nn;           // BSDFs but the layered one (full wavefront)
        unsigned pad[7];
};
static_assert(sizeof(WfLevel) == 64, "WfLevel");

// Marks every entry of a replay record as not traced (the forward pass then fills the rays it traces).
__global__ __launch_bounds__(kBlock) void k_record_clear(GzHit *__restrict__ hits, GzHit *__restrict__ mhits,
                                                         uint8_t *__restrict__ occ, int64_t n) {
        for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
                hits[j] = GzHit{kRecNone, 0.0f, 0.0f, 0.0f};
                mhits[j] = GzHit{kRecNone, 0.0f, 0.0f, 0.0f};
                occ[j] = kRecNoOcc;
        }
}

__global__ void k_wf_clear(WfLevel *ctl, int levels) {
        for (int k = threadIdx.x; k < levels; k += blockDim.x) {
                WfLevel z;
                unsigned *u = reinterpret_cast<unsigned *>(&z);
                for (int j = 0; j < 16; ++j) u[j] = 0u;
                ctl[k] = z;
        }
}

template <bool kStats>
__global__ __launch_bounds__(kBlock) void k_wf_trace(MisArgs a, const GzPath *__restrict__ paths, unsigned count,
                                                     const int *__restrict__ qin, WfLevel *ctl, int level,
                                                     GzHit *__restrict__ hits, int refill, unsigned chunk,
                                                     uint32_t *__restrict__ stats) {
        Stack<kLdsMax> S;
        Lane L;
        int cur = -1;       // path index in flight on this lane
        bool done = false;  // this lane finished a path this step
        unsigned resBase = 0, resCount = 0;
        bool counterDone = false, exhausted = false;
        const int laneId = (int)(threadIdx.x & 63);
        const uint64_t below = laneId == 0 ? 0ull : (~0ull >> (64 - laneId));
        WfLevel &lv = ctl[level];
        const unsigned nin = qin ? lv.qn : count;
        for (;;) {
                if (!exhausted) {
                        const uint64_t idle = __ballot(cur < 0);
                        const int nIdle = __popcll(idle);
                        if (nIdle >= refill || nIdle == 64) {
                                const unsigned rank = (unsigned)__popcll(idle & below);
                                const unsigned need = (unsigned)nIdle;
                                unsigned fresh = 0, take = 0;
                                if (resCount < need && !counterDone) {
                                        take = need - resCount > chunk ? need - resCount : chunk;
                                        const int leader = __ffsll((unsigned long long)idle) - 1;
                                        if (laneId == leader) fresh = atomicAdd(&lv.qhead, take);
                                        fresh = __shfl(fresh, leader);
                                        if (fresh >= nin || nin - fresh <= take) counterDone = true;
                                }
                                if (cur < 0) {
                                        unsigned mine = 0xffffffffu;
                                        if (rank < resCount)
                                                mine = resBase + rank;
                                        else if (rank - resCount < take)
                                                mine = fresh + (rank - resCount);
                                        if (mine < nin) {
                                                cur = qin ? qin[mine] : (int)mine;
                                                const GzPath &p = paths[cur];
                                                if (path_live(a, p)) {
                                                        lane_init(L, p.ray, INFINITY, a.rootDesc, a.sceneSafe);
                                                        done = a.empty;  // no nodes: a miss, nothing to step
                                                } else {
                                                        L.hitPrim = -1;
                                                        L.nInterior = L.nTris = 0;
                                                        done = true;  // stopped: the shade kernel clears it
                                                }
                                        }
                                }
                                if (take > 0) {
                                        const unsigned used = need - resCount;
                                        resBase = fresh + used;
                                        resCount = take - used;
                                } else {
                                        const unsigned used = need < resCount ? need : resCount;
                                        resBase += used;
                                        resCount -= used;
                                }
                                if (counterDone && (resCount == 0 || resBase >= nin)) exhausted = true;
                        }
                }
                if (cur >= 0 && !done) {
                        const int st = step<false, kStats, kLdsMax>(L, S, a.nodes, a.tris);
                        if (st != kStepGo) {
                                if (st == kStepOverflow) {
                                        atomicOr(a.err, 1);
                                        L.hitPrim = -1;
                                }
                                done = true;
                        }
                }
                // a finished path: store its hit (sorted into the class queues by k_wf_split)
                if (cur >= 0 && done) {
                        GzHit h;
                        h.prim = L.hitPrim;
                        h.t = L.hitPrim >= 0 ? L.tBest : INFINITY;
                        h.u = L.hitPrim >= 0 ? L.bu : 0.0f;
                        h.v = L.hitPrim >= 0 ? L.bv : 0.0f;
                        hits[cur] = h;
                        if (kStats) {
                                stats[3 * (size_t)cur] += L.nInterior;
                                stats[3 * (size_t)cur + 1] += L.nTris;
                        }
                        cur = -1;
                        done = false;
                }
                if (exhausted && __ballot(cur >= 0) == 0ull) break;
        }
}

// The level's queue sorted by the BSDF class of its hits into the two class queues. A block takes
// tiles of kSplitPer x 256 entries (each wave kSplitPer runs of 64) and appends each class with
// one atomic per tile: counters that every wave of the chip hits are the kernel's cost, not the
// bytes it moves.
constexpr int kSplitPer = 8;

// classes: 1 every path to the diffuse queue (scenes without other materials), 2 the diffuse and
// the all-BSDF queues, 3 the diffuse, all-BSDF (layered) and other-BSDF queues.
__global__ __launch_bounds__(kMisBlock) void k_wf_split(MisArgs a, unsigned count, const int *__restrict__ qin,
                                                        WfLevel *ctl, int level, const GzHit *__restrict__ hits,
                                                        int *__restrict__ qd, int *__restrict__ qa, int *__restrict__ qn,
                                                        int classes) {
        constexpr int kWaves = kMisBlock / 64;
        constexpr size_t kTile = (size_t)kMisBlock * kSplitPer;
        __shared__ unsigned sCnt[kWaves][3];
        __shared__ unsigned sBase[3];
        WfLevel &lv = ctl[level];
        unsigned *counts[3] = {&lv.dn, &lv.an, &lv.nn};
        int *queues[3] = {qd, qa, qn};
        const size_t nin = qin ? lv.qn : count;
        const int lane = (int)(threadIdx.x & 63);
        const int wave = (int)(threadIdx.x >> 6);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        for (size_t t0 = (size_t)blockIdx.x * kTile; t0 < nin; t0 += (size_t)gridDim.x * kTile) {
                const size_t w0 = t0 + (size_t)wave * 64 * kSplitPer;
                int idx[kSplitPer];
                unsigned bits[3] = {0u, 0u, 0u};
                unsigned cnt[3] = {0u, 0u, 0u};
#pragma unroll
                for (int k = 0; k < kSplitPer; ++k) {
                        const size_t j = w0 + (size_t)k * 64 + (size_t)lane;
                        int i = -1;
                        if (j < nin) i = qin ? qin[j] : (int)j;
                        idx[k] = i;
                        int cls = -1;
                        if (i >= 0) {
                                const int set = classes > 1 ? hit_bsdf_set(a, hits[i].prim) : kBsdfDiffuse;
                                cls = set == kBsdfDiffuse ? 0 : (set == kBsdfNoLayer && classes == 3) ? 2 : 1;
                        }
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                                bits[c] |= (unsigned)(cls == c) << k;
                                cnt[c] += (unsigned)__popcll(__ballot(cls == c));
                        }
                }
                if (lane == 0)
                        for (int c = 0; c < 3; ++c) sCnt[wave][c] = cnt[c];
                __syncthreads();
                if (threadIdx.x == 0) {
                        for (int c = 0; c < 3; ++c) {
                                unsigned t = 0;
                                for (int w = 0; w < kWaves; ++w) t += sCnt[w][c];
                                sBase[c] = t ? atomicAdd(counts[c], t) : 0u;
                        }
                }
                __syncthreads();
                unsigned off[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                        off[c] = sBase[c];
                        for (int w = 0; w < wave; ++w) off[c] += sCnt[w][c];
                }
#pragma unroll
                for (int k = 0; k < kSplitPer; ++k) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                                const bool mine = (bits[c] >> k) & 1u;
                                const uint64_t m = __ballot(mine);
                                if (mine) queues[c][off[c] + (unsigned)__popcll(m & below)] = idx[k];
                                off[c] += (unsigned)__popcll(m);
                        }
                }
                __syncthreads();  // sCnt and sBase are the next tile's
        }
}

template <bool kStats, int kAll>
__global__ __launch_bounds__(kMisBlock, mis_min_waves(kAll)) void k_wf_shade(MisArgs a, GzPath *__restrict__ paths,
                                                                      const GzHit *__restrict__ hits,
                                                                      const int *__restrict__ qin, WfLevel *ctl,
                                                                      int level, int *__restrict__ qout,
                                                                      unsigned chunk, uint32_t *__restrict__ stats) {
        WfLevel &lv = ctl[level];
        queue_run(kAll ? lv.an : lv.dn, qin, kAll ? &lv.ahead : &lv.dhead, qout, &ctl[level + 1].qn, chunk,
                  [&](int i) {
                          const GzPath p = paths[i];
                          if (!(p.flags & GZ_PATH_ACTIVE)) return false;
                          if (p.bounce > a.maxDepth) {
                                  paths[i].flags = p.flags & ~GZ_PATH_ACTIVE;
                                  return false;
                          }
                          const GzHit h = hits[i];
                          Work w;
                          return bounce_rest<kStats, kAll, false>(a, paths, i, p, h.prim >= 0, h.prim, h.u, h.v, w,
                                                                  stats, AdjArgs{nullptr, nullptr}, nullptr);
                  });
}

template <bool kStats, int kAll>
__global__ __launch_bounds__(kMisBlock, mis_min_waves(kAll)) void k_mis_queue(MisArgs a, GzPath *__restrict__ paths, unsigned count,
                                                         const int *__restrict__ qin, int *__restrict__ qout,
                                                         QueueLevel *ctl, int level, unsigned chunk,
                                                         uint32_t *__restrict__ stats) {
        queue_loop(count, qin, qout, ctl, level, chunk, [&](int i) {
                return bounce_one<kStats, kAll, false>(a, paths, i, stats, AdjArgs{nullptr, nullptr}, nullptr);
        });
}

// The remaining bounces of every path of one level's queue in one launch: a lane runs its path's
// bounces until it stops (bounce_one returns false once the path is inactive or past maxDepth, so
// at most maxDepth + 1 of them). For the last levels, where the few paths that Russian roulette
// leaves cost less than one launch sequence per level.
template <bool kStats, int kAll>
__global__ __launch_bounds__(kMisBlock, mis_min_waves(kAll)) void k_mis_drain(MisArgs a, GzPath *__restrict__ paths,
                                                                       const int *__restrict__ qin, const unsigned *nptr,
                                                                       unsigned *head, unsigned chunk,
                                                                       uint32_t *__restrict__ stats) {
        const unsigned nin = *nptr;
        const int lane = (int)(threadIdx.x & 63);
        for (;;) {
                unsigned c = 0;
                if (lane == 0) c = atomicAdd(head, chunk);
                c = __shfl(c, 0);
                if (c >= nin) break;
                const unsigned end = nin - c < chunk ? nin : c + chunk;
                for (unsigned j = c + (unsigned)lane; j < end; j += 64) {
                        const int i = qin[j];
                        for (int b = 0; b <= a.maxDepth; ++b)
                                if (!bounce_one<kStats, kAll, false>(a, paths, i, stats, AdjArgs{nullptr, nullptr}, nullptr))
                                        break;
                }
        }
}

// The replayed bounce over the same queues; a persistent block flushes its LDS gradient
// accumulators once, when the level is drained.
template <int kAll>
__global__ __launch_bounds__(kMisBlock, kAll == kBsdfAll ? 2 : 3) void k_mis_adjoint_queue(MisArgs a, GzPath *__restrict__ paths,
                                                                 unsigned count, const int *__restrict__ qin,
                                                                 int *__restrict__ qout, QueueLevel *ctl, int level,
                                                                 unsigned chunk, AdjArgs adj) {
        __shared__ double sg[kAdjLdsMats * 4];
        const bool useLds = a.nMats <= kAdjLdsMats;
        if (useLds) {
                for (int k = threadIdx.x; k < a.nMats * 4; k += blockDim.x) sg[k] = 0.0;
                __syncthreads();
        }
        queue_loop(count, qin, qout, ctl, level, chunk, [&](int i) {
                return bounce_one<false, kAll, true>(a, paths, i, nullptr, adj, useLds ? sg : nullptr);
        });
        if (useLds) {
                __syncthreads();
                for (int k = threadIdx.x; k < a.nMats * 4; k += blockDim.x) {
                        const double v = sg[k];
                        if (v != 0.0) atomicAdd(&adj.grad[k], v);
                }
        }
}

// The replayed bounces of the levels from one on in one launch (k_mis_drain's form): a lane
// replays its path's remaining bounces; the block flushes its LDS accumulators once at the end.
template <int kAll>
__global__ __launch_bounds__(kMisBlock, kAll == kBsdfAll ? 2 : 3) void k_mis_adjoint_drain(MisArgs a, GzPath *__restrict__ paths,
                                                                 const int *__restrict__ qin, const unsigned *nptr,
                                                                 unsigned *head, unsigned chunk, AdjArgs adj) {
        __shared__ double sg[kAdjLdsMats * 4];
        const bool useLds = a.nMats <= kAdjLdsMats;
        if (useLds) {
                for (int k = threadIdx.x; k < a.nMats * 4; k += blockDim.x) sg[k] = 0.0;
                __syncthreads();
        }
        const unsigned nin = *nptr;
        const int lane = (int)(threadIdx.x & 63);
        for (;;) {
                unsigned c = 0;
                if (lane == 0) c = atomicAdd(head, chunk);
                c = __shfl(c, 0);
                if (c >= nin) break;
                const unsigned end = nin - c < chunk ? nin : c + chunk;
                for (unsigned j = c + (unsigned)lane; j < end; j += 64) {
                        const int i = qin[j];
                        for (int b = 0; b <= a.maxDepth; ++b)
                                if (!bounce_one<false, kAll, true>(a, paths, i, nullptr, adj, useLds ? sg : nullptr)) break;
                }
        }
        if (useLds) {
                __syncthreads();
                for (int k = threadIdx.x; k < a.nMats * 4; k += blockDim.x) {
                        const double v = sg[k];
                        if (v != 0.0) atomicAdd(&adj.grad[k], v);
                }
        }
}

// ---- Full wavefront: every ray of a bounce traced by a traversal kernel -------------------------
//
// Per level: k_wf_trace (closest hits of the paths' rays), k_wf_split (class queues), then per
// class k_wf_phase (shade_phase: the bounce up to its shadow and BSDF rays, which it appends to
// two ray queues, and the continuation), k_wf_rays<1> and <2> (any hit of the shadow rays,
// closest hit of the BSDF rays, 7 to 8 waves per SIMD) and k_wf_finish over the class queues
// (finish_phase). The shade kernels trace nothing, so every traversal runs at the traversal
// kernels' occupancy.

// The rays of a queue of path indices, 64 per wave and chunk by static striding (no atomic),
// each lane tracing its ray to the end: kMode 0 the paths' own rays (queue of level `level`, the
// identity over `count` at level 0; stopped paths get a miss) into hits, 1 the shadow rays
// `rays[i]` (any hit below 1 - 1e-5) into occ, 2 the BSDF rays (closest hit) into hits. On the
// 5M-triangle frame, per 8M-path batch: shadow rays 2.7 ms here against 4.1 ms with k_wf_trace's
// lane refill, BSDF rays 2.8 against 3.1 ms, but the paths' own rays 7.8 against 6.4 ms; so
// k_wf_trace takes those.
template <int kMode, bool kStats>
__global__ __launch_bounds__(kBlock) void k_wf_rays(MisArgs a, const GzPath *__restrict__ paths, unsigned count,
                                                    const int *__restrict__ qin, const unsigned *nptr,
                                                    const Ray_C *__restrict__ rays, uint8_t *__restrict__ occ,
                                                    GzHit *__restrict__ hits, uint32_t *__restrict__ stats) {
        constexpr bool kAny = kMode == 1;
        Stack<kLdsMax> S;
        const unsigned nin = (kMode == 0 && !qin) ? count : *nptr;
        const int lane = (int)(threadIdx.x & 63);
        const unsigned waves = gridDim.x * (blockDim.x >> 6);
        for (unsigned g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); (size_t)g * 64 < nin; g += waves) {
                const unsigned j = g * 64 + (unsigned)lane;
                if (j >= nin) continue;
                const int idx = qin ? qin[j] : (int)j;
                Lane L;
                L.hitPrim = -1;
                L.nInterior = L.nTris = 0;
                bool trace = true;
                if (kMode == 0) {
                        const GzPath &p = paths[idx];
                        trace = path_live(a, p);
                        if (trace) lane_init(L, p.ray, INFINITY, a.rootDesc, a.sceneSafe);
                } else {
                        lane_init(L, rays[idx], kAny ? 1.0f - 0.00001f : INFINITY, a.rootDesc, a.sceneSafe);
                }
                if (trace && !a.empty) {
                        // BSDF rays stop at a hit closer than the emitter box (bsdf_ray_stop)
                        const float tStop = kMode == 2 ? bsdf_ray_stop(a, rays[idx]) : -1.0f;
                        int st;
                        do {
                                st = step<kAny, kStats, kLdsMax>(L, S, a.nodes, a.tris);
                        } while (st == kStepGo && !(kMode == 2 && L.hitPrim >= 0 && L.tBest < tStop));
                        if (st == kStepOverflow) {
                                atomicOr(a.err, 1);
                                L.hitPrim = -1;
                        }
                }
                if (kAny)
                        occ[idx] = L.hitPrim >= 0 ? 1 : 0;
                else
                        hits[idx] = L.hitPrim >= 0 ? GzHit{L.hitPrim, L.tBest, L.bu, L.bv} : GzHit{-1, INFINITY, 0.0f, 0.0f};
                if (kStats) {
                        stats[3 * (size_t)idx] += L.nInterior;
                        stats[3 * (size_t)idx + 1] += L.nTris;
                }
        }
}

template <bool kStats, int kAll>
__global__ __launch_bounds__(kMisBlock, mis_min_waves(kAll)) void k_wf_phase(
        MisArgs a, GzPath *__restrict__ paths, const GzHit *__restrict__ hits, const int *__restrict__ qin,
        WfLevel *ctl, int level, int *__restrict__ qnext, int *__restrict__ qs, int *__restrict__ qm,
        Pending *__restrict__ pend, Ray_C *__restrict__ srays, Ray_C *__restrict__ mrays, unsigned chunk,
        uint32_t *__restrict__ stats) {
        WfLevel &lv = ctl[level];
        int *const qo[3] = {qnext, qs, qm};
        unsigned *const on[3] = {&ctl[level + 1].qn, &lv.sn, &lv.mn};
        const unsigned nin = kAll == kBsdfAll ? lv.an : kAll == kBsdfNoLayer ? lv.nn : lv.dn;
        queue_run_n<3>(nin, qin, nullptr, chunk, qo, on, [&](int i) {
                const GzPath p = paths[i];
                Pending pd;
                pd.flags = 0;
                unsigned mask = 0;
                if (!(p.flags & GZ_PATH_ACTIVE)) {
                } else if (p.bounce > a.maxDepth) {
                        paths[i].flags = p.flags & ~GZ_PATH_ACTIVE;
                } else {
                        const GzHit h = hits[i];
                        Ray_C sr, mr;
                        const bool go = shade_phase<kStats, kAll>(a, paths, i, p, h.prim >= 0, h.prim, h.u, h.v, pd, sr,
                                                                  mr, stats);
                        if (pd.flags & kPendShadow) srays[i] = sr;
                        if (pd.flags & kPendMis) mrays[i] = mr;
                        mask = (go ? 1u : 0u) | ((pd.flags & kPendShadow) ? 2u : 0u) | ((pd.flags & kPendMis) ? 4u : 0u);
                }
                pend[i] = pd;
                return mask;
        });
}

__global__ __launch_bounds__(kMisBlock) void k_wf_finish(MisArgs a, GzPath *__restrict__ paths,
                                                         const int *__restrict__ qin, const unsigned *nptr,
                                                         const Pending *__restrict__ pend,
                                                         const uint8_t *__restrict__ occ,
                                                         const Ray_C *__restrict__ mrays,
                                                         const GzHit *__restrict__ mhits) {
        const size_t nin = *nptr;
        for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < nin; j += (size_t)gridDim.x * blockDim.x) {
                const int i = qin[j];
                const Pending pd = pend[i];
                if (!(pd.flags & kPendWork)) continue;
                const bool o = (pd.flags & kPendShadow) ? occ[i] != 0 : false;
                const Ray_C mr = (pd.flags & kPendMis) ? mrays[i] : Ray_C{0, 0, 0, 0, 0, 0};
                const GzHit mh = (pd.flags & kPendMis) ? mhits[i] : GzHit{-1, 0.0f, 0.0f, 0.0f};
                finish_phase(a, paths, i, pd, o, mr, mh);
        }
}

// Queue storage of a handle, per stream (grow-only): two ping-pong index arrays and the levels'
// counters. Traces queued on one stream run in order and share it; each stream has its own.
struct MisQueue {
        int *q[2] = {nullptr, nullptr};
        size_t cap = 0;
        QueueLevel *ctl = nullptr;
        int levels = 0;
        // split wavefront: class queues, closest hits, level counters
        int *qd = nullptr, *qa = nullptr;
        GzHit *hits = nullptr;
        size_t wfCap = 0;
        // full wavefront: the other-BSDF class queue, ray queues, pending terms, rays and results
        int *qn = nullptr;
        int *qs = nullptr, *qm = nullptr;
        Pending *pend = nullptr;
        Ray_C *srays = nullptr, *mrays = nullptr;
        uint8_t *occ = nullptr;
        GzHit *mhits = nullptr;
        size_t fullCap = 0;
        WfLevel *wctl = nullptr;
        int wfLevels = 0;
};

struct MisQueues {
        std::map<hipStream_t, MisQueue> perStream;
};

void release_mis_queues(MisQueues *m) {
        if (!m) return;
        for (auto &kv : m->perStream) {
                MisQueue &q = kv.second;
                void *ptrs[] = {q.q[0], q.q[1], q.ctl, q.qd, q.qa, q.hits, q.wctl, q.qn, q.qs, q.qm, q.pend, q.srays,
                                q.mrays, q.occ, q.mhits};
                for (void *p : ptrs)
                        if (p) (void)hipFree(p);
        }
        delete m;
}

// The queue storage of `stream` in `owner` (a handle's own, or one a batch shares between its
// scenes; created on first use) with room for `count` paths and `levels` counters.
static MisQueue *mis_queue(MisQueues *&owner, hipStream_t stream, size_t count, int levels) {
        if (!owner) owner = new MisQueues;
        MisQueue &q = owner->perStream[stream];
        if (q.cap < count) {
                for (int k = 0; k < 2; ++k) {
                        if (q.q[k]) (void)hipFree(q.q[k]);
                        q.q[k] = nullptr;
                }
                q.cap = 0;
                for (int k = 0; k < 2; ++k)
                        if (!hip_ok(hipMalloc(&q.q[k], count * sizeof(int)), "path queue")) return nullptr;
                q.cap = count;
        }
        if (q.levels < levels) {
                if (q.ctl) (void)hipFree(q.ctl);
                q.ctl = nullptr;
                q.levels = 0;
                if (!hip_ok(hipMalloc(&q.ctl, (size_t)levels * sizeof(QueueLevel)), "path queue counters"))
                        return nullptr;
                q.levels = levels;
        }
        return &q;
}

// mis_queue plus the split wavefront's class queues, hits and level counters.
static MisQueue *wf_queue(MisQueues *&owner, hipStream_t stream, size_t count, int levels) {
        MisQueue *q = mis_queue(owner, stream, count, levels);
        if (!q) return nullptr;
        if (q->wfCap < count) {
                void *old[] = {q->qd, q->qa, q->hits};
                for (void *p : old)
                        if (p) (void)hipFree(p);
                q->qd = q->qa = nullptr;
                q->hits = nullptr;
                q->wfCap = 0;
                if (!hip_ok(hipMalloc(&q->qd, count * sizeof(int)), "class queue") ||
                    !hip_ok(hipMalloc(&q->qa, count * sizeof(int)), "class queue") ||
                    !hip_ok(hipMalloc(&q->hits, count * sizeof(GzHit)), "closest hits"))
                        return nullptr;
                q->wfCap = count;
        }
        if (q->wfLevels < levels) {
                if (q->wctl) (void)hipFree(q->wctl);
                q->wctl = nullptr;
                q->wfLevels = 0;
                if (!hip_ok(hipMalloc(&q->wctl, (size_t)levels * sizeof(WfLevel)), "wavefront counters")) return nullptr;
                q->wfLevels = levels;
        }
        return q;
}

// wf_queue plus the full wavefront's ray queues, pending terms, rays and results.
static MisQueue *full_queue(MisQueues *&owner, hipStream_t stream, size_t count, int levels) {
        MisQueue *q = wf_queue(owner, stream, count, levels);
        if (!q) return nullptr;
        if (q->fullCap < count) {
                void *old[] = {q->qn, q->qs, q->qm, q->pend, q->srays, q->mrays, q->occ, q->mhits};
                for (void *p : old)
                        if (p) (void)hipFree(p);
                q->qn = q->qs = q->qm = nullptr;
                q->pend = nullptr;
                q->srays = q->mrays = nullptr;
                q->occ = nullptr;
                q->mhits = nullptr;
                q->fullCap = 0;
                if (!hip_ok(hipMalloc(&q->qn, count * sizeof(int)), "other-BSDF queue") ||
                    !hip_ok(hipMalloc(&q->qs, count * sizeof(int)), "shadow queue") ||
                    !hip_ok(hipMalloc(&q->qm, count * sizeof(int)), "BSDF ray queue") ||
                    !hip_ok(hipMalloc(&q->pend, count * sizeof(Pending)), "pending terms") ||
                    !hip_ok(hipMalloc(&q->srays, count * sizeof(Ray_C)), "shadow rays") ||
                    !hip_ok(hipMalloc(&q->mrays, count * sizeof(Ray_C)), "BSDF rays") ||
                    !hip_ok(hipMalloc(&q->occ, count), "shadow results") ||
                    !hip_ok(hipMalloc(&q->mhits, count * sizeof(GzHit)), "BSDF ray hits"))
                        return nullptr;
                q->fullCap = count;
        }
        return q;
}

// Grid of a persistent queue kernel: gridScale x the blocks resident per CU x CUs, and no more
// blocks than the first level has chunks of work for.
template <typename K>
static unsigned queue_grid(Scene *s, K kern, int block, int64_t count, unsigned chunk) {
        int perCU = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, block, 0) != hipSuccess || perCU < 1) perCU = 1;
        const GzMisLaunch &c = s->misLaunch;
        const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
        const int64_t wavesPerBlock = block / 64;
        const int64_t need = ((count + chunk - 1) / chunk + wavesPerBlock - 1) / wavesPerBlock;
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, need));
}

// Grid of k_wf_split: up to 8 blocks per CU, no more than the first level has tiles.
static unsigned split_grid(const Scene *s, int64_t count) {
        const int64_t tiles = (count + (int64_t)kMisBlock * kSplitPer - 1) / ((int64_t)kMisBlock * kSplitPer);
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)s->cuCount * 8, tiles));
}

// Light and material tables of one scene handle.
struct MisScene {
        GzMaterial *dMats = nullptr;
        int nMats = 0;
        DevLight *dLights = nullptr;
        int nLights = 0;
        float *dPower = nullptr;
        float *dCumulative = nullptr;
        int nChoices = 0;
        float totalPower = 0.0f;
        bool hasSky = false;
        float sky[3] = {0, 0, 0};
        uint8_t *dPrimClass = nullptr;  // per primitive: prim_class_of (depends on the materials)
        float emitLo[3] = {1.0f, 1.0f, 1.0f}, emitHi[3] = {0.0f, 0.0f, 0.0f};  // MisArgs::emitLo/Hi
        bool allDiffuse = true;  // picks the kernel built with the diffuse BSDF only
        bool anyLayered = false;  // some material is the layered (coated conductor) BSDF
        bool anyOther = false;    // some material is non-diffuse and not layered
};

void release_mis(MisScene *m) {
        if (!m) return;
        void *ptrs[] = {m->dMats, m->dLights, m->dPower, m->dCumulative, m->dPrimClass};
        for (void *p : ptrs)
                if (p) (void)hipFree(p);
        delete m;
}

// Reads one light triangle's vertices from the device-resident mesh arrays (validated).
static int light_vertices(Scene *s, int32_t i, const GzAreaLight &L, float p[3][3]) {
        if (L.mesh < 0 || L.mesh >= s->meshCount) {
                set_error("gz_gpu_set_mis_scene: light %d names mesh %d of %lld", i, L.mesh, (long long)s->meshCount);
                return GZ_E_INVALID;
        }
        DevMesh m;
        if (!hip_ok(hipMemcpy(&m, s->dMeshes + L.mesh, sizeof m, hipMemcpyDeviceToHost), "read mesh")) return GZ_E_HIP;
        const int64_t base = 3 * L.triangle;
        if (L.triangle < 0 || base + 2 >= m.vertexIndexCount) {
                set_error("gz_gpu_set_mis_scene: light %d triangle %lld is outside mesh %d", i, (long long)L.triangle,
                          L.mesh);
                return GZ_E_INVALID;
        }
        int64_t vi[3];
        if (!hip_ok(hipMemcpy(vi, m.vertexIndices + base, sizeof vi, hipMemcpyDeviceToHost), "read indices"))
                return GZ_E_HIP;
        for (int k = 0; k < 3; ++k) {
                if (vi[k] < 0 || vi[k] * 4 + 2 >= m.pointFloats) {
                        set_error("gz_gpu_set_mis_scene: light %d has a vertex index out of range", i);
                        return GZ_E_INVALID;
                }
                if (!hip_ok(hipMemcpy(p[k], m.points + vi[k] * 4, 3 * sizeof(float), hipMemcpyDeviceToHost),
                            "read points"))
                        return GZ_E_HIP;
        }
        return GZ_OK;
}

template <class T>
static bool put(T **dst, const T *src, size_t count, const char *what) {
        return hip_ok(hipMalloc((void **)dst, sizeof(T) * (count > 0 ? count : 1)), what) &&
               (count == 0 || hip_ok(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice), what));
}

__global__ __launch_bounds__(kBlock) void k_prim_tris(const PrimId_C *__restrict__ prims, int64_t n,
                                                      const DevMesh *__restrict__ meshes, int64_t meshCount,
                                                      float4 *__restrict__ out) {
        const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        f3 p0, p1, p2;
        const bool ok = tri_points(meshes, meshCount, prims[i], p0, p1, p2);
        out[3 * i] = ok ? make_float4(p0.x, p0.y, p0.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        out[3 * i + 1] = ok ? make_float4(p1.x, p1.y, p1.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        out[3 * i + 2] = ok ? make_float4(p2.x, p2.y, p2.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_prim_class(const PrimId_C *__restrict__ prims, int64_t n,
                                                       const float4 *__restrict__ primTris,
                                                       const GzMaterial *__restrict__ mats, int nMats,
                                                       uint8_t *__restrict__ out) {
        const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        out[i] = prim_class_of(prims[i], primTris + 3 * i, mats, nMats);
}

// Float keys whose integer order is the float order (for atomicMin / atomicMax).
__device__ __forceinline__ int float_key(float f) {
        const int b = __float_as_int(f);
        return b >= 0 ? b : b ^ 0x7fffffff;
}

// Bounds of the vertices of the primitives of class kClsLight or kClsFull into box[0..5] (keys of
// min x, y, z and max x, y, z); box[6] = 1 when such a primitive has no vertices (indices out of
// range), whose hits then cannot be bounded.
__global__ __launch_bounds__(kBlock) void k_emit_box(const uint8_t *__restrict__ cls, const float4 *__restrict__ primTris,
                                                     int64_t n, int *__restrict__ box) {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
                if (!(cls[i] & (kClsLight | kClsFull))) continue;
                const float4 q[3] = {primTris[3 * i], primTris[3 * i + 1], primTris[3 * i + 2]};
                if (q[0].w != 0.0f) {
                        atomicOr(box + 6, 1);
                        continue;
                }
                for (int v = 0; v < 3; ++v) {
                        const float c[3] = {q[v].x, q[v].y, q[v].z};
                        for (int k = 0; k < 3; ++k) {
                                atomicMin(box + k, float_key(c[k]));
                                atomicMax(box + 3 + k, float_key(c[k]));
                        }
                }
        }
}

static float key_float(int k) {
        const int b = k >= 0 ? k : k ^ 0x7fffffff;
        float f;
        std::memcpy(&f, &b, sizeof f);
        return f;
}

// MisScene::emitLo/Hi from its class array: the box, padded outward by 1e-4 of its largest
// coordinate magnitude (far more than the rounding of a slab or triangle test), the whole space
// when some such primitive cannot be bounded, empty when there is none.
static bool emit_box(Scene *s, MisScene *m) {
        int h[7] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN, 0};
        int *d = nullptr;
        if (alloc_fault("emitter box") || !hip_ok(hipMalloc((void **)&d, sizeof h), "hipMalloc(emitter box)")) return false;
        const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)s->cuCount * 4,
                                                                           (s->primCount + kBlock - 1) / kBlock));
        bool ok = hip_ok(hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice), "emitter box");
        if (ok) {
                hipLaunchKernelGGL(k_emit_box, dim3(g), dim3(kBlock), 0, s->stream, m->dPrimClass, s->dPrimTris,
                                   s->primCount, d);
                ok = hip_ok(hipGetLastError(), "emitter box launch") &&
                     hip_ok(hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s->stream), "emitter box") &&
                     hip_ok(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        }
        (void)hipFree(d);
        if (!ok) return false;
        if (h[6]) {
                for (int k = 0; k < 3; ++k) {
                        m->emitLo[k] = -INFINITY;
                        m->emitHi[k] = INFINITY;
                }
                return true;
        }
        if (h[0] == INT_MAX) return true;  // none: the box stays empty
        float big = 0.0f;
        for (int k = 0; k < 6; ++k) big = std::max(big, std::fabs(key_float(h[k])));
        const float pad = 1e-4f * std::max(big, 1.0f);
        for (int k = 0; k < 3; ++k) {
                m->emitLo[k] = key_float(h[k]) - pad;
                m->emitHi[k] = key_float(h[3 + k]) + pad;
        }
        return true;
}

// Fills Scene::dPrimTris once per handle (synchronously, on the handle's stream).
static int ensure_prim_tris(Scene *s) {
        if (s->dPrimTris || s->primCount <= 0) return GZ_OK;
        const size_t bytes = sizeof(float4) * 3 * (size_t)s->primCount;
        void *p = nullptr;
        if (alloc_fault("primitive vertices") || !hip_ok(hipMalloc(&p, bytes), "hipMalloc(primitive vertices)"))
                return GZ_E_ALLOC;
        float4 *d = static_cast<float4 *>(p);
        hipLaunchKernelGGL(k_prim_tris, dim3((unsigned)((s->primCount + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                           s->dPrims, s->primCount, s->dMeshes, s->meshCount, d);
        if (!hip_ok(hipGetLastError(), "primitive vertices launch") ||
            !hip_ok(hipStreamSynchronize(s->stream), "hipStreamSynchronize")) {
                (void)hipFree(p);
                return GZ_E_HIP;
        }
        s->dPrimTris = d;
        s->deviceBytes += (int64_t)bytes;
        return GZ_OK;
}

static int set_mis_scene(Scene *s, const GzMaterial *materials, int32_t materialCount, const GzAreaLight *lights,
                         int32_t lightCount, const float *skyRadiance) {
        const int trc = ensure_prim_tris(s);
        if (trc != GZ_OK) return trc;
        std::vector<DevLight> dl((size_t)lightCount);
        std::vector<float> power;
        // Light powers and the power CDF in the reference's float operation order
        // (AreaLight.swift:47-49, InfiniteLight.swift:135-138, PowerLightSampler.swift:4-20).
        for (int32_t i = 0; i < lightCount; ++i) {
                float p[3][3];
                const int rc = light_vertices(s, i, lights[i], p);
                if (rc != GZ_OK) return rc;
                const float e1x = p[1][0] - p[0][0], e1y = p[1][1] - p[0][1], e1z = p[1][2] - p[0][2];
                const float e2x = p[2][0] - p[0][0], e2y = p[2][1] - p[0][1], e2z = p[2][2] - p[0][2];
                const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
                const float area = 0.5f * sqrtf(cx * cx + cy * cy + cz * cz);  // Triangle.swift:522-525
                const float *b = lights[i].brightness;
                dl[(size_t)i].p0 = make_float4(p[0][0], p[0][1], p[0][2], area);
                dl[(size_t)i].p1 = make_float4(p[1][0], p[1][1], p[1][2], 0.0f);
                dl[(size_t)i].p2 = make_float4(p[2][0], p[2][1], p[2][2], 0.0f);
                dl[(size_t)i].L = make_float4(b[0], b[1], b[2], 0.0f);
                const float avg = (b[0] + b[1] + b[2]) / 3.0f;
                power.push_back(avg * area * kPi);
        }
        if (skyRadiance) {
                const float avg = (skyRadiance[0] + skyRadiance[1] + skyRadiance[2]) / 3.0f;
                const float r = 100.0f / 2.0f;
                power.push_back(kPi * (r * r) * avg);
        }
        std::vector<float> cum(power.size());
        float total = 0.0f;
        for (size_t k = 0; k < power.size(); ++k) total = total + power[k];
        for (size_t k = 0; k < power.size(); ++k) cum[k] = k == 0 ? power[0] : cum[k - 1] + power[k];
        if (!(total > 0.0f) || !std::isfinite(total)) {
                set_error("gz_gpu_set_mis_scene: total light power %g is not positive and finite", (double)total);
                return GZ_E_INVALID;
        }
        MisScene *m = new MisScene;
        m->nMats = materialCount;
        m->nLights = lightCount;
        m->nChoices = (int)power.size();
        m->totalPower = total;
        m->hasSky = skyRadiance != nullptr;
        for (int32_t i = 0; i < materialCount; ++i) {
                if (materials[i].type != GZ_MAT_DIFFUSE) m->allDiffuse = false;
                if (materials[i].type == GZ_MAT_COATED_CONDUCTOR)
                        m->anyLayered = true;
                else if (materials[i].type != GZ_MAT_DIFFUSE)
                        m->anyOther = true;
        }
        if (skyRadiance)
                for (int k = 0; k < 3; ++k) m->sky[k] = skyRadiance[k];
        bool ok = put(&m->dMats, materials, (size_t)materialCount, "upload materials") &&
                  put(&m->dLights, dl.data(), dl.size(), "upload lights") &&
                  put(&m->dPower, power.data(), power.size(), "upload light powers") &&
                  put(&m->dCumulative, cum.data(), cum.size(), "upload light CDF");
        if (ok && s->primCount > 0) {
                ok = !alloc_fault("primitive classes") &&
                     hip_ok(hipMalloc((void **)&m->dPrimClass, (size_t)s->primCount), "hipMalloc(primitive classes)");
                if (ok) {
                        hipLaunchKernelGGL(k_prim_class, dim3((unsigned)((s->primCount + kBlock - 1) / kBlock)), dim3(kBlock),
                                           0, s->stream, s->dPrims, s->primCount, s->dPrimTris, m->dMats, materialCount,
                                           m->dPrimClass);
                        ok = hip_ok(hipGetLastError(), "primitive classes launch") &&
                             hip_ok(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
                }
                ok = ok && emit_box(s, m);
        }
        if (!ok) {
                release_mis(m);
                return GZ_E_HIP;
        }
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        release_mis(s->mis);
        s->mis = m;
        return GZ_OK;
}

bool mis_all_diffuse(const Scene *s) { return !s->mis || s->mis->allDiffuse; }

// The BSDF set a scene's bounce kernels are compiled with: the diffuse BSDF only, every BSDF, or
// every BSDF but the layered one (a scene without coated conductors keeps its megakernel and drain
// at 143 instead of 256 registers).
int gz::mis_bsdf_set(const Scene *s) {
        if (!s->mis || s->mis->allDiffuse) return kBsdfDiffuse;
        return s->mis->anyLayered ? kBsdfAll : kBsdfNoLayer;
}

}  // namespace gz

using namespace gz;

extern "C" int gz_gpu_set_mis_scene(void *handle, const GzMaterial *materials, int32_t materialCount,
                                    const GzAreaLight *lights, int32_t lightCount, const float *skyRadiance) {
        clear_error();
        if (!handle || materialCount < 1 || !materials || lightCount < 0 || (lightCount > 0 && !lights)) {
                set_error("gz_gpu_set_mis_scene: need a handle, at least one material and, for lightCount > 0, lights");
                return GZ_E_INVALID;
        }
        if (lightCount == 0 && !skyRadiance) {
                set_error("gz_gpu_set_mis_scene: the scene has no light");
                return GZ_E_INVALID;
        }
        for (int32_t i = 0; i < materialCount; ++i) {
                const GzMaterial &m = materials[i];
                if (m.type < GZ_MAT_DIFFUSE || m.type > GZ_MAT_COATED_CONDUCTOR) {
                        set_error("gz_gpu_set_mis_scene: material %d has unknown type %d", i, m.type);
                        return GZ_E_INVALID;
                }
                // The layered BSDF's loops run nSamples x maxDepth steps per call.
                if (m.type == GZ_MAT_COATED_CONDUCTOR &&
                    (m.nSamples < 1 || m.nSamples > 64 || m.maxDepth < 0 || m.maxDepth > 256)) {
                        set_error("gz_gpu_set_mis_scene: coated conductor %d needs 1 <= nSamples <= 64 and "
                                  "0 <= maxDepth <= 256 (got %d, %d)", i, m.nSamples, m.maxDepth);
                        return GZ_E_INVALID;
                }
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        // Interface materials (materialIndex < 0 on a surface): the reference's integrator asserts
        // materialIndex >= 0 at every surface hit and then indexes its materials with it
        // (VolumePathIntegrator.swift:341-342), so a path that meets one cannot be rendered; only
        // its shadow rays continue through them (Visibility.swift:8-20). Such scenes are refused.
        if (s->interfacePrims > 0) {
                set_error("gz_gpu_set_mis_scene: %lld surface primitives carry an interface material "
                          "(materialIndex < 0), which the MIS integrator does not render",
                          (long long)s->interfacePrims);
                return GZ_E_INVALID;
        }
        if (!enter_device(s)) return GZ_E_HIP;
        const int rc = set_mis_scene(s, materials, materialCount, lights, lightCount, skyRadiance);
        if (rc == GZ_OK) ++s->misGeneration;
        return rc;
}

bool gz::mis_args(Scene *s, const GzMisParams *params, int *err, MisArgs &a) {
        const MisScene *ms = s->mis;
        if (!ms) {
                set_error("no MIS scene on this handle: call gz_gpu_set_mis_scene first");
                return false;
        }
        if (!params || params->maxDepth < 0 || params->maxDepth > 1024 || params->zsLog2Spp < 0 ||
            params->zsLog2Spp > 64 || params->zsBase4Digits < 0 || params->zsBase4Digits > 64 || params->zsSeed < 0) {
                set_error("bad MIS parameters");
                return false;
        }
        const uint32_t *table = device_sobol_table();
        if (!table) return false;
        a.nodes = s->dNodes;
        a.tris = s->dTris;
        a.rootDesc = s->rootDesc;
        a.sceneSafe = s->boundsSafe;
        a.empty = s->empty;
        a.prims = s->dPrims;
        a.primCount = s->primCount;
        a.primTris = s->dPrimTris;
        a.primClass = ms->dPrimClass;
        a.meshes = s->dMeshes;
        a.meshCount = s->meshCount;
        a.mats = ms->dMats;
        a.nMats = ms->nMats;
        a.lights = ms->dLights;
        a.nLights = ms->nLights;
        a.power = ms->dPower;
        a.cumulative = ms->dCumulative;
        a.nChoices = ms->nChoices;
        a.totalPower = ms->totalPower;
        a.hasSky = ms->hasSky;
        for (int k = 0; k < 3; ++k) {
                a.sky[k] = ms->sky[k];
                a.emitLo[k] = ms->emitLo[k];
                a.emitHi[k] = ms->emitHi[k];
        }
        a.sobol = table;
        a.zs = ZSobolParams{params->zsSeed, params->zsLog2Spp, params->zsBase4Digits};
        a.maxDepth = params->maxDepth;
        a.flags = params->flags;
        a.rrSeed = params->rrSeed;
        a.err = err;
        return true;
}

// Runs the bounces of a whole trace (levels 0 .. maxDepth) over the device queues: a clear of
// the counters, then one persistent launch per level. `launch(level, grid, qin, qout, ctl)`
// enqueues one level's kernel.
template <typename K, typename Launch>
static int queued_levels(Scene *s, MisQueues *&owner, K kern, int64_t count, int maxDepth, hipStream_t st,
                         const char *who, Launch &&launch,
                         const std::function<void(unsigned, const int *, unsigned *, unsigned *, unsigned)> &drain = {}) {
        const int levels = maxDepth + 2;
        MisQueue *q = mis_queue(owner, st, (size_t)count, levels);
        if (!q) return GZ_E_ALLOC;
        const GzMisLaunch &c = s->misLaunch;
        const unsigned chunk = (unsigned)c.chunk;
        const unsigned grid = queue_grid(s, kern, c.block, count, chunk);
        hipLaunchKernelGGL(k_queue_clear, dim3(1), dim3(256), 0, st, q->ctl, levels);
        if (!hip_ok(hipGetLastError(), who)) return GZ_E_HIP;
        for (int level = 0; level <= maxDepth; ++level) {
                const int *qin = level == 0 ? nullptr : q->q[level & 1];
                int *qout = q->q[(level + 1) & 1];
                if (drain && level >= c.drainLevel) {
                        drain(grid, qin, &q->ctl[level].n, &q->ctl[level].head, chunk);
                        return hip_ok(hipGetLastError(), who) ? GZ_OK : GZ_E_HIP;
                }
                launch(level, grid, qin, qout, q->ctl, chunk);
                if (!hip_ok(hipGetLastError(), who)) return GZ_E_HIP;
        }
        return GZ_OK;
}

// The launch a handle uses until gz_gpu_set_mis_launch sets one. The full wavefront keeps each
// traversal at the traversal kernels' occupancy where the megakernel, with every BSDF compiled
// in, runs at 2 waves per SIMD; the megakernel wins only on small all-diffuse scenes (4 waves per
// SIMD, the BVH in L2), where the wavefront's extra passes cost more than they save. Both drain
// the bounce levels from 3 on.
bool gz::default_wavefront(const Scene *s) {
        const bool allDiffuse = s->mis ? s->mis->allDiffuse : true;
        return !(allDiffuse && s->primCount <= kDefaultMegakernelPrims);
}

static void refresh_default_launch(Scene *s) {
        if (s->misLaunchSet) return;
        s->misLaunch = default_wavefront(s) ? GzMisLaunch{3, 256, 1, 64, 3, {0, 0, 0}}
                                            : GzMisLaunch{1, 256, 1, 256, 3, {0, 0, 0}};
}

constexpr int kMaxRecordLevels = 1025;  // maxDepth + 1 for the largest maxDepth (1024)

// The three arrays of a replay record of `levels` levels of `count` paths at `base`.
struct MisRecord {
        GzHit *hits = nullptr, *mhits = nullptr;
        uint8_t *occ = nullptr;
        int levels = 0;
};
static MisRecord record_at(void *base, int64_t count, int levels) {
        MisRecord r;
        r.hits = static_cast<GzHit *>(base);
        r.mhits = r.hits + (size_t)levels * (size_t)count;
        r.occ = reinterpret_cast<uint8_t *>(r.mhits + (size_t)levels * (size_t)count);
        r.levels = levels;
        return r;
}

// queues: the queue storage to use (null: the handle's own); err: the flag traversal overflows
// raise (null: the handle's, per stream); rec: a replay record to fill (full wavefront levels
// below the drain level; every other entry stays marked as not traced).
static int mis_launch(void *handle, GzPath *d_paths, int64_t count, const GzMisParams *params,
                      GzPathStats *d_stats, void *stream, bool toEnd, const char *who, MisQueues **queues = nullptr,
                      int *err = nullptr, const MisRecord *rec = nullptr) {
        clear_error();
        if (!handle || count < 0 || (count > 0 && !d_paths)) {
                set_error("%s: bad arguments", who);
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        const StreamSlot *slot = stream_slot(s, static_cast<hipStream_t>(stream));
        if (!slot) return GZ_E_ALLOC;
        MisArgs a;
        if (!mis_args(s, params, err ? err : slot->dErr, a)) return GZ_E_INVALID;
        if (count == 0) return GZ_OK;
        MisQueues *&owner = queues ? *queues : s->misQueues;
        const int64_t blocks = (count + kMisBlock - 1) / kMisBlock;
        if (blocks > 0x7fffffffLL) {
                set_error("%s: %lld paths exceed one launch", who, (long long)count);
                return GZ_E_INVALID;
        }
        refresh_default_launch(s);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (rec && rec->levels > 0) {
                const int64_t entries = (int64_t)rec->levels * count;
                const unsigned g = (unsigned)std::max<int64_t>(
                        1, std::min<int64_t>((int64_t)s->cuCount * 8, (entries + kBlock - 1) / kBlock));
                hipLaunchKernelGGL(k_record_clear, dim3(g), dim3(kBlock), 0, st, rec->hits, rec->mhits, rec->occ, entries);
                if (!hip_ok(hipGetLastError(), who)) return GZ_E_HIP;
        }
        uint32_t *stats = reinterpret_cast<uint32_t *>(d_stats);
        const bool all = !s->mis->allDiffuse;
        const int bset = mis_bsdf_set(s);
        if (toEnd && s->misLaunch.queued == 3) {
                if (count > 0x7fffffffLL) {
                        set_error("%s: %lld paths exceed one queue (2^31 - 1)", who, (long long)count);
                        return GZ_E_INVALID;
                }
                const int levels = params->maxDepth + 2;
                MisQueue *q = full_queue(owner, st, (size_t)count, levels);
                if (!q) return GZ_E_ALLOC;
                const GzMisLaunch &c = s->misLaunch;
                const GzTraceConfig &tc = s->traceConfig;
                const int refill = tc.refillLanes > 0 && tc.refillLanes <= 64 ? tc.refillLanes : 56;
                const unsigned tchunk = (unsigned)(tc.chunk >= 1 && tc.chunk <= 4096 ? tc.chunk : 64);
                auto ktrace = d_stats ? k_wf_trace<true> : k_wf_trace<false>;
                auto kany = d_stats ? k_wf_rays<1, true> : k_wf_rays<1, false>;
                auto kclose = d_stats ? k_wf_rays<2, true> : k_wf_rays<2, false>;
                auto kd = d_stats ? k_wf_phase<true, kBsdfDiffuse> : k_wf_phase<false, kBsdfDiffuse>;
                auto ka = d_stats ? k_wf_phase<true, kBsdfAll> : k_wf_phase<false, kBsdfAll>;
                auto kn = d_stats ? k_wf_phase<true, kBsdfNoLayer> : k_wf_phase<false, kBsdfNoLayer>;
                // class queues: the layered BSDF's paths and the other non-diffuse ones apart, so
                // the latter shade at 3 waves per SIMD instead of 2
                const bool layered = s->mis->anyLayered, other = s->mis->anyOther;
                const int classes = !all ? 1 : 3;
                auto tgrid_of = [&](auto kern) {
                        int perCU = 0;
                        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess ||
                            perCU < 1)
                                perCU = 1;
                        return (unsigned)std::max<int64_t>(
                                1, std::min<int64_t>((int64_t)s->cuCount * perCU * c.gridScale, (count + kBlock - 1) / kBlock));
                };
                const unsigned tgrid = tgrid_of(ktrace), agrid = tgrid_of(kany), cgrid = tgrid_of(kclose);
                const unsigned chunk = (unsigned)c.chunk;
                const unsigned gd = queue_grid(s, kd, c.block, count, chunk);
                const unsigned ga = queue_grid(s, ka, c.block, count, chunk);
                const unsigned gn = queue_grid(s, kn, c.block, count, chunk);
                const unsigned sgrid = (unsigned)std::max<int64_t>(
                        1, std::min<int64_t>((int64_t)s->cuCount * 8, (count + kMisBlock - 1) / kMisBlock));
                const unsigned spgrid = split_grid(s, count);
                auto kdrain = d_stats ? pick_set(bset, k_mis_drain<true, 0>, k_mis_drain<true, 1>, k_mis_drain<true, 2>)
                                      : pick_set(bset, k_mis_drain<false, 0>, k_mis_drain<false, 1>, k_mis_drain<false, 2>);
                const unsigned gdrain = queue_grid(s, kdrain, c.block, count, chunk);
                hipLaunchKernelGGL(k_wf_clear, dim3(1), dim3(256), 0, st, q->wctl, levels);
                for (int level = 0; level <= params->maxDepth; ++level) {
                        const int *qin = level == 0 ? nullptr : q->q[level & 1];
                        int *qout = q->q[(level + 1) & 1];
                        WfLevel *lv = q->wctl + level;
                        // this level's closest hits, shadow results and BSDF-ray hits: in the replay
                        // record when it keeps the level, else in the queue storage
                        const bool keep = rec && level < rec->levels;
                        GzHit *hits = keep ? rec->hits + (size_t)level * (size_t)count : q->hits;
                        GzHit *mhits = keep ? rec->mhits + (size_t)level * (size_t)count : q->mhits;
                        uint8_t *occ = keep ? rec->occ + (size_t)level * (size_t)count : q->occ;
                        if (c.drainLevel > 0 && level >= c.drainLevel) {
                                hipLaunchKernelGGL(kdrain, dim3(gdrain), dim3(c.block), 0, st, a, d_paths, qin, &lv->qn,
                                                   &lv->qhead, chunk, stats);
                                return hip_ok(hipGetLastError(), who) ? GZ_OK : GZ_E_HIP;
                        }
                        hipLaunchKernelGGL(ktrace, dim3(tgrid), dim3(kBlock), 0, st, a, d_paths, (unsigned)count, qin,
                                           q->wctl, level, hits, refill, tchunk, stats);
                        hipLaunchKernelGGL(k_wf_split, dim3(spgrid), dim3(kMisBlock), 0, st, a, (unsigned)count, qin,
                                           q->wctl, level, hits, q->qd, q->qa, q->qn, classes);
                        hipLaunchKernelGGL(kd, dim3(gd), dim3(c.block), 0, st, a, d_paths, hits, q->qd, q->wctl, level,
                                           qout, q->qs, q->qm, q->pend, q->srays, q->mrays, chunk, stats);
                        if (all && layered)
                                hipLaunchKernelGGL(ka, dim3(ga), dim3(c.block), 0, st, a, d_paths, hits, q->qa, q->wctl,
                                                   level, qout, q->qs, q->qm, q->pend, q->srays, q->mrays, chunk, stats);
                        if (all && other)
                                hipLaunchKernelGGL(kn, dim3(gn), dim3(c.block), 0, st, a, d_paths, hits, q->qn, q->wctl,
                                                   level, qout, q->qs, q->qm, q->pend, q->srays, q->mrays, chunk, stats);
                        hipLaunchKernelGGL(kany, dim3(agrid), dim3(kBlock), 0, st, a, d_paths, (unsigned)count, q->qs,
                                           &lv->sn, q->srays, occ, nullptr, stats);
                        hipLaunchKernelGGL(kclose, dim3(cgrid), dim3(kBlock), 0, st, a, d_paths, (unsigned)count, q->qm,
                                           &lv->mn, q->mrays, nullptr, mhits, stats);
                        hipLaunchKernelGGL(k_wf_finish, dim3(sgrid), dim3(kMisBlock), 0, st, a, d_paths, q->qd, &lv->dn,
                                           q->pend, occ, q->mrays, mhits);
                        if (all && layered)
                                hipLaunchKernelGGL(k_wf_finish, dim3(sgrid), dim3(kMisBlock), 0, st, a, d_paths, q->qa,
                                                   &lv->an, q->pend, occ, q->mrays, mhits);
                        if (all && other)
                                hipLaunchKernelGGL(k_wf_finish, dim3(sgrid), dim3(kMisBlock), 0, st, a, d_paths, q->qn,
                                                   &lv->nn, q->pend, occ, q->mrays, mhits);
                        if (!hip_ok(hipGetLastError(), who)) return GZ_E_HIP;
                }
                return GZ_OK;
        }
        if (toEnd && s->misLaunch.queued == 2) {
                if (count > 0x7fffffffLL) {
                        set_error("%s: %lld paths exceed one queue (2^31 - 1)", who, (long long)count);
                        return GZ_E_INVALID;
                }
                const int levels = params->maxDepth + 2;
                MisQueue *q = wf_queue(owner, st, (size_t)count, levels);
                if (!q) return GZ_E_ALLOC;
                const GzMisLaunch &c = s->misLaunch;
                const GzTraceConfig &tc = s->traceConfig;
                const int refill = tc.refillLanes > 0 && tc.refillLanes <= 64 ? tc.refillLanes : 56;
                const unsigned tchunk = (unsigned)(tc.chunk >= 1 && tc.chunk <= 4096 ? tc.chunk : 64);
                auto ktrace = d_stats ? k_wf_trace<true> : k_wf_trace<false>;
                auto kd = d_stats ? k_wf_shade<true, 0> : k_wf_shade<false, 0>;
                auto ka = d_stats ? (bset == kBsdfNoLayer ? k_wf_shade<true, 2> : k_wf_shade<true, 1>)
                                  : (bset == kBsdfNoLayer ? k_wf_shade<false, 2> : k_wf_shade<false, 1>);
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, ktrace, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const unsigned tgrid = (unsigned)std::max<int64_t>(
                        1, std::min<int64_t>((int64_t)s->cuCount * perCU * c.gridScale, (count + kBlock - 1) / kBlock));
                const unsigned chunk = (unsigned)c.chunk;
                const unsigned gd = queue_grid(s, kd, c.block, count, chunk);
                const unsigned ga = queue_grid(s, ka, c.block, count, chunk);
                auto kdrain = d_stats ? pick_set(bset, k_mis_drain<true, 0>, k_mis_drain<true, 1>, k_mis_drain<true, 2>)
                                      : pick_set(bset, k_mis_drain<false, 0>, k_mis_drain<false, 1>, k_mis_drain<false, 2>);
                const unsigned gdrain = queue_grid(s, kdrain, c.block, count, chunk);
                hipLaunchKernelGGL(k_wf_clear, dim3(1), dim3(256), 0, st, q->wctl, levels);
                for (int level = 0; level <= params->maxDepth; ++level) {
                        const int *qin = level == 0 ? nullptr : q->q[level & 1];
                        int *qout = q->q[(level + 1) & 1];
                        if (c.drainLevel > 0 && level >= c.drainLevel) {
                                hipLaunchKernelGGL(kdrain, dim3(gdrain), dim3(c.block), 0, st, a, d_paths, qin,
                                                   &q->wctl[level].qn, &q->wctl[level].qhead, chunk, stats);
                                return hip_ok(hipGetLastError(), who) ? GZ_OK : GZ_E_HIP;
                        }
                        hipLaunchKernelGGL(ktrace, dim3(tgrid), dim3(kBlock), 0, st, a, d_paths, (unsigned)count, qin,
                                           q->wctl, level, q->hits, refill, tchunk, stats);
                        hipLaunchKernelGGL(k_wf_split, dim3(split_grid(s, count)), dim3(kMisBlock), 0, st, a,
                                           (unsigned)count, qin, q->wctl, level, q->hits, q->qd, q->qa, q->qa,
                                           all ? 2 : 1);
                        hipLaunchKernelGGL(kd, dim3(gd), dim3(c.block), 0, st, a, d_paths, q->hits, q->qd, q->wctl,
                                           level, qout, chunk, stats);
                        if (all)
                                hipLaunchKernelGGL(ka, dim3(ga), dim3(c.block), 0, st, a, d_paths, q->hits, q->qa, q->wctl,
                                                   level, qout, chunk, stats);
                        if (!hip_ok(hipGetLastError(), who)) return GZ_E_HIP;
                }
                return GZ_OK;
        }
        if (toEnd && s->misLaunch.queued) {
                if (count > 0x7fffffffLL) {
                        set_error("%s: %lld paths exceed one queue (2^31 - 1)", who, (long long)count);
                        return GZ_E_INVALID;
                }
                auto kern = d_stats ? pick_set(bset, k_mis_queue<true, 0>, k_mis_queue<true, 1>, k_mis_queue<true, 2>)
                                    : pick_set(bset, k_mis_queue<false, 0>, k_mis_queue<false, 1>, k_mis_queue<false, 2>);
                auto kdrain = d_stats ? pick_set(bset, k_mis_drain<true, 0>, k_mis_drain<true, 1>, k_mis_drain<true, 2>)
                                      : pick_set(bset, k_mis_drain<false, 0>, k_mis_drain<false, 1>, k_mis_drain<false, 2>);
                const int block = s->misLaunch.block;
                std::function<void(unsigned, const int *, unsigned *, unsigned *, unsigned)> drain;
                if (s->misLaunch.drainLevel > 0)
                        drain = [&](unsigned grid, const int *qin, unsigned *n, unsigned *head, unsigned chunk) {
                                hipLaunchKernelGGL(kdrain, dim3(grid), dim3(block), 0, st, a, d_paths, qin, n, head, chunk,
                                                   stats);
                        };
                return queued_levels(
                        s, owner, kern, count, params->maxDepth, st, who,
                        [&](int level, unsigned grid, const int *qin, int *qout, QueueLevel *ctl, unsigned chunk) {
                                hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, a, d_paths, (unsigned)count, qin,
                                                   qout, ctl, level, chunk, stats);
                        },
                        drain);
        }
        // A path makes at most maxDepth + 1 bounces (bounce 0 .. maxDepth); a launch after the
        // last finds every path inactive.
        const int launches = toEnd ? params->maxDepth + 1 : 1;
        auto kernel = d_stats ? pick_set(bset, k_mis_bounce<true, 0>, k_mis_bounce<true, 1>, k_mis_bounce<true, 2>)
                              : pick_set(bset, k_mis_bounce<false, 0>, k_mis_bounce<false, 1>, k_mis_bounce<false, 2>);
        for (int b = 0; b < launches; ++b) {
                hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(kMisBlock), 0, st, a, d_paths, count, stats);
                if (!hip_ok(hipGetLastError(), "MIS bounce launch")) return GZ_E_HIP;
        }
        return GZ_OK;
}

extern "C" int gz_gpu_set_mis_launch(void *handle, const GzMisLaunch *launch) {
        clear_error();
        if (!handle || !launch) {
                set_error("gz_gpu_set_mis_launch: bad arguments");
                return GZ_E_INVALID;
        }
        const GzMisLaunch &c = *launch;
        if (c.queued < 0 || c.queued > 3 || (c.block != 64 && c.block != 128 && c.block != 256) ||
            c.gridScale < 1 || c.gridScale > 16 || c.chunk < 64 || c.chunk > 4096 || c.chunk % 64 != 0 ||
            c.drainLevel < 0 || c.drainLevel > 64 || (c.queued == 0 && c.drainLevel != 0) || c.reserved[0] != 0 ||
            c.reserved[1] != 0 || c.reserved[2] != 0) {
                set_error("gz_gpu_set_mis_launch: queued %d block %d gridScale %d chunk %d drainLevel %d out of range "
                          "(drainLevel needs a queued mode; reserved must be zero)",
                          c.queued, c.block, c.gridScale, c.chunk, c.drainLevel);
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        s->misLaunch = c;
        s->misLaunchSet = true;
        return GZ_OK;
}

extern "C" int gz_gpu_get_mis_launch(void *handle, GzMisLaunch *launch) {
        clear_error();
        if (!handle || !launch) {
                set_error("gz_gpu_get_mis_launch: bad arguments");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        refresh_default_launch(s);
        *launch = s->misLaunch;
        return GZ_OK;
}

extern "C" int gz_gpu_release_mis_queues(void *handle) {
        clear_error();
        if (!handle) {
                set_error("gz_gpu_release_mis_queues: NULL handle");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!s->misQueues) return GZ_OK;
        if (!enter_device(s)) return GZ_E_HIP;
        // the queues may still be in use by work queued on their streams
        for (auto &kv : s->misQueues->perStream)
                if (!hip_ok(hipStreamSynchronize(kv.first), "gz_gpu_release_mis_queues: stream synchronize"))
                        return GZ_E_HIP;
        release_mis_queues(s->misQueues);
        s->misQueues = nullptr;
        return GZ_OK;
}

extern "C" int gz_gpu_mis_bounce_device(void *handle, GzPath *d_paths, int64_t count, const GzMisParams *params,
                                        void *stream) {
        return mis_launch(handle, d_paths, count, params, nullptr, stream, false, "gz_gpu_mis_bounce_device");
}

extern "C" int gz_gpu_mis_trace_device(void *handle, GzPath *d_paths, int64_t count, const GzMisParams *params,
                                       void *stream) {
        return mis_launch(handle, d_paths, count, params, nullptr, stream, true, "gz_gpu_mis_trace_device");
}

int gz::mis_trace_shared(Scene *s, GzPath *d_paths, int64_t count, const GzMisParams *params, hipStream_t stream,
                         MisQueues **queues, int *err) {
        return mis_launch(s, d_paths, count, params, nullptr, stream, true, "batch scene trace", queues, err);
}

extern "C" int gz_gpu_mis_trace_stats_device(void *handle, GzPath *d_paths, int64_t count,
                                             const GzMisParams *params, GzPathStats *d_stats, void *stream) {
        if (count > 0 && !d_stats) {
                set_error("gz_gpu_mis_trace_stats_device: null stats array");
                return GZ_E_INVALID;
        }
        return mis_launch(handle, d_paths, count, params, d_stats, stream, true, "gz_gpu_mis_trace_stats_device");
}

// The adjoint replay; rec: the forward pass's replay record (null: every ray is traced again).
static int mis_adjoint(void *handle, GzPath *d_paths, const GzPathAdjoint *d_adjoint, int64_t count,
                       const GzMisParams *params, double *d_grad, void *stream, const MisRecord *rec) {
        clear_error();
        if (!handle || count < 0 || (count > 0 && (!d_paths || !d_adjoint || !d_grad))) {
                set_error("gz_gpu_mis_adjoint_device: bad arguments");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        const StreamSlot *slot = stream_slot(s, static_cast<hipStream_t>(stream));
        if (!slot) return GZ_E_ALLOC;
        MisArgs a;
        if (!mis_args(s, params, slot->dErr, a)) return GZ_E_INVALID;
        if (count == 0) return GZ_OK;
        const int64_t blocks = (count + kMisBlock - 1) / kMisBlock;
        if (blocks > 0x7fffffffLL) {
                set_error("gz_gpu_mis_adjoint_device: %lld paths exceed one launch", (long long)count);
                return GZ_E_INVALID;
        }
        const hipStream_t st = static_cast<hipStream_t>(stream);
        AdjArgs adj{d_adjoint, d_grad};
        if (rec && rec->levels > 0) {
                adj.recHits = rec->hits;
                adj.recMhits = rec->mhits;
                adj.recOcc = rec->occ;
                adj.recLevels = rec->levels;
                adj.recCount = count;
        }
        refresh_default_launch(s);
        if (s->misLaunch.queued) {
                if (count > 0x7fffffffLL) {
                        set_error("gz_gpu_mis_adjoint_device: %lld paths exceed one queue (2^31 - 1)", (long long)count);
                        return GZ_E_INVALID;
                }
                const int bset = mis_bsdf_set(s);
                auto kern = pick_set(bset, k_mis_adjoint_queue<0>, k_mis_adjoint_queue<1>, k_mis_adjoint_queue<2>);
                auto kdrain = pick_set(bset, k_mis_adjoint_drain<0>, k_mis_adjoint_drain<1>, k_mis_adjoint_drain<2>);
                const int block = s->misLaunch.block;
                std::function<void(unsigned, const int *, unsigned *, unsigned *, unsigned)> drain;
                if (s->misLaunch.drainLevel > 0)
                        drain = [&](unsigned grid, const int *qin, unsigned *n, unsigned *head, unsigned chunk) {
                                hipLaunchKernelGGL(kdrain, dim3(grid), dim3(block), 0, st, a, d_paths, qin, n, head, chunk,
                                                   adj);
                        };
                return queued_levels(
                        s, s->misQueues, kern, count, params->maxDepth, st, "MIS adjoint launch",
                        [&](int level, unsigned grid, const int *qin, int *qout, QueueLevel *ctl, unsigned chunk) {
                                hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, a, d_paths, (unsigned)count, qin,
                                                   qout, ctl, level, chunk, adj);
                        },
                        drain);
        }
        auto kernel = pick_set(mis_bsdf_set(s), k_mis_adjoint<0>, k_mis_adjoint<1>, k_mis_adjoint<2>);
        for (int b = 0; b <= params->maxDepth; ++b) {
                hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(kMisBlock), 0, st, a, d_paths, count, adj);
                if (!hip_ok(hipGetLastError(), "MIS adjoint launch")) return GZ_E_HIP;
        }
        return GZ_OK;
}

extern "C" int gz_gpu_mis_adjoint_device(void *handle, GzPath *d_paths, const GzPathAdjoint *d_adjoint,
                                         int64_t count, const GzMisParams *params, double *d_grad, void *stream) {
        return mis_adjoint(handle, d_paths, d_adjoint, count, params, d_grad, stream, nullptr);
}

extern "C" int64_t gz_mis_record_bytes(int64_t count, int32_t levels) {
        if (count < 0 || levels < 0 || levels > kMaxRecordLevels) return -1;
        return count * (int64_t)levels * (int64_t)(2 * sizeof(GzHit) + 1);
}

extern "C" int gz_gpu_mis_trace_record_device(void *handle, GzPath *d_paths, int64_t count, const GzMisParams *params,
                                              void *d_record, int32_t levels, void *stream) {
        if (levels < 0 || levels > kMaxRecordLevels || (levels > 0 && count > 0 && !d_record)) {
                clear_error();
                set_error("gz_gpu_mis_trace_record_device: need 0 <= levels <= %d and a record", kMaxRecordLevels);
                return GZ_E_INVALID;
        }
        const MisRecord rec = record_at(d_record, count, levels);
        return mis_launch(handle, d_paths, count, params, nullptr, stream, true, "gz_gpu_mis_trace_record_device",
                          nullptr, nullptr, &rec);
}

extern "C" int gz_gpu_mis_adjoint_record_device(void *handle, GzPath *d_paths, const GzPathAdjoint *d_adjoint,
                                                int64_t count, const GzMisParams *params, double *d_grad,
                                                const void *d_record, int32_t levels, void *stream) {
        if (levels < 0 || levels > kMaxRecordLevels || (levels > 0 && count > 0 && !d_record)) {
                clear_error();
                set_error("gz_gpu_mis_adjoint_record_device: need 0 <= levels <= %d and a record", kMaxRecordLevels);
                return GZ_E_INVALID;
        }
        const MisRecord rec = record_at(const_cast<void *>(d_record), count, levels);
        return mis_adjoint(handle, d_paths, d_adjoint, count, params, d_grad, stream, &rec);
}
