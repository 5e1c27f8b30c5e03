// BVH2 closest-hit and any-hit traversal for gfx950, and the entry points that drive it.
//
// Semantics follow the reference's traverse_bvh2_core (Sources/mojoKernel/kernel.mojo:223-364)
// step for step, so hit indices, t, u and v are bit-identical to the reference's CPU path:
//   - rdir = 1/dir, orgRdir = org*rdir, nearIsMin = rdir >= 0 (kernel.mojo:233-243);
//   - the root box is never tested; traversal starts at node 0;
//   - at an interior node both children's boxes are tested against the current closest t
//     (kernel.mojo:314-327, slab test kernel.mojo:171-206 with its select-form min/max and the
//     x1.0000003 slack); both hit -> go to the nearer (left wins ties) and push the other;
//   - leaves test their primitives in order with Moller-Trumbore (kernel.mojo:122-156); a hit
//     with t == current t is accepted, so the later primitive wins ties.
// Products, sums and differences are rounded one by one (`fp contract(off)`), divisions are
// IEEE-correct (hipcc's default), and min/max are written as the reference's selects so NaN
// propagates the same way.
//
// MI355X layout: one lane per ray; the interior visit reads one 64-byte child-pair record
// instead of two 32-byte nodes, a leaf reads 48-byte pre-gathered triangle records instead of
// PrimId -> mesh -> index -> point chains (five dependent loads per vertex in the reference).
#include <algorithm>

#include "traverse_device.h"

#pragma clang fp contract(off)

namespace gz {

// One lane per ray, one ray per lane.
template <bool kAny, bool kStats, int kLds>
__global__ __launch_bounds__(kBlock) void k_trace(const NodePair *__restrict__ nodes, const TriRec *__restrict__ tris,
                                                  int rootDesc, bool sceneSafe, const Ray_C *__restrict__ rays,
                                                  const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
                                                  uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats,
                                                  int *__restrict__ err) {
        const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        Lane L;
        lane_start(L, rays, tMaxIn, i, rootDesc, sceneSafe);
        Stack<kLds> S;
        int st;
        do {
                st = step<kAny, kStats, kLds>(L, S, nodes, tris);
        } while (st == kStepGo);
        lane_finish<kAny, kStats>(L, i, st == kStepOverflow, hits, occ, stats, err);
}

// Persistent lanes with ray refill. A wave keeps its lanes busy: once `refill` lanes are idle,
// each idle lane takes the next ray index from the wave's reserve, a range of `chunk`
// consecutive rays grabbed from the global counter with one atomic (so the counter sees one
// atomic per `chunk` rays, not one per refill). The long tail of deep rays no longer idles the
// other lanes of a wave. A wave leaves once the counter is exhausted and its lanes are idle.
template <bool kAny, bool kStats, int kLds>
__global__ __launch_bounds__(kBlock) void k_trace_persistent(
        const NodePair *__restrict__ nodes, const TriRec *__restrict__ tris, int rootDesc, bool sceneSafe,
        const Ray_C *__restrict__ rays, const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
        uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats, int *__restrict__ err,
        unsigned long long *__restrict__ counter, int refill, int chunk) {
        Stack<kLds> S;
        Lane L;
        int64_t ray = -1;
        // Wave-uniform state: the reserve [resBase, resBase + resCount), whether the global counter
        // has run past n (no further atomics), and whether no ray is left to hand out at all.
        unsigned long long resBase = 0, resCount = 0;
        bool counterDone = false, exhausted = false;
        const int laneId = (int)(threadIdx.x & 63);
        const uint64_t below = laneId == 0 ? 0ull : (~0ull >> (64 - laneId));
        const unsigned long long un = (unsigned long long)n;
        for (;;) {
                if (!exhausted) {
                        const uint64_t idle = __ballot(ray < 0);
                        const int nIdle = __popcll(idle);
                        if (nIdle >= refill || nIdle == 64) {
                                const unsigned long long rank = (unsigned long long)__popcll(idle & below);
                                const unsigned long long need = (unsigned long long)nIdle;
                                unsigned long long fresh = 0, take = 0;
                                if (resCount < need && !counterDone) {
                                        // Top up with a new chunk, at least enough for every idle lane.
                                        take = need - resCount > (unsigned long long)chunk ? need - resCount
                                                                                        : (unsigned long long)chunk;
                                        const int leader = __ffsll((unsigned long long)idle) - 1;
                                        if (laneId == leader) fresh = atomicAdd(counter, take);
                                        fresh = __shfl(fresh, leader);
                                        if (fresh + take >= un) counterDone = true;
                                }
                                if (ray < 0) {
                                        // The first resCount idle lanes drain the reserve, the others
                                        // start on the fresh chunk.
                                        unsigned long long mine = ~0ull;
                                        if (rank < resCount)
                                                mine = resBase + rank;
                                        else if (rank - resCount < take)
                                                mine = fresh + (rank - resCount);
                                        if (mine < un) {
                                                ray = (int64_t)mine;
                                                lane_start(L, rays, tMaxIn, ray, rootDesc, sceneSafe);
                                        }
                                }
                                if (take > 0) {
                                        const unsigned long long used = need - resCount;  // <= take
                                        resBase = fresh + used;
                                        resCount = take - used;
                                } else {
                                        const unsigned long long used = need < resCount ? need : resCount;
                                        resBase += used;
                                        resCount -= used;
                                }
                                if (counterDone && (resCount == 0 || resBase >= un)) exhausted = true;
                        }
                }
                if (ray >= 0) {
                        const int st = step<kAny, kStats, kLds>(L, S, nodes, tris);
                        if (st != kStepGo) {
                                lane_finish<kAny, kStats>(L, ray, st == kStepOverflow, hits, occ, stats, err);
                                ray = -1;
                        }
                }
                if (exhausted && __ballot(ray >= 0) == 0ull) break;
        }
}

// Strided persistent lanes: global lane g traces rays g, g + G, g + 2G, ... (G = lanes in the
// grid) back to back, starting the next ray the step after the previous one ends. Summing many
// rays per lane evens out the per-ray work, with no atomics; consecutive lanes still read
// consecutive rays. Each lane leaves after its last ray, so every wave exits.
template <bool kAny, bool kStats, int kLds>
__global__ __launch_bounds__(kBlock) void k_trace_strided(
        const NodePair *__restrict__ nodes, const TriRec *__restrict__ tris, int rootDesc, bool sceneSafe,
        const Ray_C *__restrict__ rays, const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
        uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats, int *__restrict__ err) {
        Stack<kLds> S;
        const int64_t G = (int64_t)gridDim.x * kBlock;
        int64_t ray = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (ray >= n) return;
        Lane L;
        lane_start(L, rays, tMaxIn, ray, rootDesc, sceneSafe);
        for (;;) {
                const int st = step<kAny, kStats, kLds>(L, S, nodes, tris);
                if (st != kStepGo) {
                        lane_finish<kAny, kStats>(L, ray, st == kStepOverflow, hits, occ, stats, err);
                        ray += G;
                        if (ray >= n) break;
                        lane_start(L, rays, tMaxIn, ray, rootDesc, sceneSafe);
                }
        }
}

__global__ void k_reset_counter(unsigned long long *counter) { *counter = 0ull; }

// Empty scene (no BVH2 nodes): every ray misses (the guard of BoundingHierarchy.swift:353,392).
__global__ __launch_bounds__(kBlock) void k_miss(const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
                                                 uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats) {
        const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        if (hits) hits[i] = GzHit{-1, tMaxIn[i], 0.0f, 0.0f};
        if (occ) occ[i] = 0;
        if (stats) stats[i] = GzTraceStats{0u, 0u};
}

// kernel.mojo:360-364: Intersection_C from a compact hit.
__global__ __launch_bounds__(kBlock) void k_expand(const GzHit *__restrict__ hits, const float *__restrict__ tMaxIn,
                                                   const PrimId_C *__restrict__ prims, int64_t primCount,
                                                   int64_t n, Intersection_C *__restrict__ out) {
        const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        const GzHit h = hits[i];
        Intersection_C r;
        memset(&r, 0, sizeof r);
        if (h.prim >= 0 && h.prim < primCount) {
                r.primId = prims[h.prim];
                r.tHit = h.t;
                r.u = h.u;
                r.v = h.v;
                r.hit = 1;
        } else {
                r.primId.id1 = -1;
                r.primId.id2 = -1;
                r.tHit = tMaxIn[i];
        }
        out[i] = r;
}

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <bool kAny, bool kStats, int kLds>
static void launch_variant(Scene *s, const StreamSlot &slot, const Ray_C *d_rays, const float *d_tMax, int64_t n,
                           GzHit *d_hits, uint8_t *d_occ, GzTraceStats *d_stats, hipStream_t stream) {
        const GzTraceConfig &c = s->traceConfig;
        if (c.persistent == 2) {
                auto kern = k_trace_strided<kAny, kStats, kLds>;
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
                const unsigned g = (unsigned)std::min<int64_t>(want, grid_for(n));
                hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, stream, s->dNodes, s->dTris, s->rootDesc,
                                   s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats, slot.dErr);
        } else if (c.persistent == 1) {
                auto kern = k_trace_persistent<kAny, kStats, kLds>;
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
                const unsigned g = (unsigned)std::min<int64_t>(want, grid_for(n));
                unsigned long long *ctr = slot.dCounter;
                hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, ctr);
                hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, stream, s->dNodes, s->dTris, s->rootDesc,
                                   s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats, slot.dErr, ctr,
                                   c.refillLanes > 0 && c.refillLanes <= 64 ? c.refillLanes : 16,
                                   c.chunk >= 1 && c.chunk <= 4096 ? c.chunk : 64);
        } else {
                hipLaunchKernelGGL((k_trace<kAny, kStats, kLds>), dim3(grid_for(n)), dim3(kBlock), 0, stream, s->dNodes,
                                   s->dTris, s->rootDesc, s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats,
                                   slot.dErr);
        }
}

template <bool kAny, bool kStats>
static void launch_lds(Scene *s, const StreamSlot &slot, const Ray_C *d_rays, const float *d_tMax, int64_t n,
                       GzHit *d_hits, uint8_t *d_occ, GzTraceStats *d_stats, hipStream_t stream) {
        if (s->traceConfig.ldsStack >= 16)
                launch_variant<kAny, kStats, 16>(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream);
        else if (s->traceConfig.ldsStack >= 8)
                launch_variant<kAny, kStats, 8>(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream);
        else
                launch_variant<kAny, kStats, 0>(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream);
}

// The launches of one traversal with the error flag and ray counter of `slot`.
static int launch_trace_slot(Scene *s, const StreamSlot *slot, const Ray_C *d_rays, const float *d_tMax, int64_t n,
                             GzHit *d_hits, uint8_t *d_occ, GzTraceStats *d_stats, bool anyHit, hipStream_t stream) {
        if (s->empty) {
                hipLaunchKernelGGL(k_miss, dim3(grid_for(n)), dim3(kBlock), 0, stream, d_tMax, n,
                                   anyHit ? nullptr : d_hits, anyHit ? d_occ : nullptr, d_stats);
        } else if (anyHit) {
                if (d_stats)
                        launch_lds<true, true>(s, *slot, d_rays, d_tMax, n, nullptr, d_occ, d_stats, stream);
                else
                        launch_lds<true, false>(s, *slot, d_rays, d_tMax, n, nullptr, d_occ, nullptr, stream);
        } else {
                if (d_stats)
                        launch_lds<false, true>(s, *slot, d_rays, d_tMax, n, d_hits, nullptr, d_stats, stream);
                else
                        launch_lds<false, false>(s, *slot, d_rays, d_tMax, n, d_hits, nullptr, nullptr, stream);
        }
        return hip_ok(hipGetLastError(), "traversal launch") ? GZ_OK : GZ_E_HIP;
}

static bool trace_size_ok(int64_t n) {
        if (n <= (int64_t)kBlock * 0x7fffffffLL) return true;
        set_error("traversal batch of %lld rays is too large", (long long)n);
        return false;
}

int launch_trace(Scene *s, const Ray_C *d_rays, const float *d_tMax, int64_t n, GzHit *d_hits, uint8_t *d_occ,
                 GzTraceStats *d_stats, bool anyHit, hipStream_t stream) {
        if (n <= 0) return GZ_OK;
        if (!trace_size_ok(n)) return GZ_E_INVALID;
        const StreamSlot *slot = stream_slot(s, stream);
        if (!slot) return GZ_E_ALLOC;
        return launch_trace_slot(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, anyHit, stream);
}

int launch_expand(Scene *s, const GzHit *d_hits, const float *d_tMax, int64_t n, Intersection_C *d_out,
                  hipStream_t stream) {
        if (n <= 0) return GZ_OK;
        hipLaunchKernelGGL(k_expand, dim3(grid_for(n)), dim3(kBlock), 0, stream, d_hits, d_tMax, s->dPrims, s->primCount, n, d_out);
        return hip_ok(hipGetLastError(), "expand launch") ? GZ_OK : GZ_E_HIP;
}

}  // namespace gz

using namespace gz;

static Scene *scene_of(void *h, const char *fn) {
        if (!h) {
                set_error("%s: null scene handle", fn);
                return nullptr;
        }
        return static_cast<Scene *>(h);
}

static int check_overflow_sync(Scene *s) { return check_stream(s, s->stream); }

extern "C" int gz_gpu_traverse_closest_device(void *handle, const Ray_C *d_rays, const float *d_tMax, int64_t count,
                                              GzHit *d_hits, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_closest_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_hits))) {
                set_error("gz_gpu_traverse_closest_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, d_hits, nullptr, nullptr, false, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_traverse_closest_stats_device(void *handle, const Ray_C *d_rays, const float *d_tMax,
                                                    int64_t count, GzHit *d_hits, GzTraceStats *d_stats,
                                                    void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_closest_stats_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_hits))) {
                set_error("gz_gpu_traverse_closest_stats_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, d_hits, nullptr, d_stats, false, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_traverse_any_device(void *handle, const Ray_C *d_rays, const float *d_tMax, int64_t count,
                                          uint8_t *d_occluded, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_any_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_occluded))) {
                set_error("gz_gpu_traverse_any_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, nullptr, d_occluded, nullptr, true, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_traverse_any_stats_device(void *handle, const Ray_C *d_rays, const float *d_tMax, int64_t count,
                                                uint8_t *d_occluded, GzTraceStats *d_stats, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_any_stats_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_occluded))) {
                set_error("gz_gpu_traverse_any_stats_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, nullptr, d_occluded, d_stats, true, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_expand_hits_device(void *handle, const GzHit *d_hits, const float *d_tMax, int64_t count,
                                         Intersection_C *d_results, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_expand_hits_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_hits || !d_tMax || !d_results))) {
                set_error("gz_gpu_expand_hits_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_expand(s, d_hits, d_tMax, count, d_results, static_cast<hipStream_t>(stream));
}

// Host-buffer batch: stage rays + tMax, trace, expand to Intersection_C, copy back, wait.
extern "C" void gz_gpu_traverse_batch(void *handle, const Ray_C *rays, const float *tMaxValues, int64_t count,
                                      Intersection_C *results) {
        clear_error();
        if (!handle || count <= 0) return;  // kernel.mojo:616-622
        if (!rays || !tMaxValues || !results) {
                set_error("gz_gpu_traverse_batch: null buffer");
                return;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return;
        const size_t n = (size_t)count;
        const size_t offT = (n * sizeof(Ray_C) + 255) & ~(size_t)255;
        const size_t offH = (offT + n * sizeof(float) + 255) & ~(size_t)255;
        const size_t offH2 = (offH + n * sizeof(GzHit) + 255) & ~(size_t)255;
        const size_t total = offH2 + n * sizeof(Intersection_C);
        if (!ensure_stage(s, total)) return;
        char *base = static_cast<char *>(s->dStage);
        Ray_C *dR = reinterpret_cast<Ray_C *>(base);
        float *dT = reinterpret_cast<float *>(base + offT);
        GzHit *dH = reinterpret_cast<GzHit *>(base + offH);
        Intersection_C *dI = reinterpret_cast<Intersection_C *>(base + offH2);
        hipStream_t st = s->stream;
        if (!hip_ok(hipMemcpyAsync(dR, rays, n * sizeof(Ray_C), hipMemcpyHostToDevice, st), "copy rays") ||
            !hip_ok(hipMemcpyAsync(dT, tMaxValues, n * sizeof(float), hipMemcpyHostToDevice, st), "copy tMax"))
                return;
        if (launch_trace(s, dR, dT, count, dH, nullptr, nullptr, false, st) != GZ_OK) return;
        if (launch_expand(s, dH, dT, count, dI, st) != GZ_OK) return;
        if (check_overflow_sync(s) != GZ_OK) return;
        if (!hip_ok(hipMemcpyAsync(results, dI, n * sizeof(Intersection_C), hipMemcpyDeviceToHost, st), "copy results"))
                return;
        (void)hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
}

extern "C" int gz_gpu_occluded_batch(void *handle, const Ray_C *rays, const float *tMaxValues, int64_t count,
                                     uint8_t *occluded) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_occluded_batch");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!rays || !tMaxValues || !occluded))) {
                set_error("gz_gpu_occluded_batch: bad arguments");
                return GZ_E_INVALID;
        }
        if (count == 0) return GZ_OK;
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        const size_t n = (size_t)count;
        const size_t offT = (n * sizeof(Ray_C) + 255) & ~(size_t)255;
        const size_t offO = (offT + n * sizeof(float) + 255) & ~(size_t)255;
        if (!ensure_stage(s, offO + n)) return GZ_E_ALLOC;
        char *base = static_cast<char *>(s->dStage);
        Ray_C *dR = reinterpret_cast<Ray_C *>(base);
        float *dT = reinterpret_cast<float *>(base + offT);
        uint8_t *dO = reinterpret_cast<uint8_t *>(base + offO);
        hipStream_t st = s->stream;
        if (!hip_ok(hipMemcpyAsync(dR, rays, n * sizeof(Ray_C), hipMemcpyHostToDevice, st), "copy rays") ||
            !hip_ok(hipMemcpyAsync(dT, tMaxValues, n * sizeof(float), hipMemcpyHostToDevice, st), "copy tMax"))
                return GZ_E_HIP;
        int rc = launch_trace(s, dR, dT, count, nullptr, dO, nullptr, true, st);
        if (rc != GZ_OK) return rc;
        rc = check_overflow_sync(s);
        if (rc != GZ_OK) return rc;
        if (!hip_ok(hipMemcpyAsync(occluded, dO, n, hipMemcpyDeviceToHost, st), "copy occluded") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        return GZ_OK;
}

// ---- Allocation-free host-buffer batch ----------------------------------------------------------
//
// gz_gpu_traverse_batch_noalloc stages its rays, tMax, hits and records, its error flag and the
// persistent kernel's ray counter in `scratch`, a file-scope device array of fixed size that every
// GPU the library runs on holds (its own copy, at its own address) until the process ends. In host
// code `scratch` is a stand-in for the symbol, not a device address: it is never passed to a
// kernel or a copy. hipGetSymbolAddress gives each device's address once, with that device
// current, and only that address is used. Layout (each part 256-byte aligned):
//   [0, 256) : the error flag (int at 0) and the ray counter (u64 at 8)
//   rays, tMax, hits, Intersection_C records for up to GZ_NOALLOC_MAX_RAYS rays
namespace gz {

static constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct NoAllocLayout {
        size_t rays, tMax, hits, records, end;
};

static constexpr NoAllocLayout noalloc_layout(size_t n) {
        const size_t r = 256;
        const size_t t = align256(r + n * sizeof(Ray_C));
        const size_t h = align256(t + n * sizeof(float));
        const size_t i = align256(h + n * sizeof(GzHit));
        return NoAllocLayout{r, t, h, i, i + n * sizeof(Intersection_C)};
}

constexpr size_t kNoAllocScratchBytes = align256(noalloc_layout((size_t)GZ_NOALLOC_MAX_RAYS).end);
static_assert(kNoAllocScratchBytes <= (size_t)48 << 20, "scratch stays within 48 MiB");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex gScratchMu[kMaxDevices];   // one call at a time per device owns its scratch
static void *gScratchPtr[kMaxDevices];       // device address of `scratch` per device, or null

__global__ void k_clear_flag(int *flag) { *flag = 0; }

// The device address of `scratch` on the current device `dev` (the caller holds gScratchMu[dev]).
static void *scratch_address(int dev) {
        if (!gScratchPtr[dev]) {
                void *p = nullptr;
                if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || !p)
                        return nullptr;
                gScratchPtr[dev] = p;
        }
        return gScratchPtr[dev];
}

}  // namespace gz

extern "C" int gz_gpu_traverse_batch_noalloc(void *handle, const Ray_C *rays, const float *tMaxValues, int64_t count,
                                             Intersection_C *results) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_batch_noalloc");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!rays || !tMaxValues || !results))) {
                set_error("gz_gpu_traverse_batch_noalloc: bad arguments");
                return GZ_E_INVALID;
        }
        if (count > GZ_NOALLOC_MAX_RAYS) {
                set_error("gz_gpu_traverse_batch_noalloc: %lld rays exceed the fixed staging of %d rays",
                          (long long)count, GZ_NOALLOC_MAX_RAYS);
                return GZ_E_INVALID;
        }
        if (count == 0) return GZ_OK;
        if (s->device < 0 || s->device >= kMaxDevices) {
                set_error("gz_gpu_traverse_batch_noalloc: device %d is beyond the %d this entry serves", s->device,
                          kMaxDevices);
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> dlock(gScratchMu[s->device]);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        char *base = static_cast<char *>(scratch_address(s->device));
        if (!base) return GZ_E_HIP;
        const size_t n = (size_t)count;
        const NoAllocLayout L = noalloc_layout(n);
        StreamSlot slot{reinterpret_cast<int *>(base), reinterpret_cast<unsigned long long *>(base + 8)};
        Ray_C *dR = reinterpret_cast<Ray_C *>(base + L.rays);
        float *dT = reinterpret_cast<float *>(base + L.tMax);
        GzHit *dH = reinterpret_cast<GzHit *>(base + L.hits);
        Intersection_C *dI = reinterpret_cast<Intersection_C *>(base + L.records);
        hipStream_t st = s->stream;
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, st, slot.dErr);
        if (!hip_ok(hipGetLastError(), "clear error flag") ||
            !hip_ok(hipMemcpyAsync(dR, rays, n * sizeof(Ray_C), hipMemcpyHostToDevice, st), "copy rays") ||
            !hip_ok(hipMemcpyAsync(dT, tMaxValues, n * sizeof(float), hipMemcpyHostToDevice, st), "copy tMax"))
                return GZ_E_HIP;
        int rc = launch_trace_slot(s, &slot, dR, dT, count, dH, nullptr, nullptr, false, st);
        if (rc != GZ_OK) return rc;
        rc = launch_expand(s, dH, dT, count, dI, st);
        if (rc != GZ_OK) return rc;
        int flag = 0;
        if (!hip_ok(hipMemcpyAsync(&flag, slot.dErr, sizeof(int), hipMemcpyDeviceToHost, st), "read error flag") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        if (flag & 1) {
                set_error("traversal stack overflow: a ray needed more than %d entries", GZ_STACK_CAPACITY);
                return GZ_E_OVERFLOW;
        }
        if (!hip_ok(hipMemcpyAsync(results, dI, n * sizeof(Intersection_C), hipMemcpyDeviceToHost, st), "copy results") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        return GZ_OK;
}
