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

#include "gz_internal.h"

#pragma clang fp contract(off)

namespace gz {

constexpr int kBlock = 256;

struct RayState {
        float ox, oy, oz, dx, dy, dz;
        float rdx, rdy, rdz, ordx, ordy, ordz;
        bool nx, ny, nz;
        bool safe;  // no slab value can be NaN: the min/max form equals the select form
};

// kernel.mojo:171-206, written as the reference's select chains: `a > b ? a : b` keeps the
// second operand when either is NaN, which decides the result for rays with a zero direction
// component (0 * inf) and must be reproduced exactly.
__device__ __forceinline__ bool slab_exact(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                           const RayState &r, float tMax, float &tNearOut) {
        const float nearX = r.nx ? mnx : mxx, farX = r.nx ? mxx : mnx;
        const float nearY = r.ny ? mny : mxy, farY = r.ny ? mxy : mny;
        const float nearZ = r.nz ? mnz : mxz, farZ = r.nz ? mxz : mnz;
        const float tNearX = nearX * r.rdx - r.ordx;
        const float tNearY = nearY * r.rdy - r.ordy;
        const float tNearZ = nearZ * r.rdz - r.ordz;
        const float tFarX = farX * r.rdx - r.ordx;
        const float tFarY = farY * r.rdy - r.ordy;
        const float tFarZ = farZ * r.rdz - r.ordz;
        float tNear = tNearX > tNearY ? tNearX : tNearY;
        tNear = tNearZ > tNear ? tNearZ : tNear;
        tNear = 0.0f > tNear ? 0.0f : tNear;
        float tFar = tFarX < tFarY ? tFarX : tFarY;
        tFar = tFarZ < tFar ? tFarZ : tFar;
        tFar = tMax < tFar ? tMax : tFar;
        tFar = tFar * 1.0000003f;
        tNearOut = tNear;
        return tNear <= tFar;
}

// Same test for rays whose slab values are all finite (RayState::safe): with no NaN operand the
// select chains are max/min (up to the sign of a zero, which no later comparison can see), so
// they map to the three-operand v_max3/v_min3 instructions. A NaN tMax is ignored by both forms.
__device__ __forceinline__ bool slab_fast(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                          const RayState &r, float tMax, float &tNearOut) {
        const float nearX = r.nx ? mnx : mxx, farX = r.nx ? mxx : mnx;
        const float nearY = r.ny ? mny : mxy, farY = r.ny ? mxy : mny;
        const float nearZ = r.nz ? mnz : mxz, farZ = r.nz ? mxz : mnz;
        const float tNearX = nearX * r.rdx - r.ordx;
        const float tNearY = nearY * r.rdy - r.ordy;
        const float tNearZ = nearZ * r.rdz - r.ordz;
        const float tFarX = farX * r.rdx - r.ordx;
        const float tFarY = farY * r.rdy - r.ordy;
        const float tFarZ = farZ * r.rdz - r.ordz;
        const float tNear = fmaxf(fmaxf(fmaxf(tNearX, tNearY), tNearZ), 0.0f);
        const float tFar = fminf(fminf(fminf(tFarX, tFarY), tFarZ), tMax) * 1.0000003f;
        tNearOut = tNear;
        return tNear <= tFar;
}

__device__ __forceinline__ bool slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, const RayState &r,
                                     float tMax, float &tNearOut) {
        if (r.safe) return slab_fast(mnx, mny, mnz, mxx, mxy, mxz, r, tMax, tNearOut);
        return slab_exact(mnx, mny, mnz, mxx, mxy, mxz, r, tMax, tNearOut);
}

// kernel.mojo:122-156 (cross :84-90 as a.yzx*b.zxy - a.zxy*b.yzx, dot :92-95 left to right).
__device__ __forceinline__ bool tri_test(const RayState &r, float4 a, float4 b, float4 c, float tMax, float &tOut,
                                         float &uOut, float &vOut) {
        const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
        const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
        const float px = r.dy * e2z - r.dz * e2y;
        const float py = r.dz * e2x - r.dx * e2z;
        const float pz = r.dx * e2y - r.dy * e2x;
        const float det = (e1x * px + e1y * py) + e1z * pz;
        if (det > -0.0000001f && det < 0.0000001f) return false;
        const float invDet = 1.0f / det;
        const float tx = r.ox - a.x, ty = r.oy - a.y, tz = r.oz - a.z;
        const float u = ((tx * px + ty * py) + tz * pz) * invDet;
        if (u < 0.0f || u > 1.0f) return false;
        const float qx = ty * e1z - tz * e1y;
        const float qy = tz * e1x - tx * e1z;
        const float qz = tx * e1y - ty * e1x;
        const float v = ((r.dx * qx + r.dy * qy) + r.dz * qz) * invDet;
        if (v < 0.0f || u + v > 1.0f) return false;
        const float t = ((e2x * qx + e2y * qy) + e2z * qz) * invDet;
        if (t <= 0.0f || t > tMax) return false;
        tOut = t;
        uOut = u;
        vOut = v;
        return true;
}

// Per-lane state of one ray in flight.
struct Lane {
        RayState r;
        float tMax, tBest, bu, bv;
        int hitPrim, cur, sp;
        uint32_t nInterior, nTris;
};

__device__ __forceinline__ void lane_start(Lane &L, const Ray_C *__restrict__ rays, const float *__restrict__ tMaxIn,
                                           int64_t i, int rootDesc, bool sceneSafe) {
        const Ray_C ray = rays[i];
        L.tMax = tMaxIn[i];
        RayState &r = L.r;
        r.ox = ray.orgX;
        r.oy = ray.orgY;
        r.oz = ray.orgZ;
        r.dx = ray.dirX;
        r.dy = ray.dirY;
        r.dz = ray.dirZ;
        r.rdx = 1.0f / r.dx;  // kernel.mojo:233-243
        r.rdy = 1.0f / r.dy;
        r.rdz = 1.0f / r.dz;
        r.ordx = r.ox * r.rdx;
        r.ordy = r.oy * r.rdy;
        r.ordz = r.oz * r.rdz;
        r.nx = r.rdx >= 0.0f;
        r.ny = r.rdy >= 0.0f;
        r.nz = r.rdz >= 0.0f;
        // 2^-40 <= |d| <= 2^40 and |o| <= 2^40 on every axis, with box coordinates <= 2^40
        // (checked at upload), bound every slab product and difference by 2^81: no inf, no NaN.
        const float lo = 0x1p-40f, hi = 0x1p40f;
        const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
        r.safe = sceneSafe && ax >= lo && ay >= lo && az >= lo && ax <= hi && ay <= hi && az <= hi &&
                 fabsf(r.ox) <= hi && fabsf(r.oy) <= hi && fabsf(r.oz) <= hi;
        L.hitPrim = -1;
        L.tBest = L.tMax;
        L.bu = 0.0f;
        L.bv = 0.0f;
        L.cur = rootDesc;
        L.sp = 0;
        L.nInterior = 0;
        L.nTris = 0;
}

// The first entries of every lane's stack, [entry][lane of the block]: a lane's pushes and pops
// hit its own column, so a wave's 64 lanes touch 64 consecutive words (no bank conflicts).
constexpr int kLdsMax = 16;
__shared__ int g_stackLds[kLdsMax * kBlock];

// Traversal stack of GZ_STACK_CAPACITY entries: the first kLds live in LDS, the rest in
// per-lane private memory.
template <int kLds>
struct Stack {
        int priv[GZ_STACK_CAPACITY - kLds > 0 ? GZ_STACK_CAPACITY - kLds : 1];
        __device__ __forceinline__ void put(int sp, int v) {
                if (kLds > 0 && sp < kLds)
                        g_stackLds[sp * kBlock + threadIdx.x] = v;
                else
                        priv[sp - kLds] = v;
        }
        __device__ __forceinline__ int get(int sp) const {
                // The volatile LDS read keeps the two loads apart: merged, they become one generic
                // (flat) load that pays the latency of both memories on every pop.
                typedef __attribute__((address_space(3))) volatile int LdsInt;
                if (kLds > 0 && sp < kLds) return *(LdsInt *)&g_stackLds[sp * kBlock + threadIdx.x];
                return priv[sp - kLds];
        }
};

enum : int { kStepGo = 0, kStepDone = 1, kStepOverflow = 2 };

// One step of traverse_bvh2_core's loop (kernel.mojo:256-358): a whole leaf, or one interior
// node with its two child boxes.
template <bool kAny, bool kStats, int kLds>
__device__ __forceinline__ int step(Lane &L, Stack<kLds> &S, const NodePair *__restrict__ nodes,
                                    const TriRec *__restrict__ tris) {
        if (L.cur < 0) {
                const TriRec *leaf = tris + ~L.cur;
                const int count = __float_as_int(leaf[0].p0.w);
                for (int j = 0; j < count; ++j) {
                        const float4 a = leaf[j].p0;
                        const float4 b = leaf[j].p1;
                        const float4 c = leaf[j].p2;
                        if (__float_as_int(c.w) & kTriSkip) continue;
                        if (kStats) ++L.nTris;
                        float t, u, v;
                        if (tri_test(L.r, a, b, c, L.tBest, t, u, v)) {
                                L.tBest = t;
                                L.bu = u;
                                L.bv = v;
                                L.hitPrim = __float_as_int(b.w);
                                if (kAny) return kStepDone;
                        }
                }
                if (L.sp == 0) return kStepDone;
                L.cur = S.get(--L.sp);
                return kStepGo;
        }
        if (kStats) ++L.nInterior;
        const NodePair np = nodes[L.cur];
        float tl, tr;
        const bool lh = slab(np.a.x, np.a.y, np.a.z, np.a.w, np.b.x, np.b.y, L.r, L.tBest, tl);
        const bool rh = slab(np.b.z, np.b.w, np.c.x, np.c.y, np.c.z, np.c.w, L.r, L.tBest, tr);
        if (lh && rh) {
                if (L.sp == GZ_STACK_CAPACITY) return kStepOverflow;
                const bool leftFirst = tl <= tr;
                L.cur = leftFirst ? np.d.x : np.d.y;
                S.put(L.sp++, leftFirst ? np.d.y : np.d.x);
        } else if (lh) {
                L.cur = np.d.x;
        } else if (rh) {
                L.cur = np.d.y;
        } else {
                if (L.sp == 0) return kStepDone;
                L.cur = S.get(--L.sp);
        }
        return kStepGo;
}

template <bool kAny, bool kStats>
__device__ __forceinline__ void lane_finish(const Lane &L, int64_t i, bool overflow, GzHit *__restrict__ hits,
                                            uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats,
                                            int *__restrict__ err) {
        int prim = L.hitPrim;
        if (overflow) {
                atomicOr(err, 1);
                prim = -1;
        }
        if (kAny) {
                occ[i] = prim >= 0 ? 1 : 0;
        } else {
                GzHit h;
                h.prim = prim;
                h.t = prim >= 0 ? L.tBest : L.tMax;  // kernel.mojo:360-364
                h.u = prim >= 0 ? L.bu : 0.0f;
                h.v = prim >= 0 ? L.bv : 0.0f;
                hits[i] = h;
        }
        if (kStats) {
                GzTraceStats st;
                st.interior = L.nInterior;
                st.tris = L.nTris;
                stats[i] = st;
        }
}

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
static void launch_variant(Scene *s, const Ray_C *d_rays, const float *d_tMax, int64_t n, GzHit *d_hits,
                           uint8_t *d_occ, GzTraceStats *d_stats, hipStream_t stream) {
        const GzTraceConfig &c = s->traceConfig;
        if (c.persistent == 2) {
                auto kern = k_trace_strided<kAny, kStats, kLds>;
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
                const unsigned g = (unsigned)std::min<int64_t>(want, grid_for(n));
                hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, stream, s->dNodes, s->dTris, s->rootDesc,
                                   s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats, s->dErr);
        } else if (c.persistent == 1) {
                auto kern = k_trace_persistent<kAny, kStats, kLds>;
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
                const unsigned g = (unsigned)std::min<int64_t>(want, grid_for(n));
                unsigned long long *ctr = s->dCounters + (s->launchSeq++ % kCounterRing);
                hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, ctr);
                hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, stream, s->dNodes, s->dTris, s->rootDesc,
                                   s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats, s->dErr, ctr,
                                   c.refillLanes > 0 && c.refillLanes <= 64 ? c.refillLanes : 16,
                                   c.chunk >= 1 && c.chunk <= 4096 ? c.chunk : 64);
        } else {
                hipLaunchKernelGGL((k_trace<kAny, kStats, kLds>), dim3(grid_for(n)), dim3(kBlock), 0, stream, s->dNodes,
                                   s->dTris, s->rootDesc, s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats,
                                   s->dErr);
        }
}

template <bool kAny, bool kStats>
static void launch_lds(Scene *s, const Ray_C *d_rays, const float *d_tMax, int64_t n, GzHit *d_hits, uint8_t *d_occ,
                       GzTraceStats *d_stats, hipStream_t stream) {
        if (s->traceConfig.ldsStack >= 16)
                launch_variant<kAny, kStats, 16>(s, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream);
        else if (s->traceConfig.ldsStack >= 8)
                launch_variant<kAny, kStats, 8>(s, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream);
        else
                launch_variant<kAny, kStats, 0>(s, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream);
}

int launch_trace(Scene *s, const Ray_C *d_rays, const float *d_tMax, int64_t n, GzHit *d_hits, uint8_t *d_occ,
                 GzTraceStats *d_stats, bool anyHit, hipStream_t stream) {
        if (n <= 0) return GZ_OK;
        if (n > (int64_t)kBlock * 0x7fffffffLL) {
                set_error("traversal batch of %lld rays is too large", (long long)n);
                return GZ_E_INVALID;
        }
        if (s->empty) {
                hipLaunchKernelGGL(k_miss, dim3(grid_for(n)), dim3(kBlock), 0, stream, d_tMax, n,
                                   anyHit ? nullptr : d_hits, anyHit ? d_occ : nullptr, d_stats);
        } else if (anyHit) {
                if (d_stats)
                        launch_lds<true, true>(s, d_rays, d_tMax, n, nullptr, d_occ, d_stats, stream);
                else
                        launch_lds<true, false>(s, d_rays, d_tMax, n, nullptr, d_occ, nullptr, stream);
        } else {
                if (d_stats)
                        launch_lds<false, true>(s, d_rays, d_tMax, n, d_hits, nullptr, d_stats, stream);
                else
                        launch_lds<false, false>(s, d_rays, d_tMax, n, d_hits, nullptr, nullptr, stream);
        }
        return hip_ok(hipGetLastError(), "traversal launch") ? GZ_OK : GZ_E_HIP;
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

static int check_overflow_sync(Scene *s) {
        int flag = 0;
        if (!hip_ok(hipMemcpyAsync(&flag, s->dErr, sizeof(int), hipMemcpyDeviceToHost, s->stream), "read overflow flag") ||
            !hip_ok(hipStreamSynchronize(s->stream), "hipStreamSynchronize"))
                return GZ_E_HIP;
        if (flag) {
                (void)hipMemsetAsync(s->dErr, 0, sizeof(int), s->stream);
                (void)hipStreamSynchronize(s->stream);
                set_error("traversal stack overflow: a ray needed more than %d entries", GZ_STACK_CAPACITY);
                return GZ_E_OVERFLOW;
        }
        return GZ_OK;
}

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
