// Device-side BVH2 traversal (one ray per lane): ray setup, the reference's slab and
// Moller-Trumbore tests, the per-lane stack and one traversal step. Shared by the batch
// traversal kernels and the path-tracing kernels that trace rays inline.
//
// Semantics follow traverse_bvh2_core (Sources/mojoKernel/kernel.mojo:223-364) step for step;
// see gz_traverse.hip for the layout and the bit-exactness argument.
#pragma once

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

// v_max3_f32 / v_min3_f32 of three values none of which is NaN. fmaxf and fminf of the same values
// give the same results (up to the sign of a zero), but the compiler quiets each operand with an
// extra v_max_f32 first (IEEE mode), which doubles the work of the slab test.
__device__ __forceinline__ float max3_finite(float a, float b, float c) {
        float r;
        asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
        return r;
}
__device__ __forceinline__ float min3_finite(float a, float b, float c) {
        float r;
        asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
        return r;
}

// v_min_f32 of a finite value and a value that is finite, infinite or a quiet NaN (which it ignores,
// as fminf does); a signalling NaN would not be ignored, so tMax enters as Lane::tBest, quieted when
// the lane starts.
__device__ __forceinline__ float min_quiet(float a, float q) {
        float r;
        asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(q));
        return r;
}

// slab_fast for the two child boxes of an interior record (safe rays only).
__device__ __forceinline__ void slab_pair_fast(const float4 &a, const float4 &b, const float4 &c, const RayState &r,
                                               float tMax, bool &lh, float &tl, bool &rh, float &tr) {
        // left: (a.x, a.y, a.z) - (a.w, b.x, b.y); right: (b.z, b.w, c.x) - (c.y, c.z, c.w)
        const float lNx = (r.nx ? a.x : a.w) * r.rdx - r.ordx, lFx = (r.nx ? a.w : a.x) * r.rdx - r.ordx;
        const float lNy = (r.ny ? a.y : b.x) * r.rdy - r.ordy, lFy = (r.ny ? b.x : a.y) * r.rdy - r.ordy;
        const float lNz = (r.nz ? a.z : b.y) * r.rdz - r.ordz, lFz = (r.nz ? b.y : a.z) * r.rdz - r.ordz;
        const float rNx = (r.nx ? b.z : c.y) * r.rdx - r.ordx, rFx = (r.nx ? c.y : b.z) * r.rdx - r.ordx;
        const float rNy = (r.ny ? b.w : c.z) * r.rdy - r.ordy, rFy = (r.ny ? c.z : b.w) * r.rdy - r.ordy;
        const float rNz = (r.nz ? c.x : c.w) * r.rdz - r.ordz, rFz = (r.nz ? c.w : c.x) * r.rdz - r.ordz;
        tl = max3_finite(lNx, lNy, max3_finite(lNz, 0.0f, 0.0f));
        tr = max3_finite(rNx, rNy, max3_finite(rNz, 0.0f, 0.0f));
        lh = tl <= min3_finite(lFx, lFy, min_quiet(lFz, tMax)) * 1.0000003f;
        rh = tr <= min3_finite(rFx, rFy, min_quiet(rFz, tMax)) * 1.0000003f;
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

__device__ __forceinline__ void lane_init(Lane &L, const Ray_C ray, float tMax, int rootDesc, bool sceneSafe);

__device__ __forceinline__ void lane_start(Lane &L, const Ray_C *__restrict__ rays, const float *__restrict__ tMaxIn,
                                           int64_t i, int rootDesc, bool sceneSafe) {
        lane_init(L, rays[i], tMaxIn[i], rootDesc, sceneSafe);
}

__device__ __forceinline__ void lane_init(Lane &L, const Ray_C ray, float tMax, int rootDesc, bool sceneSafe) {
        L.tMax = tMax;
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
        L.tBest = __builtin_canonicalizef(L.tMax);  // a signalling NaN quieted: the slab tests ignore it
        L.bu = 0.0f;
        L.bv = 0.0f;
        L.cur = rootDesc;
        L.sp = 0;
        L.nInterior = 0;
        L.nTris = 0;
}

// The first entries of every lane's stack, [entry][lane of the block]: a lane's pushes and pops
// hit its own column, so a wave's 64 lanes touch 64 consecutive words (no bank conflicts).
// Two sizes: 16 rows (16 KB per 256-lane block) for the batch traversal kernels, which run many
// waves per CU, and 32 rows for the path kernels, whose deep traversals of large scenes would
// otherwise spill to private (scratch) memory and which run 2 to 4 waves per SIMD.
constexpr int kLdsMax = 16;
constexpr int kLdsWide = 32;
__shared__ int g_stackLds[kLdsMax * kBlock];
__shared__ int g_stackLdsWide[kLdsWide * kBlock];

// Traversal stack of GZ_STACK_CAPACITY entries: the first kLds live in LDS, the rest in
// per-lane private memory.
template <int kLds>
struct Stack {
        static_assert(kLds <= kLdsWide, "LDS stack rows");
        int priv[GZ_STACK_CAPACITY - kLds > 0 ? GZ_STACK_CAPACITY - kLds : 1];
        __device__ __forceinline__ static int *rows() {
                if constexpr (kLds > kLdsMax)
                        return g_stackLdsWide;
                else
                        return g_stackLds;
        }
        __device__ __forceinline__ void put(int sp, int v) {
                if (kLds > 0 && sp < kLds)
                        rows()[sp * kBlock + threadIdx.x] = v;
                else
                        priv[sp - kLds] = v;
        }
        __device__ __forceinline__ int get(int sp) const {
                // The volatile LDS read keeps the two loads apart: merged, they become one generic
                // (flat) load that pays the latency of both memories on every pop.
                typedef __attribute__((address_space(3))) volatile int LdsInt;
                if (kLds > 0 && sp < kLds) return *(LdsInt *)&rows()[sp * kBlock + threadIdx.x];
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
        bool lh, rh;
        if (L.r.safe) {
                slab_pair_fast(np.a, np.b, np.c, L.r, L.tBest, lh, tl, rh, tr);
        } else {
                lh = slab_exact(np.a.x, np.a.y, np.a.z, np.a.w, np.b.x, np.b.y, L.r, L.tBest, tl);
                rh = slab_exact(np.b.z, np.b.w, np.c.x, np.c.y, np.c.z, np.c.w, L.r, L.tBest, tr);
        }
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

}  // namespace gz
