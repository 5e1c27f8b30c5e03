// Device code of the MIS path tracer shared by the single-scene kernels (gz_mis.hip) and the
// multi-scene batch (gz_batch.hip): the bounce of the reference's VolumePathIntegrator
// (Sources/libgonzales/Integrator/VolumePathIntegrator.swift:35-463), one path per lane.
//
// A bounce of a path, in the reference's order:
//   closest hit + surface interaction ......... BoundingHierarchy.swift:387-456, Triangle.swift:363-474
//   miss: infinite light if bounce 0/specular .. VolumePathIntegrator.swift:380-393
//   stop at maxDepth ........................... :400-402
//   emission on bounce 0/specular (one-sided) .. :331-340, AreaLight.swift:13-19
//   BSDF (area-light surfaces use materials[0]) :341-348
//   one light by power (own sampler at Morton 0) PowerLightSampler.swift:41-49
//   light sample + shadow ray (MIS weight) ...... :67-89, :173-195, Visibility.swift:5-24
//   BSDF sample + ray (MIS weight, light pdf 0
//     for triangles unless corrected) .......... :91-133, :197-211, Sampleable.swift:49-69
//   continuation, spawn offset 1e-4 n ........... :358-369, Interaction.swift:32-53
//   Russian roulette after bounce 1 ............. :269-283 (deterministic stream here)
// Arithmetic follows the reference's operation order without contraction; sqrt and division are
// IEEE-correct; cos/sin come from the device math library.
//
// The shadow ray and the BSDF ray are traced inline with the same traversal as the batch
// kernels. A BSDF ray toward a chosen triangle light is skipped when its MIS weight is 0 (the
// reference's light pdf quirk): it would contribute exactly 0.
#pragma once

#include <cmath>

#include "bsdf_device.h"
#include "traverse_device.h"
#include "zsobol_device.h"

#pragma clang fp contract(off)

namespace gz {

const uint32_t *device_sobol_table();  // gz_zsobol.hip

constexpr int kMisBlock = 256;
// Scene data of the MIS integrator on the device.
struct DevLight {
        float4 p0;  // w: area
        float4 p1;
        float4 p2;
        float4 L;   // brightness
};

struct MisArgs {
        const NodePair *nodes;
        const TriRec *tris;
        int rootDesc;
        bool sceneSafe;
        bool empty;
        const PrimId_C *prims;
        int64_t primCount;
        const float4 *primTris;  // Scene::dPrimTris
        const uint8_t *primClass;  // per primitive: prim_class_of (set with the materials)
        // Box around the primitives of class kClsLight or kClsFull, padded outward (lo > hi when
        // there are none): a BSDF ray may stop at a hit closer than its entry into this box
        // (bsdf_ray_stop).
        float emitLo[3], emitHi[3];
        const DevMesh *meshes;
        int64_t meshCount;
        const GzMaterial *mats;
        int nMats;
        const DevLight *lights;
        int nLights;
        const float *power;       // per choice (area lights, then the sky)
        const float *cumulative;  // running sum of power
        int nChoices;
        float totalPower;
        bool hasSky;
        float sky[3];
        const uint32_t *sobol;
        ZSobolParams zs;
        int maxDepth;
        int flags;
        uint64_t rrSeed;
        int *err;
};

// Per-lane traversal state shared by the rays of one bounce: one stack (so the three inline
// traversals do not each get their own private array) and the work counters (interior visits
// and triangle tests, all rays).
constexpr int kLdsMis = kLdsWide;

struct Work {
        Stack<kLdsMis> S;
        uint32_t interior = 0, tris = 0;
};

// Closest or any hit of one ray (the traversal of the batch kernels, inline).
template <bool kAny, bool kStats>
__device__ __forceinline__ bool trace_ray(const MisArgs &a, const Ray_C &ray, float tMax, int &prim, float &t,
                                          float &u, float &v, Work &w) {
        if (a.empty) return false;
        Lane L;
        lane_init(L, ray, tMax, a.rootDesc, a.sceneSafe);
        int st;
        do {
                st = step<kAny, kStats, kLdsMis>(L, w.S, a.nodes, a.tris);
        } while (st == kStepGo);
        if (kStats) {
                w.interior += L.nInterior;
                w.tris += L.nTris;
        }
        if (st == kStepOverflow) {
                atomicOr(a.err, 1);
                return false;
        }
        prim = L.hitPrim;
        t = L.tBest;
        u = L.bu;
        v = L.bv;
        return prim >= 0;
}

// Where a BSDF-sampled ray may stop looking for its closest hit. The BSDF strategy needs only to
// know whether that hit is an area light or a primitive whose surface interaction fails (both
// inside the emitter box), or an ordinary surface, which adds nothing. Any hit closer than the
// ray's entry into the padded box is closer than every such primitive, so the closest hit is an
// ordinary surface once one such hit is found: the ray stops there (the hit it reports is then an
// ordinary surface, not necessarily the closest). Returns the entry distance, +inf for a ray that
// misses the box, 0 (never stop) where the slab test is not clear-cut.
__device__ __forceinline__ float bsdf_ray_stop(const MisArgs &a, const Ray_C &r) {
        if (!(a.emitLo[0] <= a.emitHi[0])) return INFINITY;  // no such primitives
        const float o[3] = {r.orgX, r.orgY, r.orgZ}, d[3] = {r.dirX, r.dirY, r.dirZ};
        float t0 = 0.0f, t1 = INFINITY;
        for (int k = 0; k < 3; ++k) {
                if (d[k] == 0.0f) {
                        if (o[k] < a.emitLo[k] || o[k] > a.emitHi[k]) return INFINITY;
                        continue;
                }
                const float ta = (a.emitLo[k] - o[k]) / d[k], tb = (a.emitHi[k] - o[k]) / d[k];
                const float tn = ta < tb ? ta : tb, tf = ta < tb ? tb : ta;
                if (!(tn == tn) || !(tf == tf)) return 0.0f;
                t0 = tn > t0 ? tn : t0;
                t1 = tf < t1 ? tf : t1;
        }
        return t0 <= t1 ? t0 : INFINITY;
}

// trace_ray's closest hit that stops once a hit closer than tStop is found (see bsdf_ray_stop).
template <bool kStats>
__device__ __forceinline__ bool trace_ray_until(const MisArgs &a, const Ray_C &ray, float tStop, int &prim, float &t,
                                                float &u, float &v, Work &w) {
        if (a.empty) return false;
        Lane L;
        lane_init(L, ray, INFINITY, a.rootDesc, a.sceneSafe);
        int st;
        do {
                st = step<false, kStats, kLdsMis>(L, w.S, a.nodes, a.tris);
        } while (st == kStepGo && !(L.hitPrim >= 0 && L.tBest < tStop));
        if (kStats) {
                w.interior += L.nInterior;
                w.tris += L.nTris;
        }
        if (st == kStepOverflow) {
                atomicOr(a.err, 1);
                return false;
        }
        prim = L.hitPrim;
        t = L.tBest;
        u = L.bu;
        v = L.bv;
        return prim >= 0;
}

struct Interaction {
        f3 position, normal, shadingNormal, outgoing, dpdu;
        int material;
        int areaLight;
};

// The vertices of primitive pr through its mesh arrays (BoundingHierarchy.swift's triangle
// lookup); false for a mesh or index out of range. Used once per scene to fill Scene::dPrimTris.
__device__ __forceinline__ bool tri_points(const DevMesh *__restrict__ meshes, int64_t meshCount, const PrimId_C &pr,
                                           f3 &p0, f3 &p1, f3 &p2) {
        int64_t mesh, base;
        if (pr.type == 0) {
                mesh = pr.id1;
                base = pr.id2;
        } else {
                mesh = pr.id2 >> 32;
                base = (pr.id2 & 0xFFFFFFFFLL) * 3;
        }
        if (mesh < 0 || mesh >= meshCount) return false;
        const DevMesh m = meshes[mesh];
        if (base < 0 || base + 2 >= m.vertexIndexCount) return false;
        const int64_t i0 = m.vertexIndices[base], i1 = m.vertexIndices[base + 1], i2 = m.vertexIndices[base + 2];
        if (i0 < 0 || i1 < 0 || i2 < 0 || i0 * 4 + 2 >= m.pointFloats || i1 * 4 + 2 >= m.pointFloats ||
            i2 * 4 + 2 >= m.pointFloats)
                return false;
        p0 = mk(m.points[i0 * 4], m.points[i0 * 4 + 1], m.points[i0 * 4 + 2]);
        p1 = mk(m.points[i1 * 4], m.points[i1 * 4 + 1], m.points[i1 * 4 + 2]);
        p2 = mk(m.points[i2 * 4], m.points[i2 * 4 + 1], m.points[i2 * 4 + 2]);
        return true;
}

// BoundingHierarchy.intersect -> Triangle.computeSurfaceInteraction for meshes without vertex
// normals or uvs (Triangle.swift:363-474). false where the reference returns nil.
__device__ bool surface_interaction(const MisArgs &a, const Ray_C &ray, int prim, float u, float v, Interaction &si) {
        if (prim < 0 || prim >= a.primCount) return false;
        // the primId and the vertices: two loads that do not wait on each other
        const PrimId_C pr = a.prims[prim];
        const float4 q0 = a.primTris[3 * (int64_t)prim], q1 = a.primTris[3 * (int64_t)prim + 1],
                     q2 = a.primTris[3 * (int64_t)prim + 2];
        if (q0.w != 0.0f) {
                atomicOr(a.err, 2);
                return false;
        }
        const f3 p0 = mk(q0.x, q0.y, q0.z), p1 = mk(q1.x, q1.y, q1.z), p2 = mk(q2.x, q2.y, q2.z);
        const float b0 = 1.0f - u - v, b1 = u, b2 = v;
        const f3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
        const f3 pHit = add(add(mul(p0, b0), mul(p1, b1)), mul(p2, b2));
        const f3 n = normalized(cross(dp02, dp12));
        const f3 gn = cross(sub(p2, p0), sub(p1, p0));
        if (len2(gn) == 0.0f) return false;
        f3 dpdu, dpdv;
        coordinate_system(normalized(gn), dpdu, dpdv);
        const f3 sn = n;
        f3 S = normalized(dpdu);
        f3 T = cross(S, sn);
        if (len2(T) > 0.0f) {
                T = normalized(T);
                S = cross(T, sn);
        } else {
                coordinate_system(sn, S, T);
        }
        si.position = pHit;
        si.normal = normalized(n);
        si.shadingNormal = normalized(sn);
        si.outgoing = normalized(neg(mk(ray.dirX, ray.dirY, ray.dirZ)));
        si.dpdu = S;
        if (pr.type == 3) {
                si.areaLight = (int)pr.id1;
                si.material = 0;
        } else {
                si.areaLight = -1;
                si.material = (int)pr.materialIndex;
        }
        return true;
}

// The class byte of a primitive (MisScene::dPrimClass, rebuilt by every gz_gpu_set_mis_scene):
// bits 0-1 the BSDF set its hits shade with (hit_bsdf_set: the material's, materials[0] for an
// area-light triangle, an index out of range read as 0), kClsLight for an area-light triangle,
// kClsFull where surface_interaction does not return an ordinary surface (vertex indices out of
// range, or a zero geometric normal: the same float operations as there).
enum : uint8_t { kClsSetMask = 3, kClsLight = 4, kClsFull = 8 };

__device__ __forceinline__ uint8_t prim_class_of(const PrimId_C &pr, const float4 *q, const GzMaterial *mats,
                                                 int nMats) {
        int m = pr.type == 3 ? 0 : (int)pr.materialIndex;
        if (m < 0 || m >= nMats) m = 0;
        const int t = mats[m].type;
        uint8_t c = (uint8_t)(t == GZ_MAT_DIFFUSE ? kBsdfDiffuse : t == GZ_MAT_COATED_CONDUCTOR ? kBsdfAll : kBsdfNoLayer);
        if (pr.type == 3) c |= kClsLight;
        const float4 q0 = q[0], q1 = q[1], q2 = q[2];
        if (q0.w != 0.0f) return (uint8_t)(c | kClsFull);
        const f3 p0 = mk(q0.x, q0.y, q0.z), p1 = mk(q1.x, q1.y, q1.z), p2 = mk(q2.x, q2.y, q2.z);
        const f3 gn = cross(sub(p2, p0), sub(p1, p0));
        if (len2(gn) == 0.0f) c |= kClsFull;
        return c;
}

// Whether the closest hit `prim` of a BSDF-sampled ray needs its surface interaction for the BSDF
// strategy's term: an area-light triangle (its one-sided emission), or a primitive on which
// surface_interaction fails (the term then falls back to the sky, as for a miss). Any other hit
// adds nothing, and the class byte says so without loading the hit's primId and vertices.
__device__ __forceinline__ bool bsdf_hit_needs_interaction(const MisArgs &a, int prim) {
        if (prim < 0 || prim >= a.primCount) return true;
        return (a.primClass[prim] & (kClsLight | kClsFull)) != 0;
}

// The material's BSDF in its shading frame (BsdfFrame.swift:96-104, ShadingFrame.swift:27-77,
// FramedBsdf.swift:18-83); the BSDFs themselves are in bsdf_device.h.
struct Bsdf {
        f3 x, y, z, ng;
        Bx b;
};

__device__ __forceinline__ void bsdf_make(const MisArgs &a, const Interaction &si, Bsdf &b) {
        b.y = normalized(si.dpdu);
        b.z = si.shadingNormal;
        b.x = cross(b.y, b.z);
        b.ng = si.normal;
        const int mi = si.material >= 0 && si.material < a.nMats ? si.material : 0;
        b.b = bx_from_material(a.mats[mi]);
}
__device__ __forceinline__ f3 to_local(const Bsdf &b, f3 w) { return normalized(mk(dot(w, b.x), dot(w, b.y), dot(w, b.z))); }
__device__ __forceinline__ f3 to_world(const Bsdf &b, f3 l) { return add(add(mul(b.x, l.x), mul(b.y, l.y)), mul(b.z, l.z)); }
// evaluateWorldSpace: every BSDF keeps the protocol's isReflective = true and
// isTransmissive = false, so only the reflection side is evaluated.
template <int kAll>
__device__ __forceinline__ f3 bsdf_eval_world(const Bsdf &b, f3 wo, f3 wi, LayerRng &r) {
        const f3 ol = to_local(b, wo), il = to_local(b, wi);
        const bool reflect = dot(wi, b.ng) * dot(wo, b.ng) > 0.0f;
        return reflect ? bx_f<kAll>(b.b, ol, il, r) : mk(0.0f, 0.0f, 0.0f);
}
template <int kAll>
__device__ __forceinline__ float bsdf_pdf_world(const Bsdf &b, f3 wo, f3 wi, LayerRng &r) {
        const f3 il = to_local(b, wi), ol = to_local(b, wo);
        if (ol.z == 0.0f) return 0.0f;
        return bx_pdf<kAll>(b.b, ol, il, r);
}
template <int kAll>
__device__ __forceinline__ void bsdf_sample_world(const Bsdf &b, f3 wo, float u0, float u1, float u2, f3 &f, f3 &wi,
                                                  float &pdf, LayerRng &r) {
        const f3 ol = to_local(b, wo);
        const BSample s = bx_sample<kAll>(b.b, ol, u0, u1, u2, r);
        f = s.f;
        wi = to_world(b, s.wi);
        pdf = s.pdf;
}

// Streams of the layered BSDF for the BSDF calls of one bounce.
enum : int { kCallLightF = 1, kCallLightPdf = 2, kCallMisSample = 3, kCallContinue = 4 };

// PowerLightSampler.chooseLight + lowerBound (PowerLightSampler.swift:4-49)
__device__ __forceinline__ int choose_light(const MisArgs &a, float u, float &pdf) {
        const float key = u * a.totalPower;
        int lo = 0, hi = a.nChoices, idx = -1;
        while (lo < hi) {
                const int mid = lo + (hi - lo) / 2;
                const float cm = a.cumulative[mid];
                if (cm == key) {
                        idx = mid;
                        break;
                }
                if (cm < key)
                        lo = mid + 1;
                else
                        hi = mid;
        }
        if (idx < 0) idx = lo;
        if (idx >= a.nChoices) idx = a.nChoices - 1;
        pdf = a.power[idx] / a.totalPower;
        return idx;
}

// InfiniteLight.equalAreaSquareToSphere (InfiniteLight.swift:112-127); `unit` applies the factor
// sqrt(2 - r^2) to x and y that the reference omits (GZ_MIS_FIX_SKY_DIRECTION).
__device__ __forceinline__ f3 equal_area_square_to_sphere(float tu, float tv, bool unit) {
        const float u = 2.0f * tu - 1.0f, v = 2.0f * tv - 1.0f;
        const float au = fabsf(u), av = fabsf(v);
        const float sd = 1.0f - (au + av);
        const float ad = fabsf(sd);
        const float r = 1.0f - ad;
        const float phi = (r == 0.0f ? 1.0f : (av - au) / r + 1.0f) * (kPi / 4.0f);
        const float zz = 1.0f - r * r;
        const float z = copysignf(zz > 0.0f ? zz : 0.0f, sd);
        const float cp = cosf(phi), sp = sinf(phi);
        float rr = r;
        if (unit) {
                const float s = 2.0f - r * r;
                rr = r * sqrtf(s > 0.0f ? s : 0.0f);
        }
        return mk(copysignf(rr * cp, u), copysignf(rr * sp, v), z);
}

struct LightSample {
        f3 L, dir, target;
        float pdf;
};

// AreaLight.sample (AreaLight.swift:21-32) with Triangle.sample and Sampleable.sample(point:)
// (Triangle.swift:517-544, Sampleable.swift:33-44), or the constant sky (InfiniteLight.swift:39-65).
__device__ __forceinline__ void light_sample(const MisArgs &a, int choice, f3 point, float u0, float u1, LightSample &ls) {
        if (choice < a.nLights) {
                const DevLight dl = a.lights[choice];
                const f3 p0 = mk(dl.p0.x, dl.p0.y, dl.p0.z), p1 = mk(dl.p1.x, dl.p1.y, dl.p1.z);
                const f3 p2 = mk(dl.p2.x, dl.p2.y, dl.p2.z);
                const float su0 = sqrtf(u0);
                const float b0 = 1.0f - su0, b1 = u1 * su0;
                const f3 pos = add(add(mul(p0, b0), mul(p1, b1)), mul(p2, 1.0f - b0 - b1));
                const f3 ln = normalized(cross(sub(p1, p0), sub(p2, p0)));
                float pdf = 1.0f / dl.p0.w;
                const f3 inc = normalized(sub(pos, point));
                const float d2 = len2(sub(point, pos));
                const float angle = absdot(ln, neg(inc));
                pdf = pdf * (d2 / angle);
                ls.dir = normalized(sub(pos, point));
                ls.target = pos;
                ls.L = dot(ln, neg(ls.dir)) > 0.0f ? mk(dl.L.x, dl.L.y, dl.L.z) : mk(0.0f, 0.0f, 0.0f);
                ls.pdf = pdf;
                return;
        }
        const f3 d = equal_area_square_to_sphere(u0, u1, (a.flags & GZ_MIS_FIX_SKY_DIRECTION) != 0);
        ls.dir = d;
        ls.pdf = 1.0f / (4.0f * kPi);
        ls.target = add(point, mul(d, 100.0f));
        ls.L = mk(a.sky[0], a.sky[1], a.sky[2]);
}

// Visibility.occluded via spawnRay(from:target:) (Visibility.swift:5-24, Interaction.swift:15-24):
// origin from + 1e-4 (target - from), unnormalised direction, tMax 1 - 1e-5.
template <bool kStats>
__device__ __forceinline__ bool occluded(const MisArgs &a, f3 from, f3 target, Work &w) {
        const f3 o = add(from, mul(sub(target, from), 0.0001f));
        const f3 d = sub(target, o);
        Ray_C r{o.x, o.y, o.z, d.x, d.y, d.z};
        int prim;
        float t, u, v;
        return trace_ray<true, kStats>(a, r, 1.0f - 0.00001f, prim, t, u, v, w);
}

// Interaction.spawnRay(inDirection:) (Interaction.swift:32-53)
__device__ __forceinline__ Ray_C spawn(const Interaction &si, f3 dir) {
        const f3 off = mul(si.normal, 0.0001f);
        const f3 o = dot(dir, si.normal) > 0.0f ? add(si.position, off) : add(si.position, neg(off));
        return Ray_C{o.x, o.y, o.z, dir.x, dir.y, dir.z};
}

// Light pdf of a BSDF-sampled direction (Sampleable.swift:49-69): 0 for triangles in the
// reference (tHit = 0); with GZ_MIS_FIX_LIGHT_PDF the chosen triangle is intersected.
__device__ float light_pdf(const MisArgs &a, int choice, const Interaction &si, f3 dir) {
        if (choice >= a.nLights) return 1.0f / (4.0f * kPi);
        if (!(a.flags & GZ_MIS_FIX_LIGHT_PDF)) return 0.0f;
        const DevLight dl = a.lights[choice];
        const f3 p0 = mk(dl.p0.x, dl.p0.y, dl.p0.z), p1 = mk(dl.p1.x, dl.p1.y, dl.p1.z);
        const f3 p2 = mk(dl.p2.x, dl.p2.y, dl.p2.z);
        const Ray_C r = spawn(si, dir);
        const f3 e1 = sub(p1, p0), e2 = sub(p2, p0);
        const f3 d = mk(r.dirX, r.dirY, r.dirZ), o = mk(r.orgX, r.orgY, r.orgZ);
        const f3 pv = cross(d, e2);
        const float det = dot(e1, pv);
        if (det > -1e-12f && det < 1e-12f) return 0.0f;
        const float inv = 1.0f / det;
        const f3 tv = sub(o, p0);
        const float uu = dot(tv, pv) * inv;
        if (uu < 0.0f || uu > 1.0f) return 0.0f;
        const f3 qv = cross(tv, e1);
        const float vv = dot(d, qv) * inv;
        if (vv < 0.0f || uu + vv > 1.0f) return 0.0f;
        const float t = dot(e2, qv) * inv;
        if (t <= 0.0f) return 0.0f;
        const f3 hitp = add(o, mul(d, t));
        const float d2 = len2(sub(si.position, hitp));
        const f3 ln = normalized(cross(sub(p0, p2), sub(p1, p2)));
        const float angle = absdot(ln, neg(dir));
        const float density = d2 / (angle * dl.p0.w);
        return isinf(density) ? 0.0f : density;
}

// Deterministic Russian-roulette draw: top 24 bits of SplitMix64(seed, pixel, sample, bounce).
__device__ __forceinline__ float rr_draw(uint64_t seed, int px, int py, int si, int bounce) {
        uint64_t z = seed ^ ((uint64_t)(uint32_t)px << 40) ^ ((uint64_t)(uint32_t)py << 20) ^ (uint64_t)(uint32_t)si;
        z += 0x9E3779B97F4A7C15ull * (uint64_t)(uint32_t)(bounce + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

// Path-replay adjoint state (gz_gpu_mis_adjoint_device): per path its forward radiance and the
// adjoint of its radiance; gradients accumulate per material as (dL/dR r, g, b, dL/droughness).
// A replay record of the forward pass (gz_gpu_mis_trace_record_device): for bounce level L < recLevels
// and path i, at index L * recCount + i, the closest hit of the path's ray (recHits), the closest hit
// of its BSDF-sampled ray (recMhits) and its shadow ray's visibility (recOcc); prim kRecNone or occ
// kRecNoOcc where the forward pass did not trace that ray (the replay traces it then).
constexpr int kRecNone = -2;
constexpr uint8_t kRecNoOcc = 2;

struct AdjArgs {
        const GzPathAdjoint *paths;
        double *grad;  // [materials][4] in global memory
        const GzHit *recHits = nullptr;
        const GzHit *recMhits = nullptr;
        const uint8_t *recOcc = nullptr;
        int recLevels = 0;
        int64_t recCount = 0;
};

// The record's entry of bounce `bounce` of path i, or -1 when the record has no such level.
__device__ __forceinline__ int64_t rec_index(const AdjArgs &adj, int bounce, int64_t i) {
        return bounce < adj.recLevels ? (int64_t)bounce * adj.recCount + i : -1;
}
constexpr int kAdjLdsMats = 128;

// Log-derivatives in alpha of an isotropic Trowbridge-Reitz conductor (MicrofacetReflection.swift:
// 15-26, TrowbridgeReitzDistribution.swift:9-26) for local directions wo, wi:
//   f:   d ln D(h) + d ln G(wo, wi)    (F and the cosines do not depend on alpha)
//   pdf: d ln D(h) + d ln G1(wo)       (pdf = D G1(wo) / (4 |cos wo|))
// D = 1 / (pi a^2 cos^4 (1 + t/a^2)^2), t = tan^2(theta_h); Lambda = (sqrt(1 + a^2 tan^2) - 1) / 2.
struct DlogAlpha {
        float f, pdf;
};
__device__ __forceinline__ DlogAlpha dlog_dalpha(float al, f3 wo, f3 wi) {
        f3 h = add(wo, wi);
        if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return DlogAlpha{0.0f, 0.0f};
        h = normalized(h);
        const float t = tan2_theta(h);
        float dD = -2.0f / al;
        if (!isinf(t)) dD += 4.0f * t / (al * (al * al + t));
        float lam[2], dlam[2];
        const f3 ws[2] = {wo, wi};
        for (int k = 0; k < 2; ++k) {
                const float t2 = tan2_theta(ws[k]);
                if (isinf(t2)) {
                        lam[k] = 0.0f;
                        dlam[k] = 0.0f;
                        continue;
                }
                const float s = sqrtf(1.0f + al * al * t2);
                lam[k] = (s - 1.0f) / 2.0f;
                dlam[k] = al * t2 / (2.0f * s);
        }
        return DlogAlpha{dD - (dlam[0] + dlam[1]) / (1.0f + lam[0] + lam[1]), dD - dlam[0] / (1.0f + lam[0])};
}

// Adjoint of one bounce's BSDF at a vertex of material m: the light-strategy term, the
// BSDF-strategy term and everything after the vertex (cont) each carry the BSDF once; the
// derivative of the path's radiance is each term times d ln f / d theta at its direction, plus,
// for the two MIS terms, the derivative of their power-heuristic weights through the BSDF pdf
// (the reference's weights need not sum to one: its light pdf of a BSDF direction toward a
// triangle is 0). Sampled directions and Russian roulette are held fixed (detached sampling).
__device__ __forceinline__ void adjoint_vertex(const MisArgs &a, const Bsdf &b, int m, f3 adj, f3 termL, f3 wl,
                                               float misL, f3 termB, f3 wb, float misB, f3 cont, f3 wc, f3 wo,
                                               double *lds, double *grad) {
        double g[4] = {0.0, 0.0, 0.0, 0.0};
        bool any = false;
        const GzMaterial &mat = a.mats[m];
        if (b.b.type == GZ_MAT_DIFFUSE) {
                const f3 s = add(add(termL, termB), cont);
                const float R[3] = {b.b.R.x, b.b.R.y, b.b.R.z};
                const float S[3] = {s.x, s.y, s.z}, A[3] = {adj.x, adj.y, adj.z};
                for (int c = 0; c < 3; ++c)
                        if (R[c] != 0.0f) g[c] = (double)A[c] * (double)S[c] / (double)R[c];
                any = true;
        } else if (b.b.type == GZ_MAT_CONDUCTOR && mat.roughness[0] == mat.roughness[1] && mat.roughness[0] > 1e-3f) {
                const float al = b.b.d.ax;
                float rl = 0.0f, rb = 0.0f, rc = 0.0f;
                if (!cblack(termL)) {  // w_l = pl^2 / (pl^2 + pb^2): d ln w_l = -2 (1 - w_l) d ln pb
                        const DlogAlpha d = dlog_dalpha(al, wo, wl);
                        rl = d.f - 2.0f * (1.0f - misL) * d.pdf;
                }
                if (!cblack(termB)) {  // w_b = pb^2 / (pb^2 + pl^2): d ln w_b = 2 (1 - w_b) d ln pb
                        const DlogAlpha d = dlog_dalpha(al, wo, wb);
                        rb = d.f + 2.0f * (1.0f - misB) * d.pdf;
                }
                if (!cblack(cont)) rc = dlog_dalpha(al, wo, wc).f;
                const f3 s = add(add(mul(termL, rl), mul(termB, rb)), mul(cont, rc));
                g[3] = (double)adj.x * s.x + (double)adj.y * s.y + (double)adj.z * s.z;
                any = true;
        }
        if (!any) return;
        double *dst = lds ? lds : grad;
        for (int k = 0; k < 4; ++k)
                if (g[k] != 0.0) atomicAdd(&dst[4 * m + k], g[k]);
}

// One bounce of path i. kStats: also adds the bounce's traversal work (all its rays) to
// stats[i] and counts the bounce (interior, tris, bounces). kAdj: replays the bounce of the
// forward pass and adds the gradient of the path's radiance (lds: per-block accumulator).
// The light strategy's term of estimateDirect for an unoccluded light sample: f |cos| L weighted
// by the power heuristic against the BSDF's pdf of the light direction, over the light pdf
// (VolumePathIntegrator.swift sampleLightSource / MultipleImportanceSampler).
template <int kAll>
__device__ __forceinline__ f3 light_term(const MisArgs &a, const GzPath &p, const Bsdf &b, const Interaction &si,
                                         const LightSample &ls) {
        LayerRng r1 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallLightF);
        const f3 scatter = mul(bsdf_eval_world<kAll>(b, si.outgoing, ls.dir, r1), absdot(ls.dir, si.shadingNormal));
        const f3 Lest = cmul(scatter, ls.L);
        const float lpdf = ls.pdf;
        LayerRng r2 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallLightPdf);
        const float bpdfL = bsdf_pdf_world<kAll>(b, si.outgoing, ls.dir, r2);
        const float wl = power_heuristic(lpdf, bpdfL);
        return lpdf == 0.0f ? mk(0.0f, 0.0f, 0.0f) : divs(mul(Lest, wl), lpdf);
}

// Whether a path record takes part in a bounce (VolumePathIntegrator.swift:400-402 stops it past
// maxDepth).
__device__ __forceinline__ bool path_live(const MisArgs &a, const GzPath &p) {
        return (p.flags & GZ_PATH_ACTIVE) && p.bounce <= a.maxDepth;
}

// The bounce of path i (record p, already loaded and live) after its closest hit: `traced` and
// (prim, hu, hv) are the closest hit of p.ray. Returns whether the path goes on.
template <bool kStats, int kAll, bool kAdj>
__device__ __forceinline__ bool bounce_rest(const MisArgs &a, GzPath *__restrict__ paths, int64_t i, GzPath p,
                                            bool traced, int prim, float hu, float hv, Work &w,
                                            uint32_t *__restrict__ stats, const AdjArgs &adj, double *lds) {
        f3 T = mk(p.throughput[0], p.throughput[1], p.throughput[2]);
        f3 E = mk(p.estimate[0], p.estimate[1], p.estimate[2]);
        bool specular = (p.flags & GZ_PATH_SPECULAR) != 0;
        bool go = false;
        Interaction si;
        const bool hit = traced && surface_interaction(a, p.ray, prim, hu, hv, si);
        ZSobolState zs;
        zs.morton = (a.zs.log2spp >= 64 ? 0ull : zs_morton2((uint32_t)p.pixelX, (uint32_t)p.pixelY) << a.zs.log2spp) |
                    (uint64_t)(uint32_t)p.sampleIndex;
        zs.dim = p.dim;
        ZSobolState lzs;
        lzs.morton = 0;  // the light chooser's sampler never starts a pixel sample
        lzs.dim = p.lightDim;
        if (!hit) {
                if ((p.bounce == 0 || specular) && a.hasSky) E = cadd(E, cmul(T, mk(a.sky[0], a.sky[1], a.sky[2])));
        } else if (!cblack(T) && p.bounce < a.maxDepth) {
                if ((p.bounce == 0 || specular) && si.areaLight >= 0 && si.areaLight < a.nLights) {
                        const float4 Lb = a.lights[si.areaLight].L;
                        const f3 Le = dot(si.normal, si.outgoing) > 0.0f ? mk(Lb.x, Lb.y, Lb.z) : mk(0.0f, 0.0f, 0.0f);
                        E = cadd(E, cmul(T, Le));
                }
                Bsdf b;
                bsdf_make(a, si, b);
                if (p.bounce == 0) {
                        const f3 al = bx_albedo<kAll>(b.b);
                        p.albedo[0] = al.x;
                        p.albedo[1] = al.y;
                        p.albedo[2] = al.z;
                        p.firstNormal[0] = si.normal.x;
                        p.firstNormal[1] = si.normal.y;
                        p.firstNormal[2] = si.normal.z;
                }
                // sampleOneLight -> estimateDirect -> sampleMultipleImportance
                float choicePdf;
                const float ul = zs_get1d(a.zs, lzs, a.sobol);
                const int choice = choose_light(a, ul, choicePdf);
                float u0, u1;
                zs_get2d(a.zs, zs, a.sobol, u0, u1);
                LightSample ls;
                light_sample(a, choice, si.position, u0, u1, ls);
                f3 Lest = mk(0.0f, 0.0f, 0.0f), linc = mk(0.0f, 0.0f, 0.0f);
                float lpdf = 0.0f, wl;
                f3 lc = mk(0.0f, 0.0f, 0.0f);
                if (kAdj) {
                        // the reference's order: the shadow ray first (the gradient of a zero term
                        // can be non-zero, so the adjoint needs its visibility)
                        bool vis = false;
                        if (!cblack(ls.L) && !isinf(ls.pdf)) {
                                const int64_t ri = rec_index(adj, p.bounce, i);
                                const uint8_t o = ri >= 0 ? adj.recOcc[ri] : kRecNoOcc;
                                vis = o != kRecNoOcc ? o == 0 : !occluded<kStats>(a, si.position, ls.target, w);
                        }
                        if (vis) {
                                LayerRng r1 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallLightF);
                                const f3 scatter = mul(bsdf_eval_world<kAll>(b, si.outgoing, ls.dir, r1),
                                                       absdot(ls.dir, si.shadingNormal));
                                Lest = cmul(scatter, ls.L);
                                lpdf = ls.pdf;
                                linc = ls.dir;
                        }
                        LayerRng r2 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallLightPdf);
                        const float bpdfL = bsdf_pdf_world<kAll>(b, si.outgoing, linc, r2);
                        wl = power_heuristic(lpdf, bpdfL);
                        lc = lpdf == 0.0f ? mk(0.0f, 0.0f, 0.0f) : divs(mul(Lest, wl), lpdf);
                } else {
                        // the term as if unoccluded, then the shadow ray only where that term is not
                        // exactly zero: an occluded light gives the zero term as well
                        wl = 0.0f;
                        if (!cblack(ls.L) && !isinf(ls.pdf)) {
                                const f3 lcU = light_term<kAll>(a, p, b, si, ls);
                                if (!cblack(lcU) && !occluded<kStats>(a, si.position, ls.target, w)) lc = lcU;
                        }
                }
                float v0, v1, v2;
                zs_get3d(a.zs, zs, a.sobol, v0, v1, v2);
                f3 bf, bwi;
                float bpdf;
                LayerRng r3 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallMisSample);
                bsdf_sample_world<kAll>(b, si.outgoing, v0, v1, v2, bf, bwi, bpdf, r3);
                bf = mul(bf, absdot(bwi, si.shadingNormal));
                f3 bc = mk(0.0f, 0.0f, 0.0f);
                float wbMis = 0.0f;
                const bool weighted = choice >= a.nLights || (a.flags & GZ_MIS_FIX_LIGHT_PDF);
                if (weighted && !cblack(bf) && bpdf > 0.0f) {
                        const Ray_C r = spawn(si, bwi);
                        int bprim = -1;
                        float bt, bu, bv;
                        Interaction bsi;
                        f3 bL = mk(0.0f, 0.0f, 0.0f);
                        bool ok = false;
                        bool bhit;
                        const int64_t ri = kAdj ? rec_index(adj, p.bounce, i) : -1;
                        const GzHit rh = ri >= 0 ? adj.recMhits[ri] : GzHit{kRecNone, 0.0f, 0.0f, 0.0f};
                        if (rh.prim != kRecNone) {
                                bhit = rh.prim >= 0;
                                bprim = rh.prim;
                                bu = rh.u;
                                bv = rh.v;
                        } else {
                                bhit = trace_ray_until<kStats>(a, r, bsdf_ray_stop(a, r), bprim, bt, bu, bv, w);
                        }
                        if (bhit && !bsdf_hit_needs_interaction(a, bprim)) {
                                // a surface that emits nothing: no term
                        } else if (bhit && surface_interaction(a, r, bprim, bu, bv, bsi)) {
                                if (bsi.areaLight >= 0 && bsi.areaLight < a.nLights) {
                                        const float4 Lb = a.lights[bsi.areaLight].L;
                                        bL = dot(bsi.normal, bwi) > 0.0f ? mk(Lb.x, Lb.y, Lb.z) : mk(0.0f, 0.0f, 0.0f);
                                        ok = true;
                                }
                        } else if (a.hasSky) {
                                bL = mk(a.sky[0], a.sky[1], a.sky[2]);
                                ok = true;
                        }
                        if (ok) {
                                const f3 best = cmul(bf, bL);
                                const float lden = light_pdf(a, choice, si, bwi);
                                const float wb = power_heuristic(bpdf, lden);
                                wbMis = wb;
                                bc = divs(mul(best, wb), bpdf);
                        }
                }
                const f3 est = cadd(lc, bc);
                E = cadd(E, cmul(T, divs(est, choicePdf)));
                // continuation: bsdf.sampleWorldSpace(outgoing:uSample: get3D())
                float w0, w1, w2;
                zs_get3d(a.zs, zs, a.sobol, w0, w1, w2);
                f3 f, wi;
                float pdf;
                LayerRng r4 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallContinue);
                bsdf_sample_world<kAll>(b, si.outgoing, w0, w1, w2, f, wi, pdf, r4);
                if (kAdj) {
                        const GzPathAdjoint &pa = adj.paths[i];
                        const f3 Ltot = mk(pa.radiance[0], pa.radiance[1], pa.radiance[2]);
                        const f3 av = mk(pa.adjoint[0], pa.adjoint[1], pa.adjoint[2]);
                        const int mi = si.material >= 0 && si.material < a.nMats ? si.material : 0;
                        adjoint_vertex(a, b, mi, av, cmul(T, divs(lc, choicePdf)), to_local(b, linc), wl,
                                       cmul(T, divs(bc, choicePdf)), to_local(b, bwi), wbMis, sub(Ltot, E),
                                       to_local(b, wi), to_local(b, si.outgoing), lds, adj.grad);
                }
                if (pdf != 0.0f && !isnan(pdf)) {
                        specular = bx_specular<kAll>(b.b);
                        T = cmul(T, divs(mul(f, absdot(wi, si.normal)), pdf));
                        p.ray = spawn(si, wi);
                        go = true;
                        if (cmax(T) < 1.0f && p.bounce > 1) {
                                const float m = cmax(T);
                                const float q = 1.0f - m > 0.0f ? 1.0f - m : 0.0f;
                                const float u = rr_draw(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce);
                                if (u < q)
                                        go = false;
                                else
                                        T = divs(T, 1.0f - q);
                        }
                }
        }
        p.throughput[0] = T.x;
        p.throughput[1] = T.y;
        p.throughput[2] = T.z;
        p.estimate[0] = E.x;
        p.estimate[1] = E.y;
        p.estimate[2] = E.z;
        p.dim = zs.dim;
        p.lightDim = lzs.dim;
        p.flags = (go ? GZ_PATH_ACTIVE : 0) | (specular ? GZ_PATH_SPECULAR : 0);
        if (go) p.bounce += 1;
        paths[i] = p;
        if (kStats) {
                stats[3 * i] += w.interior;
                stats[3 * i + 1] += w.tris;
                stats[3 * i + 2] += 1;
        }
        return go;
}

// ---- The bounce in phases (the full wavefront launch) ------------------------------------------
//
// bounce_rest with its shadow ray and BSDF ray traced by other kernels: shade_phase does everything
// up to those two rays and the continuation, leaving the light-strategy term (valid if the shadow
// ray is unoccluded), the BSDF-strategy factors and the throughput before the continuation in a
// Pending record; finish_phase adds the two terms once the rays are traced. The arithmetic is
// bounce_rest's, operation for operation, so the records equal the one-kernel bounce's.
enum : int { kPendWork = 1, kPendShadow = 2, kPendMis = 4 };

struct Pending {
        float T[3];   // throughput before the continuation
        float lc[3];  // light-strategy term if the shadow ray is unoccluded
        float bf[3];  // BSDF-strategy f |cos| (times the radiance the BSDF ray finds)
        float wb;     // BSDF-strategy MIS weight
        float bpdf;
        float choicePdf;
        int flags;
        int pad[3];
};
static_assert(sizeof(Pending) == 64, "Pending");

template <bool kStats, int kAll>
__device__ __forceinline__ bool shade_phase(const MisArgs &a, GzPath *__restrict__ paths, int64_t i, GzPath p,
                                            bool traced, int prim, float hu, float hv, Pending &pd, Ray_C &sray,
                                            Ray_C &mray, uint32_t *__restrict__ stats) {
        f3 T = mk(p.throughput[0], p.throughput[1], p.throughput[2]);
        f3 E = mk(p.estimate[0], p.estimate[1], p.estimate[2]);
        bool specular = (p.flags & GZ_PATH_SPECULAR) != 0;
        bool go = false;
        pd.flags = 0;
        Interaction si;
        const bool hit = traced && surface_interaction(a, p.ray, prim, hu, hv, si);
        ZSobolState zs;
        zs.morton = (a.zs.log2spp >= 64 ? 0ull : zs_morton2((uint32_t)p.pixelX, (uint32_t)p.pixelY) << a.zs.log2spp) |
                    (uint64_t)(uint32_t)p.sampleIndex;
        zs.dim = p.dim;
        ZSobolState lzs;
        lzs.morton = 0;
        lzs.dim = p.lightDim;
        if (!hit) {
                if ((p.bounce == 0 || specular) && a.hasSky) E = cadd(E, cmul(T, mk(a.sky[0], a.sky[1], a.sky[2])));
        } else if (!cblack(T) && p.bounce < a.maxDepth) {
                if ((p.bounce == 0 || specular) && si.areaLight >= 0 && si.areaLight < a.nLights) {
                        const float4 Lb = a.lights[si.areaLight].L;
                        const f3 Le = dot(si.normal, si.outgoing) > 0.0f ? mk(Lb.x, Lb.y, Lb.z) : mk(0.0f, 0.0f, 0.0f);
                        E = cadd(E, cmul(T, Le));
                }
                Bsdf b;
                bsdf_make(a, si, b);
                if (p.bounce == 0) {
                        const f3 al = bx_albedo<kAll>(b.b);
                        p.albedo[0] = al.x;
                        p.albedo[1] = al.y;
                        p.albedo[2] = al.z;
                        p.firstNormal[0] = si.normal.x;
                        p.firstNormal[1] = si.normal.y;
                        p.firstNormal[2] = si.normal.z;
                }
                float choicePdf;
                const float ul = zs_get1d(a.zs, lzs, a.sobol);
                const int choice = choose_light(a, ul, choicePdf);
                float u0, u1;
                zs_get2d(a.zs, zs, a.sobol, u0, u1);
                LightSample ls;
                light_sample(a, choice, si.position, u0, u1, ls);
                pd.flags = kPendWork;
                f3 lcU = mk(0.0f, 0.0f, 0.0f);
                if (!cblack(ls.L) && !isinf(ls.pdf)) {
                        // bounce_rest's unoccluded term; the shadow ray, where it is not zero, decides
                        lcU = light_term<kAll>(a, p, b, si, ls);
                        if (!cblack(lcU)) {
                                // occluded(): origin from + 1e-4 (target - from), direction target - origin
                                const f3 o = add(si.position, mul(sub(ls.target, si.position), 0.0001f));
                                const f3 d = sub(ls.target, o);
                                sray = Ray_C{o.x, o.y, o.z, d.x, d.y, d.z};
                                pd.flags |= kPendShadow;
                        }
                }
                float v0, v1, v2;
                zs_get3d(a.zs, zs, a.sobol, v0, v1, v2);
                f3 bf, bwi;
                float bpdf;
                LayerRng r3 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallMisSample);
                bsdf_sample_world<kAll>(b, si.outgoing, v0, v1, v2, bf, bwi, bpdf, r3);
                bf = mul(bf, absdot(bwi, si.shadingNormal));
                const bool weighted = choice >= a.nLights || (a.flags & GZ_MIS_FIX_LIGHT_PDF);
                float wb = 0.0f;
                if (weighted && !cblack(bf) && bpdf > 0.0f) {
                        mray = spawn(si, bwi);
                        const float lden = light_pdf(a, choice, si, bwi);
                        wb = power_heuristic(bpdf, lden);
                        pd.flags |= kPendMis;
                }
                pd.T[0] = T.x;
                pd.T[1] = T.y;
                pd.T[2] = T.z;
                pd.lc[0] = lcU.x;
                pd.lc[1] = lcU.y;
                pd.lc[2] = lcU.z;
                pd.bf[0] = bf.x;
                pd.bf[1] = bf.y;
                pd.bf[2] = bf.z;
                pd.wb = wb;
                pd.bpdf = bpdf;
                pd.choicePdf = choicePdf;
                float w0, w1, w2;
                zs_get3d(a.zs, zs, a.sobol, w0, w1, w2);
                f3 f, wi;
                float pdf;
                LayerRng r4 = layer_rng(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce, kCallContinue);
                bsdf_sample_world<kAll>(b, si.outgoing, w0, w1, w2, f, wi, pdf, r4);
                if (pdf != 0.0f && !isnan(pdf)) {
                        specular = bx_specular<kAll>(b.b);
                        T = cmul(T, divs(mul(f, absdot(wi, si.normal)), pdf));
                        p.ray = spawn(si, wi);
                        go = true;
                        if (cmax(T) < 1.0f && p.bounce > 1) {
                                const float m = cmax(T);
                                const float q = 1.0f - m > 0.0f ? 1.0f - m : 0.0f;
                                const float u = rr_draw(a.rrSeed, p.pixelX, p.pixelY, p.sampleIndex, p.bounce);
                                if (u < q)
                                        go = false;
                                else
                                        T = divs(T, 1.0f - q);
                        }
                }
        }
        p.throughput[0] = T.x;
        p.throughput[1] = T.y;
        p.throughput[2] = T.z;
        p.estimate[0] = E.x;
        p.estimate[1] = E.y;
        p.estimate[2] = E.z;
        p.dim = zs.dim;
        p.lightDim = lzs.dim;
        p.flags = (go ? GZ_PATH_ACTIVE : 0) | (specular ? GZ_PATH_SPECULAR : 0);
        if (go) p.bounce += 1;
        paths[i] = p;
        if (kStats) stats[3 * i + 2] += 1;
        return go;
}

// Adds the bounce's light- and BSDF-strategy terms once the shadow ray (occ) and the BSDF ray
// (mray, its closest hit mh) are traced.
__device__ __forceinline__ void finish_phase(const MisArgs &a, GzPath *__restrict__ paths, int64_t i,
                                             const Pending &pd, bool occ, const Ray_C &mray, const GzHit &mh) {
        if (!(pd.flags & kPendWork)) return;
        f3 lc = mk(0.0f, 0.0f, 0.0f);
        if ((pd.flags & kPendShadow) && !occ) lc = mk(pd.lc[0], pd.lc[1], pd.lc[2]);
        f3 bc = mk(0.0f, 0.0f, 0.0f);
        if (pd.flags & kPendMis) {
                const f3 bwi = mk(mray.dirX, mray.dirY, mray.dirZ);
                Interaction bsi;
                f3 bL = mk(0.0f, 0.0f, 0.0f);
                bool ok = false;
                if (mh.prim >= 0 && !bsdf_hit_needs_interaction(a, mh.prim)) {
                        // a surface that emits nothing: no term
                } else if (mh.prim >= 0 && surface_interaction(a, mray, mh.prim, mh.u, mh.v, bsi)) {
                        if (bsi.areaLight >= 0 && bsi.areaLight < a.nLights) {
                                const float4 Lb = a.lights[bsi.areaLight].L;
                                bL = dot(bsi.normal, bwi) > 0.0f ? mk(Lb.x, Lb.y, Lb.z) : mk(0.0f, 0.0f, 0.0f);
                                ok = true;
                        }
                } else if (a.hasSky) {
                        bL = mk(a.sky[0], a.sky[1], a.sky[2]);
                        ok = true;
                }
                if (ok) {
                        const f3 best = cmul(mk(pd.bf[0], pd.bf[1], pd.bf[2]), bL);
                        bc = divs(mul(best, pd.wb), pd.bpdf);
                }
        }
        const f3 est = cadd(lc, bc);
        const GzPath &p = paths[i];
        f3 E = mk(p.estimate[0], p.estimate[1], p.estimate[2]);
        E = cadd(E, cmul(mk(pd.T[0], pd.T[1], pd.T[2]), divs(est, pd.choicePdf)));
        paths[i].estimate[0] = E.x;
        paths[i].estimate[1] = E.y;
        paths[i].estimate[2] = E.z;
}

// One bounce of path i; returns whether the path goes on to another bounce.
template <bool kStats, int kAll, bool kAdj>
__device__ __forceinline__ bool bounce_one(const MisArgs &a, GzPath *__restrict__ paths, int64_t i,
                                           uint32_t *__restrict__ stats, const AdjArgs &adj, double *lds) {
        const GzPath p = paths[i];
        if (!(p.flags & GZ_PATH_ACTIVE)) return false;
        if (p.bounce > a.maxDepth) {
                paths[i].flags = p.flags & ~GZ_PATH_ACTIVE;
                return false;
        }
        Work w;
        int prim = -1;
        float th, hu = 0.0f, hv = 0.0f;
        bool traced;
        const int64_t ri = kAdj ? rec_index(adj, p.bounce, i) : -1;
        const GzHit rh = ri >= 0 ? adj.recHits[ri] : GzHit{kRecNone, 0.0f, 0.0f, 0.0f};
        if (rh.prim != kRecNone) {
                traced = rh.prim >= 0;
                prim = rh.prim;
                hu = rh.u;
                hv = rh.v;
        } else {
                traced = trace_ray<false, kStats>(a, p.ray, INFINITY, prim, th, hu, hv, w);
        }
        return bounce_rest<kStats, kAll, kAdj>(a, paths, i, p, traced, prim, hu, hv, w, stats, adj, lds);
}

// The BSDF set a hit shades with (area-light surfaces shade with materials[0], as in
// surface_interaction and bsdf_make): kBsdfDiffuse for a diffuse material or a miss, kBsdfAll for
// the layered BSDF, kBsdfNoLayer for the other BSDFs.
__device__ __forceinline__ int hit_bsdf_set(const MisArgs &a, int prim) {
        if (prim < 0 || prim >= a.primCount) return kBsdfDiffuse;
        return a.primClass[prim] & kClsSetMask;
}
// Whether the hit needs a kernel with more than the diffuse BSDF.
__device__ __forceinline__ bool hit_needs_all_bsdfs(const MisArgs &a, int prim) {
        return hit_bsdf_set(a, prim) != kBsdfDiffuse;
}

// Waves per SIMD a bounce kernel compiled with BSDF set kAll asks of the compiler.
constexpr int mis_min_waves(int kAll) { return kAll == kBsdfDiffuse ? 4 : 3; }

// ---- Queue appends ----------------------------------------------------------------------------

// Entries of up to four runs of 64 queue entries held per lane until their appends are written
// (k is wave-uniform, so the selects stay in registers).
struct Keep4 {
        int e0 = -1, e1 = -1, e2 = -1, e3 = -1;
        __device__ __forceinline__ void set(int k, int v) {
                if (k == 0)
                        e0 = v;
                else if (k == 1)
                        e1 = v;
                else if (k == 2)
                        e2 = v;
                else
                        e3 = v;
        }
        __device__ __forceinline__ int get(int k) const { return k == 0 ? e0 : k == 1 ? e1 : k == 2 ? e2 : e3; }
};

// Writes the appends of `runs` held runs: bits[q] bit k = the lane's entry of run k goes to queue q.
// One atomic per queue and flush: a counter that every wave of the chip hits serialises them.
template <int kQ>
__device__ __forceinline__ void flush_appends(const Keep4 &keep, const unsigned (&bits)[kQ], int runs,
                                              int *const (&qout)[kQ], unsigned *const (&outN)[kQ], int lane,
                                              uint64_t below) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
                unsigned total = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                        if (k < runs) total += (unsigned)__popcll(__ballot((bits[q] >> k) & 1u));
                if (total == 0) continue;
                unsigned base = 0;
                if (lane == 0) base = atomicAdd(outN[q], total);
                base = __shfl(base, 0);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                        if (k >= runs) break;
                        const bool want = (bits[q] >> k) & 1u;
                        const uint64_t m = __ballot(want);
                        if (want) qout[q][base + (unsigned)__popcll(m & below)] = keep.get(k);
                        base += (unsigned)__popcll(m);
                }
        }
}

// Host side (gz_mis.hip): the device arguments of a scene's MIS tables and parameters (false
// with the error set when the handle has none or the parameters are out of range), and whether
// every material is diffuse (the kernels built without the other BSDFs then suffice).
bool mis_args(Scene *s, const GzMisParams *params, int *err, MisArgs &a);
bool mis_all_diffuse(const Scene *s);
// Whether a scene's default launch is the full wavefront: every scene but the small all-diffuse
// ones (at most kDefaultMegakernelPrims primitives), which the queued megakernel serves better.
constexpr int64_t kDefaultMegakernelPrims = 65536;
bool default_wavefront(const Scene *s);
// The BSDF set (kBsdfDiffuse, kBsdfAll or kBsdfNoLayer) a scene's megakernel-form kernels use.
int mis_bsdf_set(const Scene *s);
// One of three kernel instantiations by BSDF set.
template <typename K>
inline K pick_set(int set, K diffuse, K all, K noLayer) {
        return set == kBsdfDiffuse ? diffuse : set == kBsdfAll ? all : noLayer;
}

}  // namespace gz
