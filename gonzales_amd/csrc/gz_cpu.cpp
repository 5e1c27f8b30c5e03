// The CPU half of the reference's C boundary (mojo_kernel.h:60-64, 86-91): single-ray and batch
// closest-hit traversal over the reference's own BVH2 / PrimId / TriangleMesh layout, and the
// mode-A diffuse shade step, as plain host C++. A Swift build that links this library instead of
// libmojo keeps its CPU path (BoundingHierarchy.swift:308, :334, :374, :413).
//
// Semantics follow kernel.mojo:
//   traversal  traverse_bvh2_core :223-364 (slab test :171-206, Moller-Trumbore :122-156);
//   shade      shade_core :676-773 with PCG32 :99-118.
// Every product and sum rounds on its own (this file is compiled with -ffp-contract=off), in the
// reference's order, so the results equal the device kernels' bit for bit where those are
// bit-exact (traversal, PCG state, throughput); directions go through the C library's cosf/sinf.
//
// Like the reference, the batch functions are sequential loops with no global state: callers run
// them concurrently (TileRenderer.swift:77-107 runs one per tile task). Unlike the reference, the
// 64-entry traversal stack is checked: a ray that would overflow it stops with the hit found so far
// and the calling thread's gz_last_error() says so.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "gonzales_amd.h"

namespace gz {
void set_error(const char *fmt, ...);
void clear_error();
}

namespace {

struct V3 {
        float x, y, z;
};

inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
// a.yzx * b.zxy - a.zxy * b.yzx (kernel.mojo:84-90)
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// (a*b)[0] + (a*b)[1] + (a*b)[2], left to right (kernel.mojo:92-95)
inline float dot(V3 a, V3 b) {
        const float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
        return (px + py) + pz;
}

// The reference's PrimId decode (kernel.mojo:266-277): false for the primitives it skips.
inline bool decode(const PrimId_C &p, int64_t &mesh, int64_t &base) {
        if (p.type == 0) {
                mesh = p.id1;
                base = p.id2;
                return true;
        }
        if (p.type == 1 || p.type == 2 || p.type == 3) {
                if (p.id2 == -1) return false;
                mesh = p.id2 >> 32;
                base = (p.id2 & 0xFFFFFFFFLL) * 3;
                return true;
        }
        return false;
}

inline V3 point(const TriangleMesh_C &m, int64_t k) {
        const int64_t v = m.vertexIndices[k];
        return V3{m.points[v * 4], m.points[v * 4 + 1], m.points[v * 4 + 2]};
}

// Moller-Trumbore with the reference's rejections (kernel.mojo:122-156).
inline bool triangle(V3 o, V3 d, V3 p0, V3 p1, V3 p2, float tMax, float &t, float &u, float &v) {
        const V3 e1 = sub(p1, p0), e2 = sub(p2, p0);
        const V3 pvec = cross(d, e2);
        const float det = dot(e1, pvec);
        if (det > -0.0000001f && det < 0.0000001f) return false;
        const float invDet = 1.0f / det;
        const V3 tvec = sub(o, p0);
        u = dot(tvec, pvec) * invDet;
        if (u < 0.0f || u > 1.0f) return false;
        const V3 qvec = cross(tvec, e1);
        v = dot(d, qvec) * invDet;
        if (v < 0.0f || u + v > 1.0f) return false;
        t = dot(e2, qvec) * invDet;
        return !(t <= 0.0f || t > tMax);
}

struct RayPre {
        float rx, ry, rz, ox, oy, oz;
        bool nx, ny, nz;
};

// Slab test with select-form maxima and minima (`a if a > b else b`: NaN takes the second
// operand) and the x1.0000003 slack (kernel.mojo:171-206).
inline bool box(const BVH2Node &n, const RayPre &r, float tMax, float &tNearOut) {
        const float nearX = r.nx ? n.boundsMinX : n.boundsMaxX, farX = r.nx ? n.boundsMaxX : n.boundsMinX;
        const float nearY = r.ny ? n.boundsMinY : n.boundsMaxY, farY = r.ny ? n.boundsMaxY : n.boundsMinY;
        const float nearZ = r.nz ? n.boundsMinZ : n.boundsMaxZ, farZ = r.nz ? n.boundsMaxZ : n.boundsMinZ;
        const float tnx = nearX * r.rx - r.ox, tny = nearY * r.ry - r.oy, tnz = nearZ * r.rz - r.oz;
        const float tfx = farX * r.rx - r.ox, tfy = farY * r.ry - r.oy, tfz = farZ * r.rz - r.oz;
        float tNear = tnx > tny ? tnx : tny;
        tNear = tnz > tNear ? tnz : tNear;
        tNear = 0.0f > tNear ? 0.0f : tNear;
        float tFar = tfx < tfy ? tfx : tfy;
        tFar = tfz < tFar ? tfz : tFar;
        tFar = tMax < tFar ? tMax : tFar;
        tFar = tFar * 1.0000003f;
        tNearOut = tNear;
        return tNear <= tFar;
}

// traverse_bvh2_core (kernel.mojo:223-364). Returns false on a stack overflow.
bool traverse(const SceneDescriptor2_C &sc, const Ray_C &ray, float tMax, Intersection_C &out) {
        RayPre r;
        r.rx = 1.0f / ray.dirX;
        r.ry = 1.0f / ray.dirY;
        r.rz = 1.0f / ray.dirZ;
        r.ox = ray.orgX * r.rx;
        r.oy = ray.orgY * r.ry;
        r.oz = ray.orgZ * r.rz;
        r.nx = r.rx >= 0.0f;
        r.ny = r.ry >= 0.0f;
        r.nz = r.rz >= 0.0f;
        const V3 o{ray.orgX, ray.orgY, ray.orgZ}, d{ray.dirX, ray.dirY, ray.dirZ};
        int64_t hitIndex = -1;
        float tHit = tMax, bu = 0.0f, bv = 0.0f;
        int64_t stack[GZ_STACK_CAPACITY];
        int top = 0;
        int64_t current = 0;
        bool ok = true;
        const BVH2Node *nodes = sc.bvh2Nodes;
        while (nodes) {
                const BVH2Node &node = nodes[current];
                bool pop = false;
                if (node.count > 0) {
                        for (int64_t j = 0; j < node.count; ++j) {
                                const PrimId_C &p = sc.primIds[node.offset + j];
                                int64_t mesh, base;
                                if (!decode(p, mesh, base)) continue;
                                const TriangleMesh_C &m = sc.meshes[mesh];
                                float t, u, v;
                                if (triangle(o, d, point(m, base), point(m, base + 1), point(m, base + 2), tHit, t, u,
                                             v)) {
                                        tHit = t;
                                        bu = u;
                                        bv = v;
                                        hitIndex = node.offset + j;
                                }
                        }
                        pop = true;
                } else {
                        const int64_t li = current + 1, ri = node.offset;
                        float tl, tr;
                        const bool hl = box(nodes[li], r, tHit, tl);
                        const bool hr = box(nodes[ri], r, tHit, tr);
                        if (hl && hr) {
                                if (top == GZ_STACK_CAPACITY) {
                                        ok = false;
                                        break;
                                }
                                const bool leftFirst = tl <= tr;
                                current = leftFirst ? li : ri;
                                stack[top++] = leftFirst ? ri : li;
                        } else if (hl) {
                                current = li;
                        } else if (hr) {
                                current = ri;
                        } else {
                                pop = true;
                        }
                }
                if (pop) {
                        if (top == 0) break;
                        current = stack[--top];
                }
        }
        std::memset(&out, 0, sizeof out);
        if (hitIndex != -1) {
                out.primId = sc.primIds[hitIndex];
                out.tHit = tHit;
                out.u = bu;
                out.v = bv;
                out.hit = 1;
        } else {
                out.primId.id1 = -1;
                out.primId.id2 = -1;
                out.tHit = tMax;
        }
        return ok;
}

// PCG32 (kernel.mojo:99-118): pcg32_srandom seeding, XSH-RR output.
struct Pcg32 {
        uint64_t state, inc;
        Pcg32(uint64_t initstate, uint64_t initseq) : state(0), inc((initseq << 1) | 1u) {
                (void)next();
                state += initstate;
                (void)next();
        }
        uint32_t next() {
                const uint64_t old = state;
                state = old * 6364136223846793005ull + inc;
                const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
                const uint32_t rot = (uint32_t)(old >> 59);
                return (xs >> rot) | (xs << ((0u - rot) & 31u));
        }
        float next_float() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// shade_core (kernel.mojo:676-773).
void shade(PathState_C &p, const Intersection_C &in, const TriangleMesh_C *meshes, const Material_C *materials) {
        if (p.active == 0) return;
        if (in.hit == 0) {
                p.active = 0;
                return;
        }
        const Material_C &mat = materials[in.primId.materialIndex];
        if (mat.type == 2) {
                for (int c = 0; c < 3; ++c) p.estimate[c] += p.throughput[c] * mat.emission[c];
                p.active = 0;
                return;
        }
        if (mat.type != 1) {
                p.active = 0;
                return;
        }
        int64_t mesh, base;
        const PrimId_C &pid = in.primId;
        if (pid.type == 0) {
                mesh = pid.id1;
                base = pid.id2;
        } else if (pid.type == 1 || pid.type == 2 || pid.type == 3) {
                mesh = pid.id2 >> 32;
                base = (pid.id2 & 0xFFFFFFFFLL) * 3;
        } else {
                p.active = 0;
                return;
        }
        const TriangleMesh_C &m = meshes[mesh];
        const V3 p0 = point(m, base), p1 = point(m, base + 1), p2 = point(m, base + 2);
        V3 n = cross(sub(p1, p0), sub(p2, p0));
        const float nlen = dot(n, n);
        if (nlen > 0.0f) n = scale(n, 1.0f / std::sqrt(nlen));
        const V3 rd{p.ray.dirX, p.ray.dirY, p.ray.dirZ};
        if (dot(n, rd) > 0.0f) n = scale(n, -1.0f);
        Pcg32 pcg(p.pcgState, p.pcgInc);
        const float u1 = pcg.next_float();
        const float u2 = pcg.next_float();
        p.pcgState = pcg.state;
        const float r = std::sqrt(u1);
        const float theta = 2.0f * 3.14159265359f * u2;
        const float x = r * std::cos(theta);
        const float y = r * std::sin(theta);
        const float z2 = 1.0f - u1;
        const float z = std::sqrt(z2 > 0.0f ? z2 : 0.0f);
        // Duff et al.'s basis about n (kernel.mojo:752-756)
        const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
        const float a = -1.0f / (sign + n.z);
        const float b = n.x * n.y * a;
        const V3 t{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
        const V3 bt{b, sign + n.y * n.y * a, -n.y};
        V3 dir = add(add(scale(t, x), scale(bt, y)), scale(n, z));
        const float dlen = dot(dir, dir);
        if (dlen > 0.0f) dir = scale(dir, 1.0f / std::sqrt(dlen));
        const V3 org = add(add(V3{p.ray.orgX, p.ray.orgY, p.ray.orgZ}, scale(rd, in.tHit)), scale(n, 0.0001f));
        p.ray = Ray_C{org.x, org.y, org.z, dir.x, dir.y, dir.z};
        for (int c = 0; c < 3; ++c) p.throughput[c] *= mat.albedo[c];
}

}  // namespace

extern "C" void gz_traverse_bvh2(const struct SceneDescriptor2_C *scene, const struct Ray_C *ray, float tMax,
                                 struct Intersection_C *result) {
        gz::clear_error();
        if (!scene || !ray || !result) return;
        if (!traverse(*scene, *ray, tMax, *result))
                gz::set_error("gz_traverse_bvh2: traversal stack overflow (more than %d entries)", GZ_STACK_CAPACITY);
}

extern "C" void gz_cpu_traverse_batch(const struct SceneDescriptor2_C *scene, const struct Ray_C *rays,
                                      const float *tMaxValues, int64_t count, struct Intersection_C *results) {
        gz::clear_error();
        if (!scene || count <= 0 || !rays || !tMaxValues || !results) return;
        int64_t overflow = 0;
        for (int64_t i = 0; i < count; ++i)
                if (!traverse(*scene, rays[i], tMaxValues[i], results[i])) ++overflow;
        if (overflow)
                gz::set_error("gz_cpu_traverse_batch: %lld rays overflowed the %d-entry traversal stack",
                              (long long)overflow, GZ_STACK_CAPACITY);
}

extern "C" void gz_cpu_shade_batch(struct PathState_C *paths, int64_t count, const struct Intersection_C *intersections,
                                   const struct TriangleMesh_C *meshes, const struct Material_C *materials) {
        gz::clear_error();
        if (count <= 0 || !paths || !intersections || !meshes || !materials) return;
        for (int64_t i = 0; i < count; ++i) shade(paths[i], intersections[i], meshes, materials);
}
