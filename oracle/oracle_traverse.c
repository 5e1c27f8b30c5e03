/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's BVH2 traversal, used by
 * tests/, the smoke check and bench.py's CPU-baseline leg as the checker; never shipped and
 * never on the product path.
 *
 * Restates Sources/mojoKernel/kernel.mojo (the reference's hot path):
 *   cross / dot ............ kernel.mojo:84-95
 *   intersect_triangle ..... kernel.mojo:122-156   (Moller-Trumbore, |det| < 1e-7 reject)
 *   intersect_aabb ......... kernel.mojo:171-206   (select-form min/max, x 1.0000003 slack)
 *   traverse_bvh2_core ..... kernel.mojo:223-364   (near-first, 64-entry stack, root box untested)
 *   mojo_traverse_bvh2 ..... kernel.mojo:369-373
 *   mojo_cpu_traverse_batch  kernel.mojo:378-390   (here: chunked over worker threads, the
 *                            reference's tile-task concurrency, TileRenderer.swift:77)
 * Compiled with -ffp-contract=off so every product and sum rounds separately, as written.
 *
 * Pinning (tests/test_oracle_traversal.py): brute force over all primitives (layout-independent
 * closest t), closed-form rays against axis-aligned triangles and boxes, M-T invariants, and
 * "any-hit == closest hit exists" equivalence.
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

/* Byte layouts of mojo_kernel.h:4-58, declared independently of the product header. */
typedef struct { int64_t id1, id2, materialIndex; int8_t type; int8_t pad[7]; } o_prim_id;
typedef struct { const float *points; const int64_t *faceIndices; const int64_t *vertexIndices; } o_mesh;
typedef struct { float ox, oy, oz, dx, dy, dz; } o_ray;
typedef struct { o_prim_id prim; float tHit, u, v; int8_t hit; int8_t pad[3]; } o_isect;
typedef struct { float minx, miny, minz, maxx, maxy, maxz; int32_t offset, count; } o_node;
typedef struct { const o_node *nodes; const o_prim_id *prims; const o_mesh *meshes; int64_t meshCount; } o_scene;
typedef struct { uint32_t interior, tris; } o_stats;

_Static_assert(sizeof(o_prim_id) == 32, "PrimId_C");
_Static_assert(sizeof(o_isect) == 48, "Intersection_C");
_Static_assert(offsetof(o_isect, tHit) == 32, "Intersection_C.tHit");
_Static_assert(sizeof(o_node) == 32, "BVH2Node");
_Static_assert(sizeof(o_ray) == 24, "Ray_C");
_Static_assert(sizeof(o_scene) == 32, "SceneDescriptor2_C");

typedef struct { float x, y, z; } v3;

/* kernel.mojo:84-90: a.yzx * b.zxy - a.zxy * b.yzx */
static inline v3 o_cross(v3 a, v3 b) {
        v3 r;
        r.x = a.y * b.z - a.z * b.y;
        r.y = a.z * b.x - a.x * b.z;
        r.z = a.x * b.y - a.y * b.x;
        return r;
}

/* kernel.mojo:92-95: products summed left to right */
static inline float o_dot(v3 a, v3 b) {
        float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
        return (px + py) + pz;
}

static inline v3 o_sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }

/* kernel.mojo:122-156. Returns 1 on hit and writes t, u, v. */
int orc_intersect_triangle(const float *org, const float *dir, const float *p0, const float *p1,
                           const float *p2, float tMax, float *tOut, float *uOut, float *vOut) {
        v3 o = {org[0], org[1], org[2]}, d = {dir[0], dir[1], dir[2]};
        v3 a = {p0[0], p0[1], p0[2]}, b = {p1[0], p1[1], p1[2]}, c = {p2[0], p2[1], p2[2]};
        v3 e1 = o_sub(b, a);
        v3 e2 = o_sub(c, a);
        v3 pvec = o_cross(d, e2);
        float det = o_dot(e1, pvec);
        if (det > -0.0000001f && det < 0.0000001f) return 0;
        float invDet = 1.0f / det;
        v3 tvec = o_sub(o, a);
        float u = o_dot(tvec, pvec) * invDet;
        if (u < 0.0f || u > 1.0f) return 0;
        v3 qvec = o_cross(tvec, e1);
        float v = o_dot(d, qvec) * invDet;
        if (v < 0.0f || u + v > 1.0f) return 0;
        float t = o_dot(e2, qvec) * invDet;
        if (t <= 0.0f || t > tMax) return 0;
        *tOut = t; *uOut = u; *vOut = v;
        return 1;
}

/* kernel.mojo:171-206. Returns hit flag, writes tNear. */
static inline int o_aabb(const o_node *n, float rdx, float rdy, float rdz, float ordx, float ordy,
                         float ordz, int nxm, int nym, int nzm, float tMax, float *tNearOut) {
        float nearX = nxm ? n->minx : n->maxx, farX = nxm ? n->maxx : n->minx;
        float nearY = nym ? n->miny : n->maxy, farY = nym ? n->maxy : n->miny;
        float nearZ = nzm ? n->minz : n->maxz, farZ = nzm ? n->maxz : n->minz;
        float tNearX = nearX * rdx - ordx, tNearY = nearY * rdy - ordy, tNearZ = nearZ * rdz - ordz;
        float tFarX = farX * rdx - ordx, tFarY = farY * rdy - ordy, tFarZ = farZ * rdz - ordz;
        float tNear = tNearX > tNearY ? tNearX : tNearY;
        tNear = tNearZ > tNear ? tNearZ : tNear;
        tNear = 0.0f > tNear ? 0.0f : tNear;
        float tFar = tFarX < tFarY ? tFarX : tFarY;
        tFar = tFarZ < tFar ? tFarZ : tFar;
        tFar = tMax < tFar ? tMax : tFar;
        tFar = tFar * 1.0000003f;
        *tNearOut = tNear;
        return tNear <= tFar;
}

/* Exported for the box known-answer tests. bounds: minx, miny, minz, maxx, maxy, maxz. */
int orc_intersect_aabb(const float *bounds, const float *org, const float *dir, float tMax,
                       float *tNearOut) {
        o_node n;
        memset(&n, 0, sizeof n);
        n.minx = bounds[0]; n.miny = bounds[1]; n.minz = bounds[2];
        n.maxx = bounds[3]; n.maxy = bounds[4]; n.maxz = bounds[5];
        float rdx = 1.0f / dir[0], rdy = 1.0f / dir[1], rdz = 1.0f / dir[2];
        return o_aabb(&n, rdx, rdy, rdz, org[0] * rdx, org[1] * rdy, org[2] * rdz, rdx >= 0.0f,
                      rdy >= 0.0f, rdz >= 0.0f, tMax, tNearOut);
}

/* kernel.mojo:266-300: PrimId decode and vertex gather. Returns 0 when the reference skips the
 * primitive (`continue`). */
static inline int o_gather(const o_scene *s, const o_prim_id *p, v3 *a, v3 *b, v3 *c) {
        int64_t mesh, base;
        if (p->type == 0) {
                mesh = p->id1;
                base = p->id2;
        } else if (p->type == 1 || p->type == 2 || p->type == 3) {
                if (p->id2 == -1) return 0;
                mesh = p->id2 >> 32;
                base = (p->id2 & 0xFFFFFFFFLL) * 3;
        } else {
                return 0;
        }
        const o_mesh *m = &s->meshes[mesh];
        int64_t i0 = m->vertexIndices[base], i1 = m->vertexIndices[base + 1], i2 = m->vertexIndices[base + 2];
        a->x = m->points[i0 * 4]; a->y = m->points[i0 * 4 + 1]; a->z = m->points[i0 * 4 + 2];
        b->x = m->points[i1 * 4]; b->y = m->points[i1 * 4 + 1]; b->z = m->points[i1 * 4 + 2];
        c->x = m->points[i2 * 4]; c->y = m->points[i2 * 4 + 1]; c->z = m->points[i2 * 4 + 2];
        return 1;
}

/* kernel.mojo:223-364. `anyHit` stops at the first accepted triangle (occlusion form; the
 * boolean equals "the closest-hit traversal reports a hit" because both follow the same order
 * with the same tMax up to that point). Returns -1 on stack overflow (the reference's stack is
 * unchecked), else the hit index or -1 for a miss via *hitIndexOut. */
static int o_traverse(const o_scene *s, const o_ray *ray, float tMax, int anyHit, int64_t *hitIndexOut,
                      float *tOut, float *uOut, float *vOut, o_stats *st) {
        float rdx = 1.0f / ray->dx, rdy = 1.0f / ray->dy, rdz = 1.0f / ray->dz;
        float ordx = ray->ox * rdx, ordy = ray->oy * rdy, ordz = ray->oz * rdz;
        int nxm = rdx >= 0.0f, nym = rdy >= 0.0f, nzm = rdz >= 0.0f;
        int64_t hitIndex = -1;
        float localT = tMax, bestU = 0.0f, bestV = 0.0f;
        int64_t stack[64];
        int toVisit = 0;
        int64_t current = 0;
        uint32_t nInterior = 0, nTris = 0;
        float org[3] = {ray->ox, ray->oy, ray->oz}, dir[3] = {ray->dx, ray->dy, ray->dz};
        int overflow = 0;
        for (;;) {
                const o_node *node = &s->nodes[current];
                if (node->count > 0) {
                        int64_t off = node->offset;
                        int done = 0;
                        for (int32_t j = 0; j < node->count; ++j) {
                                v3 a, b, c;
                                if (!o_gather(s, &s->prims[off + j], &a, &b, &c)) continue;
                                float pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {c.x, c.y, c.z};
                                float t, u, v;
                                ++nTris;
                                if (orc_intersect_triangle(org, dir, pa, pb, pc, localT, &t, &u, &v)) {
                                        localT = t; bestU = u; bestV = v;
                                        hitIndex = off + j;
                                        if (anyHit) { done = 1; break; }
                                }
                        }
                        if (done) break;
                        if (toVisit == 0) break;
                        current = stack[--toVisit];
                } else {
                        ++nInterior;
                        int64_t li = current + 1, ri = node->offset;
                        float tl, tr;
                        int lh = o_aabb(&s->nodes[li], rdx, rdy, rdz, ordx, ordy, ordz, nxm, nym, nzm, localT, &tl);
                        int rh = o_aabb(&s->nodes[ri], rdx, rdy, rdz, ordx, ordy, ordz, nxm, nym, nzm, localT, &tr);
                        if (lh && rh) {
                                if (toVisit == 64) { overflow = 1; break; }
                                if (tl <= tr) { current = li; stack[toVisit] = ri; }
                                else { current = ri; stack[toVisit] = li; }
                                ++toVisit;
                        } else if (lh) {
                                current = li;
                        } else if (rh) {
                                current = ri;
                        } else {
                                if (toVisit == 0) break;
                                current = stack[--toVisit];
                        }
                }
        }
        if (st) { st->interior = nInterior; st->tris = nTris; }
        *hitIndexOut = hitIndex;
        *tOut = hitIndex != -1 ? localT : tMax;
        *uOut = hitIndex != -1 ? bestU : 0.0f;
        *vOut = hitIndex != -1 ? bestV : 0.0f;
        return overflow ? -1 : 0;
}

/* kernel.mojo:360-364: result record. */
static void o_write(const o_scene *s, int64_t hitIndex, float t, float u, float v, o_isect *out) {
        memset(out, 0, sizeof *out);
        if (hitIndex != -1) {
                out->prim = s->prims[hitIndex];
                out->tHit = t; out->u = u; out->v = v; out->hit = 1;
        } else {
                out->prim.id1 = -1; out->prim.id2 = -1;
                out->tHit = t;
        }
}

/* mojo_traverse_bvh2 (kernel.mojo:369-373) with the empty-scene guard of
 * BoundingHierarchy.swift:392 (nodeCount == 0 -> miss). Returns -1 on stack overflow. */
int orc_traverse_one(const o_scene *s, int64_t nodeCount, const o_ray *ray, float tMax, o_isect *out,
                     o_stats *st) {
        int64_t hi = -1;
        float t = tMax, u = 0.0f, v = 0.0f;
        int rc = 0;
        if (nodeCount > 0) rc = o_traverse(s, ray, tMax, 0, &hi, &t, &u, &v, st);
        else if (st) { st->interior = 0; st->tris = 0; }
        o_write(s, hi, t, u, v, out);
        return rc;
}

typedef struct {
        const o_scene *s;
        int64_t nodeCount;
        const o_ray *rays;
        const float *tMax;
        int64_t n;
        o_isect *out;          /* closest-hit output (or NULL) */
        uint8_t *occluded;     /* any-hit output (or NULL) */
        o_stats *stats;        /* per-ray counters (or NULL) */
        atomic_long next;
        atomic_int overflow;
} o_batch;

enum { O_CHUNK = 4096 };

static void *o_worker(void *arg) {
        o_batch *b = (o_batch *)arg;
        for (;;) {
                int64_t start = atomic_fetch_add(&b->next, O_CHUNK);
                if (start >= b->n) break;
                int64_t end = start + O_CHUNK < b->n ? start + O_CHUNK : b->n;
                for (int64_t i = start; i < end; ++i) {
                        int64_t hi = -1;
                        float t = b->tMax[i], u = 0.0f, v = 0.0f;
                        o_stats st = {0, 0};
                        int rc = 0;
                        if (b->nodeCount > 0)
                                rc = o_traverse(b->s, &b->rays[i], b->tMax[i], b->occluded != NULL, &hi, &t, &u, &v, &st);
                        if (rc) atomic_store(&b->overflow, 1);
                        if (b->out) o_write(b->s, hi, t, u, v, &b->out[i]);
                        if (b->occluded) b->occluded[i] = hi != -1;
                        if (b->stats) b->stats[i] = st;
                }
        }
        return NULL;
}

static int o_run(o_batch *b, int nthreads) {
        atomic_init(&b->next, 0);
        atomic_init(&b->overflow, 0);
        if (nthreads <= 1) {
                o_worker(b);
        } else {
                pthread_t th[256];
                if (nthreads > 256) nthreads = 256;
                int started = 0;
                for (int i = 0; i < nthreads; ++i)
                        if (pthread_create(&th[i], NULL, o_worker, b) == 0) ++started;
                if (started == 0) o_worker(b);
                for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
        }
        return atomic_load(&b->overflow) ? -1 : 0;
}

/* mojo_cpu_traverse_batch (kernel.mojo:378-390): closest hit for n rays; stats may be NULL. */
int orc_traverse_batch(const o_scene *s, int64_t nodeCount, const o_ray *rays, const float *tMax,
                       int64_t n, o_isect *out, o_stats *stats, int nthreads) {
        o_batch b = {s, nodeCount, rays, tMax, n, out, NULL, stats, 0, 0};
        return o_run(&b, nthreads);
}

/* Occlusion (BoundingHierarchy.swift:348-384): occluded[i] = closest hit exists, computed with
 * an early exit at the first accepted triangle; stats counted up to that exit. */
int orc_occluded_batch(const o_scene *s, int64_t nodeCount, const o_ray *rays, const float *tMax,
                       int64_t n, uint8_t *occluded, o_stats *stats, int nthreads) {
        o_batch b = {s, nodeCount, rays, tMax, n, NULL, occluded, stats, 0, 0};
        return o_run(&b, nthreads);
}

/* Brute force over every primitive in primIds order with a shrinking tMax: the closest t is
 * independent of the BVH layout (each triangle's t is computed by the same arithmetic), so a
 * traversal must reproduce it bit for bit; the index may differ only on equal-t ties. */
void orc_brute_closest(const o_scene *s, int64_t primCount, const o_ray *rays, const float *tMax,
                       int64_t n, int64_t *hitIndex, float *tOut) {
        for (int64_t i = 0; i < n; ++i) {
                float org[3] = {rays[i].ox, rays[i].oy, rays[i].oz};
                float dir[3] = {rays[i].dx, rays[i].dy, rays[i].dz};
                float best = tMax[i];
                int64_t hi = -1;
                for (int64_t k = 0; k < primCount; ++k) {
                        v3 a, b, c;
                        if (!o_gather(s, &s->prims[k], &a, &b, &c)) continue;
                        float pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {c.x, c.y, c.z};
                        float t, u, v;
                        if (orc_intersect_triangle(org, dir, pa, pb, pc, best, &t, &u, &v)) { best = t; hi = k; }
                }
                hitIndex[i] = hi;
                tOut[i] = best;
        }
}
