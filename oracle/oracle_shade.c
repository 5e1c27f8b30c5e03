/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's wavefront diffuse shade
 * (the shade step of Core/Tile.swift's bounce loop), used as the checker by tests/; never
 * shipped and never on the product path.
 *
 * Restates Sources/mojoKernel/kernel.mojo:
 *   PCG32 (pcg32_srandom seeding, XSH-RR output) .. kernel.mojo:99-118
 *   shade_core ..................................... kernel.mojo:676-773
 *   mojo_cpu_shade_batch ........................... kernel.mojo:393-403
 * and the PathState_C / Material_C layouts of mojo_kernel.h:11-15, 36-44.
 * sqrt is correctly rounded; cos/sin come from the C library (the reference takes them from the
 * Mojo standard library, which is not available here), so directions are compared with a
 * tolerance. Pinning: PCG32 against O'Neill's published pcg32 demo output, cosine-hemisphere and
 * ONB closed forms (tests/test_oracle_shade.py).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef struct { int64_t id1, id2, materialIndex; int8_t type; int8_t pad[7]; } s_prim_id;
typedef struct { const float *points; const int64_t *faceIndices; const int64_t *vertexIndices; } s_mesh;
typedef struct { s_prim_id prim; float tHit, u, v; int8_t hit; int8_t pad[3]; } s_isect;
typedef struct {
        float ox, oy, oz, dx, dy, dz;
        float thr[3], est[3], alb[3];
        int32_t pad0;
        uint64_t pcgState, pcgInc;
        int8_t active;
        int8_t pad1[7];
} s_path;
typedef struct { int8_t type; int8_t pad[3]; float albedo[3]; float emission[3]; } s_material;

_Static_assert(sizeof(s_path) == 88, "PathState_C");
_Static_assert(offsetof(s_path, pcgState) == 64, "PathState_C.pcgState");
_Static_assert(offsetof(s_path, active) == 80, "PathState_C.active");
_Static_assert(sizeof(s_material) == 28, "Material_C");

typedef struct { uint64_t state, inc; } s_pcg;

/* kernel.mojo:107-112 */
static uint32_t s_next(s_pcg *g) {
        uint64_t old = g->state;
        g->state = old * 6364136223846793005ull + g->inc;
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

/* kernel.mojo:100-105 */
static void s_seed(s_pcg *g, uint64_t initstate, uint64_t initseq) {
        g->state = 0;
        g->inc = (initseq << 1) | 1u;
        (void)s_next(g);
        g->state += initstate;
        (void)s_next(g);
}

/* kernel.mojo:114-115 */
static float s_float(s_pcg *g) { return (float)(s_next(g) >> 8) * (1.0f / 16777216.0f); }

/* Exposed for the PCG32 known-answer test: n outputs of pcg32 seeded (initstate, initseq). */
void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t *out) {
        s_pcg g;
        s_seed(&g, initstate, initseq);
        for (int i = 0; i < n; ++i) out[i] = s_next(&g);
}

typedef struct { float x, y, z; } f3;
static inline f3 f3sub(f3 a, f3 b) { f3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline f3 f3cross(f3 a, f3 b) {
        f3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
        return r;
}
static inline float f3dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline f3 f3scale(f3 a, float s) { f3 r = {a.x * s, a.y * s, a.z * s}; return r; }

/* kernel.mojo:676-773, one path. */
static void s_shade(s_path *p, const s_isect *is, const s_mesh *meshes, const s_material *mats) {
        if (p->active == 0) return;
        if (is->hit == 0) { p->active = 0; return; }
        const s_material *m = &mats[is->prim.materialIndex];
        if (m->type == 2) {
                p->est[0] += p->thr[0] * m->emission[0];
                p->est[1] += p->thr[1] * m->emission[1];
                p->est[2] += p->thr[2] * m->emission[2];
                p->active = 0;
                return;
        }
        if (m->type != 1) { p->active = 0; return; }
        int64_t mesh, base;
        if (is->prim.type == 0) { mesh = is->prim.id1; base = is->prim.id2; }
        else if (is->prim.type == 1 || is->prim.type == 2 || is->prim.type == 3) {
                mesh = is->prim.id2 >> 32;
                base = (is->prim.id2 & 0xFFFFFFFFLL) * 3;
        } else { p->active = 0; return; }
        const s_mesh *me = &meshes[mesh];
        int64_t i0 = me->vertexIndices[base], i1 = me->vertexIndices[base + 1], i2 = me->vertexIndices[base + 2];
        f3 p0 = {me->points[i0 * 4], me->points[i0 * 4 + 1], me->points[i0 * 4 + 2]};
        f3 p1 = {me->points[i1 * 4], me->points[i1 * 4 + 1], me->points[i1 * 4 + 2]};
        f3 p2 = {me->points[i2 * 4], me->points[i2 * 4 + 1], me->points[i2 * 4 + 2]};
        f3 n = f3cross(f3sub(p1, p0), f3sub(p2, p0));
        float nlen = f3dot(n, n);
        if (nlen > 0.0f) n = f3scale(n, 1.0f / sqrtf(nlen));
        f3 d = {p->dx, p->dy, p->dz};
        if (f3dot(n, d) > 0.0f) n = f3scale(n, -1.0f);
        s_pcg g = {p->pcgState, 0};
        s_seed(&g, p->pcgState, p->pcgInc);
        float u1 = s_float(&g), u2 = s_float(&g);
        p->pcgState = g.state;
        float r = sqrtf(u1);
        float theta = 2.0f * 3.14159265359f * u2;
        float x = r * cosf(theta), y = r * sinf(theta);
        float z2 = 1.0f - u1;
        float z = sqrtf(z2 > 0.0f ? z2 : 0.0f);
        float sign = n.z >= 0.0f ? 1.0f : -1.0f;
        float a = -1.0f / (sign + n.z);
        float b = n.x * n.y * a;
        f3 t = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
        f3 bt = {b, sign + n.y * n.y * a, -n.y};
        f3 dir = {t.x * x + bt.x * y + n.x * z, t.y * x + bt.y * y + n.y * z, t.z * x + bt.z * y + n.z * z};
        float dlen = f3dot(dir, dir);
        if (dlen > 0.0f) dir = f3scale(dir, 1.0f / sqrtf(dlen));
        f3 o = {p->ox, p->oy, p->oz};
        f3 org = {o.x + d.x * is->tHit + n.x * 0.0001f, o.y + d.y * is->tHit + n.y * 0.0001f,
                  o.z + d.z * is->tHit + n.z * 0.0001f};
        p->ox = org.x; p->oy = org.y; p->oz = org.z;
        p->dx = dir.x; p->dy = dir.y; p->dz = dir.z;
        p->thr[0] *= m->albedo[0];
        p->thr[1] *= m->albedo[1];
        p->thr[2] *= m->albedo[2];
}

/* mojo_cpu_shade_batch (kernel.mojo:393-403) */
void orc_shade_batch(void *paths, int64_t n, const void *isects, const void *meshes, const void *mats) {
        s_path *p = (s_path *)paths;
        const s_isect *is = (const s_isect *)isects;
        for (int64_t i = 0; i < n; ++i) s_shade(&p[i], &is[i], (const s_mesh *)meshes, (const s_material *)mats);
}
