/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's MIS path integrator
 * (VolumePathIntegrator), used as the checker by tests/; never shipped, never on the product path.
 *
 * One path at a time, bounce after bounce, as evaluateRayPath/oneBounce/shadeBounce do
 * (Integrator/VolumePathIntegrator.swift:35-463):
 *   closest hit + Triangle.computeSurfaceInteraction ...... BoundingHierarchy.swift:387-456,
 *                                                          Shape/Triangle.swift:363-474
 *   miss: infinite-light radiance if bounce 0 / specular .. VolumePathIntegrator.swift:380-393
 *   bounce >= maxDepth stops .............................. :400-402
 *   surfaceEstimate: emission (bounce 0 / specular), BSDF,
 *     albedo/first normal, one light with MIS, continuation  :325-371
 *   sampleOneLight / estimateDirect / sampleMultipleImportance :173-266
 *   sampleLightSource, sampleDistributionFunction ......... :67-133
 *   Russian roulette (bounce > 1) ......................... :269-283
 *   PowerLightSampler.chooseLight (own sampler: Morton 0) . Light/PowerLightSampler.swift:4-49
 *   AreaLight.sample / emittedRadiance / power ............ Light/AreaLight.swift:13-49
 *   Triangle.sample / area, Sampleable.sample(point:) ..... Triangle.swift:517-544, Sampleable.swift:33-44
 *   light pdf of a BSDF direction (always 0 for triangles,
 *     tHit = 0) ........................................... Sampleable.swift:49-69
 *   Visibility.occluded, spawn-ray offsets ................ Core/Visibility.swift:5-24,
 *                                                          Core/Interaction.swift:15-53
 *   ShadingFrame, BsdfFrame, FramedBsdf ................... Core/ShadingFrame.swift,
 *                                                          Reflection/BsdfFrame.swift, FramedBsdf.swift
 *   the BSDFs in their local frame ......................... oracle_bsdf.c
 *   power heuristic ....................................... Sampler/Sampling.swift:27-30
 *   InfiniteLight (constant radiance map: uniform equal-area
 *     sphere sampling, pdf 1/(4 pi), power pi 50^2 avg) .... Light/InfiniteLight.swift:39-138
 * Deterministic replacement for the reference's system-random Russian roulette: u = the top 24
 * bits of a SplitMix64 hash of (seed, pixel, sample, bounce), times 2^-24.
 * Compiled with -ffp-contract=off. Transcendentals (cos, sin) come from the C library.
 *
 * Pinning (tests/test_oracle_mis.py): the reference's KATs (SamplingTests.swift:9-89,
 * BsdfTests.swift:18-49, InteractionTests.swift:33-40), closed forms (triangle sampling is
 * uniform in area; light pdf = d^2 / (cos A)), and the white-furnace closed form
 * L = Le (1 - a^(D+1)) / (1 - a) with the light-pdf correction switched on.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "oracle_bsdf.h"

/* ---- shared layouts (match include/gonzales_amd.h byte for byte) -------------------------- */
typedef struct { int64_t id1, id2, materialIndex; int8_t type; int8_t pad[7]; } mb_prim;
typedef struct { const float *points; const int64_t *faceIndices; const int64_t *vertexIndices; } mb_mesh;
typedef struct { float ox, oy, oz, dx, dy, dz; } mb_ray;
typedef struct { mb_prim prim; float tHit, u, v; int8_t hit; int8_t pad[3]; } mb_isect;
typedef struct { const void *nodes; const mb_prim *prims; const mb_mesh *meshes; int64_t meshCount; } mb_scene;

typedef struct {
        mb_ray ray;
        float throughput[3], estimate[3], albedo[3], firstNormal[3];
        float filterWeight;
        int32_t pixelX, pixelY, sampleIndex, dim, lightDim, bounce, flags, pad0, pad1;
} mb_path;
_Static_assert(sizeof(mb_path) == 112, "GzPath");

typedef struct { float brightness[3]; int32_t mesh; int64_t triangle; } mb_light;
_Static_assert(sizeof(mb_light) == 24, "GzAreaLight");

typedef struct {
        int32_t maxDepth;
        int32_t flags;          /* FIX_LIGHT_PDF | FIX_SKY_DIRECTION */
        int64_t zsSeed;
        int32_t zsLog2Spp, zsBase4Digits;
        uint64_t rrSeed;        /* Russian roulette and the layered BSDF's streams */
        float sky[3];           /* constant infinite light radiance */
        int32_t hasSky;
} mb_params;
_Static_assert(sizeof(mb_params) == 48, "GzMisParams");

enum { PATH_ACTIVE = 1, PATH_SPECULAR = 2 };
enum { FIX_LIGHT_PDF = 1, FIX_SKY_DIRECTION = 2 };
/* the layered BSDF's stream of each BSDF call of a bounce */
enum { CALL_LIGHT_F = 1, CALL_LIGHT_PDF = 2, CALL_MIS_SAMPLE = 3, CALL_CONTINUE = 4 };

/* oracle_traverse.c / oracle_zsobol.c */
typedef struct { uint32_t interior, tris; } mb_stats;
int orc_traverse_one(const mb_scene *s, int64_t nodeCount, const mb_ray *ray, float tMax, mb_isect *out, mb_stats *st);
uint32_t orc_mix_bits(uint64_t v);
float orc_sobol_sample(int64_t index, int dim, uint32_t seed, const uint32_t *m);
uint64_t orc_encode_morton2(uint32_t x, uint32_t y);
typedef struct { int64_t seed; int32_t log2spp, nBase4Digits; uint64_t morton; int32_t dim, pad; } mb_zs;
uint64_t orc_zsobol_sample_index(const mb_zs *z);
float orc_zsobol_get1d(mb_zs *z, const uint32_t *m);
void orc_zsobol_get2d(mb_zs *z, const uint32_t *m, float *out);
void orc_zsobol_get3d(mb_zs *z, const uint32_t *m, float *out);
/* oracle_bsdf.c */
void orc_concentric_disk_xy(float u0, float u1, float *x, float *y);
v3 orc_cosine_hemisphere_v(float u0, float u1);

/* ---- scene access --------------------------------------------------------------------------- */
typedef struct {
        const mb_scene *s;
        int64_t nodeCount;
        const mb_material *mats;
        int32_t nMats;
        const mb_light *lights;
        int32_t nLights;
        const uint32_t *table;
        mb_params p;
        /* power light sampler: area lights in index order, then the sky */
        float power[4096 + 1];
        float cumulative[4096 + 1];
        int32_t nChoices;
        float totalPower;
} mb_ctx;

static void tri_points(const mb_ctx *c, int64_t mesh, int64_t base, v3 *p0, v3 *p1, v3 *p2) {
        const mb_mesh *m = &c->s->meshes[mesh];
        int64_t i0 = m->vertexIndices[base], i1 = m->vertexIndices[base + 1], i2 = m->vertexIndices[base + 2];
        *p0 = V(m->points[i0 * 4], m->points[i0 * 4 + 1], m->points[i0 * 4 + 2]);
        *p1 = V(m->points[i1 * 4], m->points[i1 * 4 + 1], m->points[i1 * 4 + 2]);
        *p2 = V(m->points[i2 * 4], m->points[i2 * 4 + 1], m->points[i2 * 4 + 2]);
}

/* Triangle.area (Triangle.swift:522-525) */
static float tri_area(v3 p0, v3 p1, v3 p2) { return 0.5f * sqrtf(len2(cross(sub(p1, p0), sub(p2, p0)))); }

typedef struct {
        v3 position, normal, shadingNormal, outgoing, dpdu;
        int32_t material;   /* materials[] index */
        int32_t areaLight;  /* -1 if none */
} mb_si;

/* Closest hit -> SurfaceInteraction (BoundingHierarchy.swift:420-452 + Triangle.swift:363-474
 * without vertex normals or uvs). Returns 0 when the reference's intersect returns nil. */
static int surface_interaction(const mb_ctx *c, const mb_ray *ray, const mb_isect *is, mb_si *si) {
        if (!is->hit) return 0;
        int64_t mesh, base;
        const mb_prim *pr = &is->prim;
        if (pr->type == 0) { mesh = pr->id1; base = pr->id2; }
        else { mesh = pr->id2 >> 32; base = (pr->id2 & 0xFFFFFFFFLL) * 3; }
        v3 p0, p1, p2;
        tri_points(c, mesh, base, &p0, &p1, &p2);
        float b0 = 1.0f - is->u - is->v, b1 = is->u, b2 = is->v;
        v3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
        v3 pHit = add(add(mul(p0, b0), mul(p1, b1)), mul(p2, b2));
        v3 n = normalized(cross(dp02, dp12));
        /* no uvs: the uv determinant is 0, dpdu comes from the geometric normal */
        v3 gn = cross(sub(p2, p0), sub(p1, p0));
        if (len2(gn) == 0.0f) return 0;
        v3 dpdu, dpdv;
        coordinate_system(normalized(gn), &dpdu, &dpdv);
        v3 sn = n;
        v3 S = normalized(dpdu);
        v3 T = cross(S, sn);
        if (len2(T) > 0.0f) {
                T = normalized(T);
                S = cross(T, sn);
        } else {
                coordinate_system(sn, &S, &T);
        }
        si->position = pHit;
        si->normal = normalized(n);
        si->shadingNormal = normalized(sn);
        si->outgoing = normalized(neg(V(ray->dx, ray->dy, ray->dz)));
        si->dpdu = S;
        if (pr->type == 3) { si->areaLight = (int32_t)pr->id1; si->material = 0; }
        else { si->areaLight = -1; si->material = (int32_t)pr->materialIndex; }
        return 1;
}

/* ---- BSDF in the shading frame (BsdfFrame / FramedBsdf) ------------------------------------ */
typedef struct { v3 x, y, z, ng; bx b; } mb_bsdf;

static void bsdf_make(const mb_ctx *c, const mb_si *si, mb_bsdf *b) {
        b->y = normalized(si->dpdu);            /* BsdfFrame(interaction:) */
        b->z = si->shadingNormal;
        b->x = cross(b->y, b->z);                /* ShadingFrame(y:z:) */
        b->ng = si->normal;
        bx_from_material(&c->mats[si->material], &b->b);
}
static v3 to_local(const mb_bsdf *b, v3 w) { return normalized(V(dot(w, b->x), dot(w, b->y), dot(w, b->z))); }
static v3 to_world(const mb_bsdf *b, v3 l) { return add(add(mul(b->x, l.x), mul(b->y, l.y)), mul(b->z, l.z)); }

/* FramedBsdf.evaluateWorldSpace (FramedBsdf.swift:18-34): every BSDF keeps the protocol's
 * isReflective = true, isTransmissive = false, so only the reflection side is evaluated. */
static rgb bsdf_eval_world(const mb_bsdf *b, v3 wo, v3 wi, lrng *r) {
        v3 ol = to_local(b, wo), il = to_local(b, wi);
        int reflect = dot(wi, b->ng) * dot(wo, b->ng) > 0.0f;
        return reflect ? bx_f(&b->b, ol, il, r) : C(0, 0, 0);
}
/* FramedBsdf.probabilityDensityWorldSpace (FramedBsdf.swift:36-44) */
static float bsdf_pdf_world(const mb_bsdf *b, v3 wo, v3 wi, lrng *r) {
        v3 il = to_local(b, wi), ol = to_local(b, wo);
        if (ol.z == 0.0f) return 0.0f;
        return bx_pdf(&b->b, ol, il, r);
}
/* FramedBsdf.sampleWorldSpace (FramedBsdf.swift:46-57) */
static void bsdf_sample_world(const mb_bsdf *b, v3 wo, const float *u, rgb *f, v3 *wi, float *pdf, lrng *r) {
        v3 ol = to_local(b, wo);
        bsample s = bx_sample(&b->b, ol, u[0], u[1], u[2], r);
        *f = s.f;
        *wi = to_world(b, s.wi);
        *pdf = s.pdf;
}

static float power_heuristic(float f, float g) {  /* Sampling.swift:27-30 */
        if (f == 0.0f || g == 0.0f) return 0.0f;
        return (f * f) / (f * f + g * g);
}

/* ---- lights ---------------------------------------------------------------------------------- */
static void light_points(const mb_ctx *c, int32_t li, v3 *p0, v3 *p1, v3 *p2) {
        tri_points(c, c->lights[li].mesh, 3 * c->lights[li].triangle, p0, p1, p2);
}

static void build_light_sampler(mb_ctx *c) {
        c->nChoices = 0;
        for (int32_t i = 0; i < c->nLights; ++i) {
                v3 p0, p1, p2;
                light_points(c, i, &p0, &p1, &p2);
                rgb L = C(c->lights[i].brightness[0], c->lights[i].brightness[1], c->lights[i].brightness[2]);
                c->power[c->nChoices++] = cavg(L) * tri_area(p0, p1, p2) * PI_F;  /* AreaLight.swift:47-49 */
        }
        if (c->p.hasSky) {
                rgb L = C(c->p.sky[0], c->p.sky[1], c->p.sky[2]);
                float r = 100.0f / 2.0f;
                c->power[c->nChoices++] = PI_F * (r * r) * cavg(L);  /* InfiniteLight.swift:135-138 */
        }
        float total = 0.0f;
        for (int32_t i = 0; i < c->nChoices; ++i) total = total + c->power[i];
        c->totalPower = total;
        for (int32_t i = 0; i < c->nChoices; ++i)
                c->cumulative[i] = i == 0 ? c->power[0] : c->cumulative[i - 1] + c->power[i];
}

/* PowerLightSampler.chooseLight + lowerBound (PowerLightSampler.swift:4-49) */
static int32_t choose_light(const mb_ctx *c, float u, float *pdf) {
        float key = u * c->totalPower;
        int32_t lo = 0, hi = c->nChoices, idx = -1;
        while (lo < hi) {
                int32_t mid = lo + (hi - lo) / 2;
                if (c->cumulative[mid] == key) { idx = mid; break; }
                if (c->cumulative[mid] < key) lo = mid + 1;
                else hi = mid;
        }
        if (idx < 0) idx = lo;
        if (idx >= c->nChoices) idx = c->nChoices - 1;  /* the reference would read past the end */
        *pdf = c->power[idx] / c->totalPower;
        return idx;
}

/* InfiniteLight.equalAreaSquareToSphere (InfiniteLight.swift:112-127); `unit` applies the factor
 * sqrt(2 - r^2) to x and y that the reference omits (FIX_SKY_DIRECTION). */
static v3 equal_area_square_to_sphere(float tu, float tv, int unit) {
        float u = 2.0f * tu - 1.0f, v = 2.0f * tv - 1.0f;
        float au = fabsf(u), av = fabsf(v);
        float sd = 1.0f - (au + av);
        float ad = fabsf(sd);
        float r = 1.0f - ad;
        float phi = (r == 0.0f ? 1.0f : (av - au) / r + 1.0f) * (PI_F / 4.0f);
        float zz = 1.0f - r * r;
        float z = copysignf(zz > 0.0f ? zz : 0.0f, sd);
        float cp = cosf(phi), sp = sinf(phi);
        float rr = r;
        if (unit) {
                float s = 2.0f - r * r;
                rr = r * sqrtf(s > 0.0f ? s : 0.0f);
        }
        return V(copysignf(rr * cp, u), copysignf(rr * sp, v), z);
}

typedef struct { rgb L; v3 dir; float pdf; v3 from, target; } mb_lsample;

/* light.sample(point:samples:) for an area light (AreaLight.swift:21-32) or the sky */
static int light_sample(const mb_ctx *c, int32_t choice, v3 point, const float *u, mb_lsample *ls) {
        ls->from = point;
        if (choice < c->nLights) {
                v3 p0, p1, p2;
                light_points(c, choice, &p0, &p1, &p2);
                float su0 = sqrtf(u[0]);
                float b0 = 1.0f - su0, b1 = u[1] * su0;
                v3 pos = add(add(mul(p0, b0), mul(p1, b1)), mul(p2, 1.0f - b0 - b1));
                v3 ln = normalized(cross(sub(p1, p0), sub(p2, p0)));
                float pdf = 1.0f / tri_area(p0, p1, p2);
                v3 inc = normalized(sub(pos, point));
                float d2 = len2(sub(point, pos));
                float angle = absdot(ln, neg(inc));
                pdf = pdf * (d2 / angle);
                ls->dir = normalized(sub(pos, point));
                ls->target = pos;
                const mb_light *L = &c->lights[choice];
                ls->L = dot(ln, neg(ls->dir)) > 0.0f ? C(L->brightness[0], L->brightness[1], L->brightness[2]) : C(0, 0, 0);
                ls->pdf = pdf;
                return 1;
        }
        v3 d = equal_area_square_to_sphere(u[0], u[1], (c->p.flags & FIX_SKY_DIRECTION) != 0);
        ls->dir = d;
        ls->pdf = 1.0f / (4.0f * PI_F);
        ls->target = add(point, mul(d, 100.0f));
        ls->L = C(c->p.sky[0], c->p.sky[1], c->p.sky[2]);
        return 1;
}

/* Visibility.occluded (Visibility.swift:5-24) via spawnRay(from:target:) (Interaction.swift:15-24) */
static int occluded(const mb_ctx *c, v3 from, v3 target) {
        v3 dirFT = sub(target, from);
        v3 origin = add(from, mul(dirFT, 0.0001f));
        v3 d = sub(target, origin);
        mb_ray r = {origin.x, origin.y, origin.z, d.x, d.y, d.z};
        mb_isect is;
        orc_traverse_one(c->s, c->nodeCount, &r, 1.0f - 0.00001f, &is, NULL);
        return is.hit != 0;
}

/* Interaction.spawnRay(inDirection:) (Interaction.swift:32-53) */
static mb_ray spawn(const mb_si *si, v3 dir) {
        v3 off = mul(si->normal, 0.0001f);
        v3 o = dot(dir, si->normal) > 0.0f ? add(si->position, off) : add(si->position, neg(off));
        mb_ray r = {o.x, o.y, o.z, dir.x, dir.y, dir.z};
        return r;
}

/* Light pdf of a BSDF-sampled direction (AreaLight.probabilityDensityFor -> Sampleable.swift:49-69).
 * The reference passes tHit = 0, so the triangle test rejects every hit and the pdf is 0; with
 * FIX_LIGHT_PDF the direction is intersected with the chosen triangle. */
static float light_pdf(const mb_ctx *c, int32_t choice, const mb_si *si, v3 dir) {
        if (choice >= c->nLights) return 1.0f / (4.0f * PI_F);  /* constant map: mapPdf = 1 */
        if (!(c->p.flags & FIX_LIGHT_PDF)) return 0.0f;
        v3 p0, p1, p2;
        light_points(c, choice, &p0, &p1, &p2);
        mb_ray r = spawn(si, dir);
        /* Moller-Trumbore against the single light triangle */
        v3 e1 = sub(p1, p0), e2 = sub(p2, p0);
        v3 d = V(r.dx, r.dy, r.dz), o = V(r.ox, r.oy, r.oz);
        v3 pv = cross(d, e2);
        float det = dot(e1, pv);
        if (det > -1e-12f && det < 1e-12f) return 0.0f;
        float inv = 1.0f / det;
        v3 tv = sub(o, p0);
        float uu = dot(tv, pv) * inv;
        if (uu < 0.0f || uu > 1.0f) return 0.0f;
        v3 qv = cross(tv, e1);
        float vv = dot(d, qv) * inv;
        if (vv < 0.0f || uu + vv > 1.0f) return 0.0f;
        float t = dot(e2, qv) * inv;
        if (t <= 0.0f) return 0.0f;
        v3 hitp = add(o, mul(d, t));
        float d2 = len2(sub(si->position, hitp));
        v3 ln = normalized(cross(sub(p0, p2), sub(p1, p2)));
        float angle = absdot(ln, neg(dir));
        float density = d2 / (angle * tri_area(p0, p1, p2));
        return isinf(density) ? 0.0f : density;
}

/* deterministic Russian-roulette draw */
static float rr_draw(uint64_t seed, int32_t px, int32_t py, int32_t si, int32_t bounce) {
        uint64_t z = seed ^ ((uint64_t)(uint32_t)px << 40) ^ ((uint64_t)(uint32_t)py << 20) ^ (uint64_t)(uint32_t)si;
        z += 0x9E3779B97F4A7C15ull * (uint64_t)(uint32_t)(bounce + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

/* One bounce (oneBounce + shadeBounce). Returns 1 to continue. */
static int one_bounce(const mb_ctx *c, mb_path *p, mb_zs *zs, mb_zs *lzs) {
        mb_isect is;
        orc_traverse_one(c->s, c->nodeCount, &p->ray, INFINITY, &is, NULL);
        mb_si si;
        int hit = surface_interaction(c, &p->ray, &is, &si);
        rgb T = C(p->throughput[0], p->throughput[1], p->throughput[2]);
        rgb E = C(p->estimate[0], p->estimate[1], p->estimate[2]);
        int specular = (p->flags & PATH_SPECULAR) != 0;
        int result = 0;
        if (!hit) {
                if ((p->bounce == 0 || specular) && c->p.hasSky)
                        E = cadd(E, cmul(T, C(c->p.sky[0], c->p.sky[1], c->p.sky[2])));
                goto done;
        }
        if (cblack(T)) goto done;
        if (!(p->bounce < c->p.maxDepth)) goto done;
        /* surfaceEstimate */
        if ((p->bounce == 0 || specular) && si.areaLight >= 0) {
                const mb_light *L = &c->lights[si.areaLight];
                rgb Le = dot(si.normal, si.outgoing) > 0.0f ? C(L->brightness[0], L->brightness[1], L->brightness[2]) : C(0, 0, 0);
                E = cadd(E, cmul(T, Le));
        }
        mb_bsdf b;
        bsdf_make(c, &si, &b);
#define RNG(call) lrng_make(c->p.rrSeed, p->pixelX, p->pixelY, p->sampleIndex, p->bounce, (call))
        if (p->bounce == 0) {
                rgb al = bx_albedo(&b.b);
                p->albedo[0] = al.r; p->albedo[1] = al.g; p->albedo[2] = al.b;
                p->firstNormal[0] = si.normal.x; p->firstNormal[1] = si.normal.y; p->firstNormal[2] = si.normal.z;
        }
        {
                /* sampleOneLight */
                zs->dim = p->dim;
                lzs->dim = p->lightDim;
                float choicePdf;
                float ul = orc_zsobol_get1d(lzs, c->table);
                p->lightDim = lzs->dim;
                int32_t choice = choose_light(c, ul, &choicePdf);
                /* sampleMultipleImportance: light sample */
                float u2[2];
                orc_zsobol_get2d(zs, c->table, u2);
                mb_lsample ls;
                light_sample(c, choice, si.position, u2, &ls);
                rgb Lest = C(0, 0, 0);
                float lpdf = 0.0f;
                v3 linc = V(0, 0, 0);
                if (!cblack(ls.L) && !isinf(ls.pdf) && !occluded(c, ls.from, ls.target)) {
                        lrng r1 = RNG(CALL_LIGHT_F);
                        rgb scatter = cscale(bsdf_eval_world(&b, si.outgoing, ls.dir, &r1), absdot(ls.dir, si.shadingNormal));
                        Lest = cmul(scatter, ls.L);
                        lpdf = ls.pdf;
                        linc = ls.dir;
                }
                lrng r2 = RNG(CALL_LIGHT_PDF);
                float bpdfL = bsdf_pdf_world(&b, si.outgoing, linc, &r2);
                float wl = power_heuristic(lpdf, bpdfL);
                rgb lc = lpdf == 0.0f ? C(0, 0, 0) : cdiv(cscale(Lest, wl), lpdf);
                /* sampleMultipleImportance: BSDF sample */
                float u3[3];
                orc_zsobol_get3d(zs, c->table, u3);
                rgb bf;
                v3 bwi;
                float bpdf;
                lrng r3 = RNG(CALL_MIS_SAMPLE);
                bsdf_sample_world(&b, si.outgoing, u3, &bf, &bwi, &bpdf, &r3);
                bf = cscale(bf, absdot(bwi, si.shadingNormal));
                rgb bc = C(0, 0, 0);
                if (!cblack(bf) && bpdf > 0.0f) {
                        mb_ray r = spawn(&si, bwi);
                        mb_isect bis;
                        orc_traverse_one(c->s, c->nodeCount, &r, INFINITY, &bis, NULL);
                        mb_si bsi;
                        rgb bL = C(0, 0, 0);
                        int ok = 0;
                        if (surface_interaction(c, &r, &bis, &bsi)) {
                                if (bsi.areaLight >= 0) {
                                        const mb_light *L = &c->lights[bsi.areaLight];
                                        /* emittedRadiance(from: hit, inDirection: incoming) */
                                        bL = dot(bsi.normal, bwi) > 0.0f ? C(L->brightness[0], L->brightness[1], L->brightness[2]) : C(0, 0, 0);
                                        ok = 1;
                                }
                        } else if (c->p.hasSky) {
                                bL = C(c->p.sky[0], c->p.sky[1], c->p.sky[2]);
                                ok = 1;
                        }
                        if (ok) {
                                rgb best = cmul(bf, bL);
                                float lden = light_pdf(c, choice, &si, bwi);
                                float wb = power_heuristic(bpdf, lden);
                                bc = cdiv(cscale(best, wb), bpdf);
                        }
                }
                rgb est = cadd(lc, bc);
                E = cadd(E, cmul(T, cdiv(est, choicePdf)));
        }
        {
                /* continuation: bsdf.sampleWorldSpace(outgoing:uSample: get3D()) */
                float u3[3];
                orc_zsobol_get3d(zs, c->table, u3);
                p->dim = zs->dim;
                rgb f;
                v3 wi;
                float pdf;
                lrng r4 = RNG(CALL_CONTINUE);
                bsdf_sample_world(&b, si.outgoing, u3, &f, &wi, &pdf, &r4);
                if (pdf == 0.0f || isnan(pdf)) goto done;
                specular = bx_specular(&b.b);
                T = cmul(T, cdiv(cscale(f, absdot(wi, si.normal)), pdf));
                p->ray = spawn(&si, wi);
        }
        /* Russian roulette (VolumePathIntegrator.swift:269-283) */
        if (cmax(T) < 1.0f && p->bounce > 1) {
                float m = cmax(T);
                float q = 1.0f - m > 0.0f ? 1.0f - m : 0.0f;
                float u = rr_draw(c->p.rrSeed, p->pixelX, p->pixelY, p->sampleIndex, p->bounce);
                if (u < q) goto done;
                T = cdiv(T, 1.0f - q);
        }
        result = 1;
done:
        p->throughput[0] = T.r; p->throughput[1] = T.g; p->throughput[2] = T.b;
        p->estimate[0] = E.r; p->estimate[1] = E.g; p->estimate[2] = E.b;
        p->flags = (result ? PATH_ACTIVE : 0) | (specular ? PATH_SPECULAR : 0);
        return result;
}

typedef struct {
        mb_ctx *c;
        mb_path *paths;
        int64_t n;
        atomic_long next;
} mb_job;

static void trace_path(const mb_ctx *c, mb_path *p) {
        mb_zs zs, lzs;
        zs.seed = c->p.zsSeed; zs.log2spp = c->p.zsLog2Spp; zs.nBase4Digits = c->p.zsBase4Digits;
        uint64_t pm = orc_encode_morton2((uint32_t)p->pixelX, (uint32_t)p->pixelY);
        zs.morton = (zs.log2spp >= 64 ? 0 : pm << zs.log2spp) | (uint64_t)(uint32_t)p->sampleIndex;
        lzs = zs;
        lzs.morton = 0;  /* the light chooser's sampler never starts a pixel sample */
        while ((p->flags & PATH_ACTIVE) && p->bounce <= c->p.maxDepth) {
                int go = one_bounce(c, p, &zs, &lzs);
                if (!go) break;
                p->bounce += 1;
        }
        p->flags &= ~PATH_ACTIVE;
}

static void *mb_worker(void *arg) {
        mb_job *j = (mb_job *)arg;
        for (;;) {
                int64_t i = atomic_fetch_add(&j->next, 64);
                if (i >= j->n) break;
                int64_t e = i + 64 < j->n ? i + 64 : j->n;
                for (int64_t k = i; k < e; ++k) trace_path(j->c, &j->paths[k]);
        }
        return NULL;
}

/* Runs every active path to completion (bounce loop of VolumePathIntegrator.bounces,
 * :448-463). Returns 0, or -1 on bad arguments. */
int orc_mis_trace_paths(const mb_scene *s, int64_t nodeCount, const mb_material *mats, int32_t nMats,
                        const mb_light *lights, int32_t nLights, const uint32_t *table, const mb_params *params,
                        mb_path *paths, int64_t n, int nthreads) {
        if (nLights > 4096 || nMats < 1) return -1;
        static mb_ctx cstore;  /* large; callers run one batch at a time */
        mb_ctx *c = &cstore;
        c->s = s; c->nodeCount = nodeCount; c->mats = mats; c->nMats = nMats;
        c->lights = lights; c->nLights = nLights; c->table = table; c->p = *params;
        build_light_sampler(c);
        if (c->nChoices == 0) return -1;
        mb_job j;
        j.c = c; j.paths = paths; j.n = n;
        atomic_init(&j.next, 0);
        if (nthreads <= 1) { mb_worker(&j); return 0; }
        pthread_t th[64];
        if (nthreads > 64) nthreads = 64;
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, mb_worker, &j);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
        return 0;
}

/* Exposed for the known-answer tests. */
float orc_power_heuristic(float f, float g) { return power_heuristic(f, g); }
void orc_concentric_disk(float u0, float u1, float *out) { orc_concentric_disk_xy(u0, u1, &out[0], &out[1]); }
void orc_cosine_hemisphere(float u0, float u1, float *out) {
        v3 v = orc_cosine_hemisphere_v(u0, u1);
        out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
void orc_equal_area_square_to_sphere(float u, float v, int unit, float *out) {
        v3 d = equal_area_square_to_sphere(u, v, unit);
        out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
float orc_rr_draw(uint64_t seed, int32_t px, int32_t py, int32_t si, int32_t bounce) {
        return rr_draw(seed, px, py, si, bounce);
}
