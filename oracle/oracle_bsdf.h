/*
 * ORACLE — test infrastructure only. Shared declarations of the oracle's MIS integrator
 * (oracle_mis.c) and its BSDFs (oracle_bsdf.c): float vector and RGB helpers written in the
 * reference's operation order, and the local-frame BSDF ("bx") interface.
 */
#ifndef ORACLE_BSDF_H
#define ORACLE_BSDF_H

#include <math.h>
#include <stdint.h>

typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 smul(float s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline v3 dvs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }   /* Three.swift:131-133 */
static inline float absdot(v3 a, v3 b) { return fabsf(dot(a, b)); }
static inline v3 cross(v3 a, v3 b) {                                                  /* Three.swift:119-129 */
        return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float len2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }        /* Three.swift:15-17 */
static inline v3 normalized(v3 a) {                                                   /* Three.swift:108-117 */
        float l = sqrtf(len2(a));
        return l == 0.0f ? a : dvs(a, l);
}
/* Geometry.swift:63-77 */
static inline void coordinate_system(v3 v1, v3 *v2, v3 *v3o) {
        if (fabsf(v1.x) > fabsf(v1.y))
                *v2 = dvs(V(-v1.z, 0.0f, v1.x), sqrtf(v1.x * v1.x + v1.z * v1.z));
        else
                *v2 = dvs(V(0.0f, v1.z, -v1.y), sqrtf(v1.y * v1.y + v1.z * v1.z));
        *v3o = cross(v1, *v2);
}
/* Swift max(x, y) = y >= x ? y : x; min(x, y) = y < x ? y : x */
static inline float smax(float x, float y) { return y >= x ? y : x; }
static inline float smin(float x, float y) { return y < x ? y : x; }
/* Utility.swift:28-30 */
static inline float sclamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

typedef struct { float r, g, b; } rgb;
static inline rgb C(float r, float g, float b) { rgb c = {r, g, b}; return c; }
static inline rgb C1(float i) { return C(i, i, i); }
static inline rgb cadd(rgb a, rgb b) { return C(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline rgb csub(rgb a, rgb b) { return C(a.r - b.r, a.g - b.g, a.b - b.b); }
static inline rgb cmul(rgb a, rgb b) { return C(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline rgb cdivc(rgb a, rgb b) { return C(a.r / b.r, a.g / b.g, a.b / b.b); }
static inline rgb cscale(rgb a, float s) { return C(a.r * s, a.g * s, a.b * s); }
static inline rgb csadd(rgb a, float s) { return C(a.r + s, a.g + s, a.b + s); }
static inline rgb cssub(rgb a, float s) { return C(a.r - s, a.g - s, a.b - s); }
static inline rgb cdiv(rgb a, float s) { return C(a.r / s, a.g / s, a.b / s); }
static inline rgb csqrt3(rgb a) { return C(sqrtf(a.r), sqrtf(a.g), sqrtf(a.b)); }
static inline int cblack(rgb a) { return a.r == 0.0f && a.g == 0.0f && a.b == 0.0f; }
static inline float cmax(rgb a) {                                                      /* Spectrum.swift:357-359 */
        float gb = a.b >= a.g ? a.b : a.g;
        return gb >= a.r ? gb : a.r;
}
static inline float cavg(rgb a) { return (a.r + a.g + a.b) / 3.0f; }                  /* Spectrum.swift:353-355 */

static const float PI_F = 3.14159265358979323846f;

/* Deterministic stand-in for the reference's RandomSampler (system random) inside the layered
 * BSDF: a SplitMix64 stream per (seed, pixel, sample, bounce, call); each draw is the top 24 bits
 * of the next output times 2^-24. */
typedef struct { uint64_t s; } lrng;
static inline uint64_t splitmix_mix(uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
}
static inline lrng lrng_make(uint64_t seed, int32_t px, int32_t py, int32_t si, int32_t bounce, int32_t call) {
        uint64_t z = seed ^ ((uint64_t)(uint32_t)px << 40) ^ ((uint64_t)(uint32_t)py << 20) ^ (uint64_t)(uint32_t)si;
        z += 0x9E3779B97F4A7C15ull * (uint64_t)(uint32_t)(bounce + 1);
        z ^= (uint64_t)(uint32_t)call << 58;
        lrng r = {splitmix_mix(z)};
        return r;
}
static inline float lrng_next(lrng *r) {
        r->s += 0x9E3779B97F4A7C15ull;
        return (float)(uint32_t)(splitmix_mix(r->s) >> 40) * (1.0f / 16777216.0f);
}

/* Local-frame BSDF of one material at one shading point (the reference's BxDFs). */
enum { BX_DIFFUSE = 1, BX_CONDUCTOR = 2, BX_DIELECTRIC = 3, BX_COATED_CONDUCTOR = 4 };
typedef struct {
        int type;
        rgb R;                 /* diffuse reflectance; conductor: white; coated: conductor reflectance */
        float ax, ay;          /* conductor alpha, dielectric alpha, or the coating's interface alpha */
        rgb eta, k;            /* conductor (plain or under the coating) */
        float ieta;            /* dielectric / coating refractive index */
        float cax, cay;        /* coated conductor: the conductor's alpha */
        float thickness;       /* coated conductor layer */
        rgb albedo;
        float g;
        int maxDepth, nSamples;
} bx;

typedef struct { rgb f; v3 wi; float pdf; } bsample;

rgb bx_f(const bx *b, v3 wo, v3 wi, lrng *rng);
float bx_pdf(const bx *b, v3 wo, v3 wi, lrng *rng);
bsample bx_sample(const bx *b, v3 wo, float u0, float u1, float u2, lrng *rng);
int bx_specular(const bx *b);
rgb bx_albedo(const bx *b);

/* Material record (include/gonzales_amd.h GzMaterial, 96 bytes) and its BSDF parameters as the
 * reference's getBsdf derives them (Conductor.swift:3-17, Dielectric.swift:44-54,
 * CoatedConductor.swift:31-75). */
typedef struct {
        int32_t type;
        int32_t flags;            /* bit 0: remap roughness */
        float reflectance[3];
        float eta;
        float roughness[2];
        float conductorRoughness[2];
        float conductorEta[3];
        float conductorK[3];
        float thickness;
        float albedo[3];
        float g;
        int32_t maxDepth;
        int32_t nSamples;
        int32_t reserved;
} mb_material;
_Static_assert(sizeof(mb_material) == 96, "GzMaterial");

void bx_from_material(const mb_material *m, bx *b);

#endif
