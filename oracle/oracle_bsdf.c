/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's BSDFs in their local
 * shading frame, used by the oracle's MIS integrator (oracle_mis.c) as the checker of the GPU
 * kernels; never shipped, never on the product path.
 *
 *   spherical helpers, reflect, refract ........ Reflection/Reflection.swift:1-52
 *   FresnelDielectric ........................... Reflection/FresnelDielectric.swift:3-21
 *   FresnelConductor ............................ Reflection/FresnelConductor.swift:4-36
 *   Trowbridge-Reitz D, Lambda, visible-normal
 *     slope sampling ........................... Reflection/TrowbridgeReitzDistribution.swift:9-114
 *   G1, G, VNDF pdf ............................. Reflection/MicrofacetDistribution.swift:26-39
 *   DiffuseBsdf + default cosine sampling ...... Bsdf/DiffuseBsdf.swift, Reflection/Bsdf.swift:23-51
 *   MicrofacetReflection ........................ Bsdf/MicrofacetReflection.swift:1-40
 *   DielectricBsdf (smooth and rough) .......... Bsdf/DielectricBsdf.swift:19-276
 *   LayeredBsdf f / sample / pdf ................ Bsdf/LayeredBsdf.swift:55-589, 847-854
 *   CoatedConductorBsdf ......................... Bsdf/CoatedConductorBsdf.swift:3-56
 *   HenyeyGreenstein ............................ Medium/HenyeyGreenstein.swift:1-46
 * The layered BSDF draws from a deterministic stream (lrng) where the reference uses system
 * randomness. Compiled with -ffp-contract=off; cos, sin, exp, log from the C library.
 *
 * Pinning (tests/test_oracle_bsdf.py): Fresnel closed forms (normal incidence, Brewster angle,
 * total internal reflection, conductor with k = 0 = dielectric), Snell's law for refract, the
 * normalisation of D (integral of D cos = 1) and of the visible normals (integral of D G1 <wo,h>
 * = cos wo), sampler-vs-pdf consistency (importance-sampled integrals equal quadratures) for the
 * conductor, the rough dielectric and the layered BSDF, and energy bounds.
 */
#include "oracle_bsdf.h"

#include <float.h>
#include <string.h>

/* ---- spherical geometry in the local frame (Reflection.swift) ----------------------------- */
static inline float cos_theta(v3 v) { return v.z; }
static inline float cos2_theta(v3 v) { return v.z * v.z; }
static inline float abs_cos_theta(v3 v) { return fabsf(v.z); }
static inline float sin2_theta(v3 v) { return smax(0.0f, 1.0f - cos2_theta(v)); }
static inline float sin_theta(v3 v) { return sqrtf(sin2_theta(v)); }
static inline float tan2_theta(v3 v) { return sin2_theta(v) / cos2_theta(v); }
static inline float tan_theta(v3 v) { return sin_theta(v) / cos_theta(v); }
static inline float cos_phi(v3 v) {
        float st = sin_theta(v);
        return st == 0.0f ? 1.0f : sclamp(v.x / st, -1.0f, 1.0f);
}
static inline float sin_phi(v3 v) {
        float st = sin_theta(v);
        return st == 0.0f ? 0.0f : sclamp(v.y / st, -1.0f, 1.0f);
}
static inline float cos2_phi(v3 v) { return cos_phi(v) * cos_phi(v); }
static inline float sin2_phi(v3 v) { return sin_phi(v) * sin_phi(v); }
static inline int same_hemisphere(v3 a, v3 b) { return a.z * b.z > 0.0f; }
static inline float sq(float x) { return x * x; }

/* reflect(vector:by:) = -v + 2 dot(v, n) n */
static inline v3 reflect_v(v3 v, v3 n) { return add(neg(v), smul(2.0f * dot(v, n), n)); }

/* refract(incident:normal:eta:) — returns 0 on total internal reflection */
static int refract_v(v3 wi, v3 n, float eta, v3 *wt, float *etap) {
        float cosI = dot(n, wi);
        v3 fn = n;
        if (cosI < 0.0f) {
                eta = 1.0f / eta;
                cosI = -cosI;
                fn = neg(fn);
        }
        float sin2I = smax(0.0f, 1.0f - sq(cosI));
        float sin2T = sin2I / sq(eta);
        if (sin2T >= 1.0f) return 0;
        float cosT = sqrtf(1.0f - sin2T);
        *wt = add(dvs(neg(wi), eta), smul(cosI / eta - cosT, fn));
        *etap = eta;
        return 1;
}

/* ---- Fresnel -------------------------------------------------------------------------------- */
float orc_fresnel_dielectric(float cosI, float eta) {
        cosI = sclamp(cosI, -1.0f, 1.0f);
        if (cosI < 0.0f) {
                eta = 1.0f / eta;
                cosI = -cosI;
        }
        float sin2I = 1.0f - sq(cosI);
        float sin2T = sin2I / sq(eta);
        if (sin2T >= 1.0f) return 1.0f;
        float cosT = sqrtf(1.0f - sin2T);
        float rI = eta * cosI;
        float rT = eta * cosT;
        float par = (rI - cosT) / (rI + cosT);
        float perp = (cosI - rT) / (cosI + rT);
        return (sq(par) + sq(perp)) / 2.0f;
}

/* FresnelConductor.evaluate(cosTheta:) with etaI = white */
static rgb fresnel_conductor(float cosTheta, rgb etaT, rgb kT) {
        float c = sclamp(fabsf(cosTheta), -1.0f, 1.0f);
        rgb eta = cdivc(etaT, C1(1.0f));
        rgb etak = cdivc(kT, C1(1.0f));
        float cos2 = c * c;
        float sin2 = 1.0f - cos2;
        rgb eta2 = cmul(eta, eta);
        rgb etak2 = cmul(etak, etak);
        rgb term0 = cssub(csub(eta2, etak2), sin2);
        rgb a2b2 = csqrt3(cadd(cmul(term0, term0), cmul(cscale(eta2, 4.0f), etak2)));
        rgb term1 = csadd(a2b2, cos2);
        rgb termA = csqrt3(cscale(cadd(a2b2, term0), 0.5f));
        rgb term2 = cscale(termA, 2.0f * c);
        rgb rs = cdivc(csub(term1, term2), cadd(term1, term2));
        rgb term3 = csadd(cscale(a2b2, cos2), sin2 * sin2);
        rgb term4 = cscale(term2, sin2);
        rgb rp = cdivc(cmul(rs, csub(term3, term4)), cadd(term3, term4));
        return cscale(cadd(rp, rs), 0.5f);
}
void orc_fresnel_conductor(float cosTheta, const float *eta, const float *k, float *out) {
        rgb r = fresnel_conductor(cosTheta, C(eta[0], eta[1], eta[2]), C(k[0], k[1], k[2]));
        out[0] = r.r; out[1] = r.g; out[2] = r.b;
}

/* ---- Trowbridge-Reitz ------------------------------------------------------------------------ */
typedef struct { float ax, ay; } trd;

static float tr_D(trd d, v3 h) {
        float t2 = tan2_theta(h);
        if (isinf(t2)) return 0.0f;
        float cos4 = cos2_theta(h) * cos2_theta(h);
        float e = t2 * (sq(cos_phi(h) / d.ax) + sq(sin_phi(h) / d.ay));
        return 1.0f / (PI_F * d.ax * d.ay * cos4 * sq(1.0f + e));
}
static float tr_lambda(trd d, v3 v) {
        float at = fabsf(tan_theta(v));
        if (isinf(at)) return 0.0f;
        float aterm = sqrtf(cos2_phi(v) * d.ax * d.ax + sin2_phi(v) * d.ay * d.ay);
        float a2t2 = (aterm * at) * (aterm * at);
        return (-1.0f + sqrtf(1.0f + a2t2)) / 2.0f;
}
static float tr_G1(trd d, v3 v) { return 1.0f / (1.0f + tr_lambda(d, v)); }
static float tr_G(trd d, v3 wo, v3 wi) { return 1.0f / (1.0f + tr_lambda(d, wo) + tr_lambda(d, wi)); }
static float tr_pdf(trd d, v3 wo, v3 h) { return tr_D(d, h) * tr_G1(d, wo) * absdot(wo, h) / abs_cos_theta(wo); }
static int tr_smooth(trd d) { return smax(d.ax, d.ay) < 0.001f; }

static void tr_sample11(float cosTheta, float u0, float u1, float *sx, float *sy) {
        if (cosTheta > 0.9999f) {
                float r = sqrtf(u0 / (1.0f - u0));
                float phi = 2.0f * PI_F * u1;
                *sx = r * cosf(phi);
                *sy = r * sinf(phi);
                return;
        }
        float sinTheta = sqrtf(smax(0.0f, 1.0f - cosTheta * cosTheta));
        float tanTheta = sinTheta / cosTheta;
        float invTan = 1.0f / tanTheta;
        float G1 = 2.0f / (1.0f + sqrtf(1.0f + 1.0f / (invTan * invTan)));
        float A = 2.0f * u0 / G1 - 1.0f;
        float tmp = 1.0f / (A * A - 1.0f);
        if (tmp > 1e10f) tmp = 1e10f;
        float B = tanTheta;
        float D = sqrtf(smax(0.0f, B * B * tmp * tmp - (A * A - B * B) * tmp));
        float x1 = B * tmp - D;
        float x2 = B * tmp + D;
        float s0 = (A < 0.0f || x2 > 1.0f / tanTheta) ? x1 : x2;
        float S, u = u1;
        if (u > 0.5f) {
                S = 1.0f;
                u = 2.0f * (u - 0.5f);
        } else {
                S = -1.0f;
                u = 2.0f * (0.5f - u);
        }
        float z = (u * (u * (u * 0.27385f - 0.73369f) + 0.46341f)) /
                  (u * (u * (u * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        *sx = s0;
        *sy = S * z * sqrtf(1.0f + s0 * s0);
}
static v3 tr_sample_h(trd d, v3 wo, float u0, float u1) {
        int flip = wo.z < 0.0f;
        v3 w = flip ? neg(wo) : wo;
        v3 st = normalized(V(d.ax * w.x, d.ay * w.y, w.z));
        float s0, s1;
        tr_sample11(cos_theta(st), u0, u1, &s0, &s1);
        float cp = cos_phi(st), sp = sin_phi(st);
        float tmp = cp * s0 - sp * s1;
        s1 = sp * s0 + cp * s1;
        s0 = tmp;
        s0 = d.ax * s0;
        s1 = d.ay * s1;
        v3 h = normalized(V(-s0, -s1, 1.0f));
        return flip ? neg(h) : h;
}

/* ---- diffuse -------------------------------------------------------------------------------- */
/* Sampling.swift:3-25 */
void orc_concentric_disk_xy(float u0, float u1, float *x, float *y) {
        float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
        if (ox == 0.0f && oy == 0.0f) { *x = 0.0f; *y = 0.0f; return; }
        float theta, r;
        if (fabsf(ox) > fabsf(oy)) { r = ox; theta = (PI_F / 4.0f) * (oy / ox); }
        else { r = oy; theta = (PI_F / 2.0f) - (PI_F / 4.0f) * (ox / oy); }
        *x = r * cosf(theta);
        *y = r * sinf(theta);
}
v3 orc_cosine_hemisphere_v(float u0, float u1) {
        float x, y;
        orc_concentric_disk_xy(u0, u1, &x, &y);
        float z2 = 1.0f - x * x - y * y;
        return V(x, y, sqrtf(z2 > 0.0f ? z2 : 0.0f));
}

/* ---- MicrofacetReflection (conductor) -------------------------------------------------------- */
typedef struct { rgb R; trd d; rgb eta, k; } mfr;

static rgb mfr_f(const mfr *m, v3 wo, v3 wi) {
        float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
        v3 h = add(wo, wi);
        if (co == 0.0f || ci == 0.0f) return C1(0.0f);
        if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return C1(0.0f);
        h = normalized(h);
        rgb F = fresnel_conductor(dot(wi, h), m->eta, m->k);
        float D = tr_D(m->d, h);
        float G = tr_G(m->d, wo, wi);
        return cdiv(cmul(cscale(cscale(m->R, D), G), F), 4.0f * ci * co);
}
static float mfr_pdf(const mfr *m, v3 wo, v3 wi) {
        if (!same_hemisphere(wo, wi)) return 0.0f;
        v3 h = normalized(add(wo, wi));
        return tr_pdf(m->d, wo, h) / (4.0f * dot(wo, h));
}
static bsample mfr_sample(const mfr *m, v3 wo, float u0, float u1) {
        bsample s = {C1(0.0f), V(0.0f, 0.0f, 0.0f), 0.0f};
        if (wo.z == 0.0f) return s;
        v3 h = tr_sample_h(m->d, wo, u0, u1);
        v3 wi = reflect_v(wo, h);
        s.f = mfr_f(m, wo, wi);
        s.wi = wi;
        s.pdf = mfr_pdf(m, wo, wi);
        return s;
}

/* ---- DielectricBsdf -------------------------------------------------------------------------- */
typedef struct { float eta; trd d; } diel;

static int diel_specular(const diel *b) { return tr_smooth(b->d); }

static rgb diel_f(const diel *b, v3 wo, v3 wi) {
        if (b->eta == 1.0f || tr_smooth(b->d)) return C1(0.0f);
        float co = cos_theta(wo), ci = cos_theta(wi);
        int reflect = ci * co > 0.0f;
        float etap = 1.0f;
        if (!reflect) etap = co > 0.0f ? b->eta : (1.0f / b->eta);
        v3 h = add(mul(wi, etap), wo);
        if (ci == 0.0f || co == 0.0f || len2(h) == 0.0f) return C1(0.0f);
        h = normalized(h);
        if (h.x * 0.0f + h.y * 0.0f + h.z * 1.0f < 0.0f) h = neg(h);
        if (dot(h, wi) * cos_theta(wi) < 0.0f || dot(h, wo) * cos_theta(wo) < 0.0f) return C1(0.0f);
        float F = orc_fresnel_dielectric(dot(wo, h), b->eta);
        if (reflect) {
                float D = tr_D(b->d, h);
                float G = tr_G(b->d, wo, wi);
                float num = D * G * F;
                float den = fabsf(4.0f * ci * co);
                return C1(num / den);
        }
        float den = sq(dot(wi, h) + dot(wo, h) / etap) * abs_cos_theta(wi) * abs_cos_theta(wo);
        float D = tr_D(b->d, h);
        float G = tr_G(b->d, wo, wi);
        float Ft = 1.0f - F;
        float adi = fabsf(dot(wi, h));
        float ado = fabsf(dot(wo, h));
        float adio = adi * ado;
        float num = D * Ft * G * adio;
        num /= sq(etap);
        return C1(num / den);
}

static bsample diel_sample(const diel *b, v3 wo, float u0, float u1, float u2, int radiance) {
        bsample s = {C1(0.0f), V(0.0f, 0.0f, 0.0f), 0.0f};
        if (b->eta == 1.0f || tr_smooth(b->d)) {
                float R = orc_fresnel_dielectric(cos_theta(wo), b->eta);
                float T = 1.0f - R;
                float pR = R / (R + T);
                float pT = T / (R + T);
                if (u0 < pR) {
                        v3 wi = V(-wo.x, -wo.y, wo.z);
                        s.f = C1(R / abs_cos_theta(wi));
                        s.wi = wi;
                        s.pdf = pR;
                        return s;
                }
                v3 wi;
                float etap;
                if (!refract_v(wo, V(0.0f, 0.0f, 1.0f), b->eta, &wi, &etap)) return s;
                rgb f = C1(T / abs_cos_theta(wi));
                if (radiance) f = cdiv(f, sq(etap));
                s.f = f;
                s.wi = wi;
                s.pdf = pT;
                return s;
        }
        v3 h = tr_sample_h(b->d, wo, u0, u1);
        float R = orc_fresnel_dielectric(dot(wo, h), b->eta);
        float T = 1.0f - R;
        float pR = R / (R + T);
        float pT = T / (R + T);
        if (u2 < pR) {
                v3 wi = reflect_v(wo, h);
                if (!same_hemisphere(wo, wi)) return s;
                float pdf = tr_pdf(b->d, wo, h) / (4.0f * absdot(wo, h)) * pR;
                float D = tr_D(b->d, h);
                float G = tr_G(b->d, wo, wi);
                float num = D * G * R;
                float den = 4.0f * cos_theta(wi) * cos_theta(wo);
                s.f = C1(num / den);
                s.wi = wi;
                s.pdf = pdf;
                return s;
        }
        v3 wi;
        float etap;
        if (!refract_v(wo, h, b->eta, &wi, &etap)) return s;
        if (same_hemisphere(wo, wi) || wi.z == 0.0f) return s;
        float denom = sq(dot(wi, h) + dot(wo, h) / etap);
        float dwm = absdot(wi, h) / denom;
        float pdf = tr_pdf(b->d, wo, h) * dwm * pT;
        float D = tr_D(b->d, h);
        float G = tr_G(b->d, wo, wi);
        rgb f = C1(T * D * G *
                   fabsf(dot(wi, h) * dot(wo, h) / (abs_cos_theta(wi) * abs_cos_theta(wo) * denom)));
        if (radiance) f = cdiv(f, sq(etap));
        s.f = f;
        s.wi = wi;
        s.pdf = pdf;
        return s;
}

static float diel_pdf(const diel *b, v3 wo, v3 wi) {
        if (b->eta == 1.0f || diel_specular(b)) return 0.0f;
        float co = cos_theta(wo), ci = cos_theta(wi);
        int reflect = ci * co > 0.0f;
        float etap = 1.0f;
        if (!reflect) etap = co > 0.0f ? b->eta : 1.0f / b->eta;
        v3 h = add(mul(wi, etap), wo);
        if (co == 0.0f || ci == 0.0f || len2(h) == 0.0f) return 0.0f;
        h = normalized(h);
        if (h.x * 0.0f + h.y * 0.0f + h.z * 1.0f < 0.0f) h = neg(h);
        if (dot(h, wi) * ci < 0.0f || dot(h, wo) * co < 0.0f) return 0.0f;
        float R = orc_fresnel_dielectric(dot(wo, h), b->eta);
        float T = 1.0f - R;
        if (reflect) {
                float dp = tr_pdf(b->d, wo, h);
                float four = 4.0f * absdot(wo, h);
                float p = R / (R + T);
                return dp / four * p;
        }
        float denom = sq(dot(wi, h) + dot(wo, h) / etap);
        float dwm = absdot(wi, h) / denom;
        return tr_pdf(b->d, wo, h) * dwm * T / (R + T);
}

/* ---- Henyey-Greenstein -------------------------------------------------------------------- */
static float hg_phase(float cosTheta, float g) {
        float denom = 1.0f + g * g + 2.0f * g * cosTheta;
        return (1.0f / (4.0f * PI_F)) * (1.0f - g * g) / (denom * sqrtf(denom));
}
static float hg_sample(float g, v3 wo, lrng *r, v3 *wi) {
        float u0 = lrng_next(r), u1 = lrng_next(r);
        float cosTheta;
        if (fabsf(g) < 1e-3f) {
                cosTheta = 1.0f - 2.0f * u0;
        } else {
                float sqrTerm = (1.0f - g * g) / (1.0f + g - 2.0f * g * u0);
                cosTheta = -(1.0f + g * g - sqrTerm * sqrTerm) / (2.0f * g);
        }
        float sinTheta = sqrtf(smax(0.0f, 1.0f - cosTheta * cosTheta));
        float phi = 2.0f * PI_F * u1;
        v3 e1, e2;
        coordinate_system(wo, &e1, &e2);
        v3 vx = mul(e1, sinTheta * cosf(phi));
        v3 vy = mul(e2, sinTheta * sinf(phi));
        v3 vz = mul(wo, cosTheta);
        *wi = add(add(vx, vy), vz);
        return hg_phase(cosTheta, g);
}

/* ---- LayeredBsdf<DielectricBsdf, MicrofacetReflection>, two-sided ----------------------------- */
typedef struct {
        diel top;
        mfr bottom;
        float thickness;
        rgb albedo;
        float g;
        int maxDepth, nSamples;
} layered;

/* the interfaces: z == 0 is the bottom (conductor), z == thickness the top (dielectric) */
static rgb lay_if_f(const layered *L, int bottom, v3 wo, v3 wi) {
        return bottom ? mfr_f(&L->bottom, wo, wi) : diel_f(&L->top, wo, wi);
}
static float lay_if_pdf(const layered *L, int bottom, v3 wo, v3 wi) {
        return bottom ? mfr_pdf(&L->bottom, wo, wi) : diel_pdf(&L->top, wo, wi);
}
static bsample lay_if_sample(const layered *L, int bottom, v3 wo, float u0, float u1, float u2, int radiance) {
        return bottom ? mfr_sample(&L->bottom, wo, u0, u1) : diel_sample(&L->top, wo, u0, u1, u2, radiance);
}
static bsample lay_if_sample_rng(const layered *L, int bottom, v3 wo, lrng *r, int radiance) {
        float u0 = lrng_next(r), u1 = lrng_next(r), u2 = lrng_next(r);
        return lay_if_sample(L, bottom, wo, u0, u1, u2, radiance);
}
static int valid(const bsample *s) { return !cblack(s->f) && s->wi.z != 0.0f && s->pdf != 0.0f; }
static int is_reflection(const bsample *s, v3 wo) { return same_hemisphere(s->wi, wo); }

static rgb lay_tr(const layered *L, float dz, v3 w) {
        (void)L;
        if (fabsf(dz) < FLT_MIN) return C1(1.0f);
        float val = fabsf(dz / w.z);
        return C1(expf(-val));
}

static rgb lay_f(const layered *L, v3 wo, v3 wi, lrng *r) {
        const int topSpec = diel_specular(&L->top), botSpec = 0;
        rgb result = C1(0.0f);
        if (wo.z < 0.0f) {
                wo = neg(wo);
                wi = neg(wi);
        }
        const int enteredTop = 1;
        const int same = same_hemisphere(wo, wi);
        const float exitZ = (same != enteredTop) ? 0.0f : L->thickness;
        if (same) result = cmul(C1((float)L->nSamples), lay_if_f(L, 0, wo, wi));
        for (int s = 0; s < L->nSamples; ++s) {
                bsample wos = lay_if_sample_rng(L, 0, wo, r, 1);
                if (!valid(&wos) || is_reflection(&wos, wo) || wos.wi.z == 0.0f) continue;
                bsample wis = (same != enteredTop) ? lay_if_sample_rng(L, 1, wi, r, 0) : lay_if_sample_rng(L, 0, wi, r, 0);
                if (!valid(&wis) || is_reflection(&wis, wi) || wis.wi.z == 0.0f) continue;
                rgb beta = cdiv(cscale(wos.f, abs_cos_theta(wos.wi)), wos.pdf);
                float z = enteredTop ? L->thickness : 0.0f;
                v3 w = wos.wi;
                for (int depth = 0; depth < L->maxDepth; ++depth) {
                        if (depth > 3 && cmax(beta) < 0.25f) {
                                float q = smax(0.0f, 1.0f - cmax(beta));
                                if (lrng_next(r) < q) break;
                                beta = cdiv(beta, 1.0f - q);
                        }
                        if (cblack(L->albedo)) {
                                z = (z == L->thickness) ? 0.0f : L->thickness;
                                beta = cmul(beta, lay_tr(L, L->thickness, w));
                        } else {
                                const float sigmaT = 1.0f;
                                float dz = -logf(1.0f - lrng_next(r)) / (sigmaT / fabsf(w.z));
                                float zp = w.z > 0.0f ? (z + dz) : (z - dz);
                                if (z == zp) continue;
                                if (zp > 0.0f && zp < L->thickness) {
                                        float misW = 1.0f;
                                        float ph = hg_phase(dot(neg(w), neg(wis.wi)), L->g);
                                        int exitSpec = (exitZ == 0.0f) ? botSpec : topSpec;
                                        if (!exitSpec) {
                                                float pf = wis.pdf, pg = ph;
                                                misW = (pf == 0.0f || pg == 0.0f) ? 0.0f : (pf * pf) / (pf * pf + pg * pg);
                                        }
                                        rgb t = cmul(cscale(cscale(cmul(beta, L->albedo), ph), misW),
                                                     lay_tr(L, zp - exitZ, wis.wi));
                                        result = cadd(result, cdiv(cmul(t, wis.f), wis.pdf));
                                        v3 nw;
                                        float ppdf = hg_sample(L->g, neg(w), r, &nw);
                                        if (ppdf == 0.0f || nw.z == 0.0f) continue;
                                        beta = cmul(beta, cdiv(cscale(L->albedo, ppdf), ppdf));
                                        w = nw;
                                        z = zp;
                                        continue;
                                }
                                z = sclamp(zp, 0.0f, L->thickness);
                        }
                        if (z == exitZ) {
                                bsample bs = lay_if_sample_rng(L, z == 0.0f, neg(w), r, 1);
                                if (!valid(&bs) || bs.pdf == 0.0f || bs.wi.z == 0.0f) break;
                                if (!is_reflection(&bs, neg(w))) break;
                                beta = cmul(beta, cdiv(cscale(bs.f, abs_cos_theta(bs.wi)), bs.pdf));
                                w = bs.wi;
                        } else {
                                int nonExitSpec = (z == 0.0f) ? botSpec : topSpec;
                                int exitSpec = (exitZ == 0.0f) ? botSpec : topSpec;
                                if (!nonExitSpec) {
                                        float misW = 1.0f;
                                        if (!exitSpec) {
                                                float pg = lay_if_pdf(L, z == 0.0f, neg(w), neg(wis.wi));
                                                float pf = wis.pdf;
                                                misW = (pf == 0.0f || pg == 0.0f) ? 0.0f : (pf * pf) / (pf * pf + pg * pg);
                                        }
                                        rgb ne = lay_if_f(L, z == 0.0f, neg(w), neg(wis.wi));
                                        rgb t = cmul(cscale(cscale(cmul(beta, ne), abs_cos_theta(wis.wi)), misW),
                                                     lay_tr(L, L->thickness, wis.wi));
                                        result = cadd(result, cdiv(cmul(t, wis.f), wis.pdf));
                                }
                                bsample bs = lay_if_sample_rng(L, z == 0.0f, neg(w), r, 1);
                                if (!valid(&bs) || bs.pdf == 0.0f || bs.wi.z == 0.0f) break;
                                if (!is_reflection(&bs, neg(w))) break;
                                beta = cmul(beta, cdiv(cscale(bs.f, abs_cos_theta(bs.wi)), bs.pdf));
                                w = bs.wi;
                                if (!exitSpec) {
                                        rgb fExit = lay_if_f(L, exitZ == 0.0f, neg(w), wi);
                                        if (!cblack(fExit)) {
                                                float misW = 1.0f;
                                                if (!nonExitSpec) {
                                                        float pg = lay_if_pdf(L, exitZ == 0.0f, neg(w), wi);
                                                        float pf = bs.pdf;
                                                        misW = (pf == 0.0f || pg == 0.0f) ? 0.0f : (pf * pf) / (pf * pf + pg * pg);
                                                }
                                                result = cadd(result, cscale(cmul(cmul(beta, lay_tr(L, L->thickness, bs.wi)), fExit),
                                                                             misW));
                                        }
                                }
                        }
                }
        }
        return cdiv(result, (float)L->nSamples);
}

static bsample lay_sample(const layered *L, v3 wo, float u0, float u1, float u2, lrng *r) {
        bsample bad = {C1(0.0f), V(0.0f, 0.0f, 0.0f), 0.0f};
        int flip = 0;
        if (wo.z < 0.0f) {
                wo = neg(wo);
                flip = 1;
        }
        bsample bs = lay_if_sample(L, 0, wo, u0, u1, u2, 1);
        if (!valid(&bs) || bs.pdf == 0.0f || bs.wi.z == 0.0f) return bad;
        if (is_reflection(&bs, wo)) {
                if (flip) bs.wi = neg(bs.wi);
                return bs;
        }
        v3 w = bs.wi;
        rgb f = cscale(bs.f, abs_cos_theta(bs.wi));
        float pdf = bs.pdf;
        float z = L->thickness;
        for (int depth = 0; depth < L->maxDepth; ++depth) {
                float rrBeta = cmax(f) / pdf;
                if (depth > 3 && rrBeta < 0.25f) {
                        float q = smax(0.0f, 1.0f - rrBeta);
                        if (lrng_next(r) < q) return bad;
                        pdf *= 1.0f - q;
                }
                if (w.z == 0.0f) return bad;
                if (!cblack(L->albedo)) {
                        const float sigmaT = 1.0f;
                        float dz = -logf(1.0f - lrng_next(r)) / (sigmaT / abs_cos_theta(w));
                        float zp = w.z > 0.0f ? (z + dz) : (z - dz);
                        if (zp == z) return bad;
                        if (zp > 0.0f && zp < L->thickness) {
                                v3 nw;
                                float ph = hg_sample(L->g, neg(w), r, &nw);
                                if (ph == 0.0f || nw.z == 0.0f) return bad;
                                f = cmul(f, cscale(L->albedo, ph));
                                pdf *= ph;
                                w = nw;
                                z = zp;
                                continue;
                        }
                        z = sclamp(zp, 0.0f, L->thickness);
                } else {
                        z = (z == L->thickness) ? 0.0f : L->thickness;
                        f = cmul(f, C1(1.0f));
                }
                bsample s = lay_if_sample_rng(L, z == 0.0f, neg(w), r, 1);
                if (!valid(&s) || s.pdf == 0.0f || s.wi.z == 0.0f) return bad;
                f = cmul(f, s.f);
                pdf *= s.pdf;
                w = s.wi;
                if (!is_reflection(&s, neg(w))) {
                        if (flip) w = neg(w);
                        bsample out = {f, w, pdf};
                        return out;
                }
                f = cscale(f, abs_cos_theta(s.wi));
        }
        return bad;
}

static float ph_w(float f, float g) { return (f == 0.0f || g == 0.0f) ? 0.0f : (f * f) / (f * f + g * g); }

static float lay_pdf(const layered *L, v3 wo, v3 wi, lrng *r) {
        const int topSpec = diel_specular(&L->top), botSpec = 0;
        if (wo.z < 0.0f) {
                wo = neg(wo);
                wi = neg(wi);
        }
        float pdfSum = 0.0f;
        if (same_hemisphere(wo, wi)) pdfSum += (float)L->nSamples * diel_pdf(&L->top, wo, wi);
        for (int s = 0; s < L->nSamples; ++s) {
                if (same_hemisphere(wo, wi)) {
                        bsample os = lay_if_sample_rng(L, 0, wo, r, 1);
                        bsample is = lay_if_sample_rng(L, 0, wi, r, 1);
                        if (valid(&os) && !is_reflection(&os, wo) && valid(&is) && !is_reflection(&is, wi)) {
                                if (topSpec) {
                                        pdfSum += mfr_pdf(&L->bottom, neg(os.wi), neg(is.wi));
                                } else {
                                        bsample rs = lay_if_sample_rng(L, 1, neg(os.wi), r, 1);
                                        if (valid(&rs) && rs.pdf > 0.0f) {
                                                if (botSpec) {
                                                        pdfSum += diel_pdf(&L->top, neg(rs.wi), wi);
                                                } else {
                                                        float rPDF = mfr_pdf(&L->bottom, neg(os.wi), neg(is.wi));
                                                        float w = ph_w(is.pdf, rPDF);
                                                        pdfSum += w * rPDF;
                                                        float tPDF = diel_pdf(&L->top, neg(rs.wi), wi);
                                                        w = ph_w(rs.pdf, tPDF);
                                                        pdfSum += w * tPDF;
                                                }
                                        }
                                }
                        }
                } else {
                        bsample os = lay_if_sample_rng(L, 0, wo, r, 1);
                        bsample is = lay_if_sample_rng(L, 1, wi, r, 1);
                        if (valid(&os) && !is_reflection(&os, wo) && valid(&is) && !is_reflection(&is, wi)) {
                                float p1 = diel_pdf(&L->top, wo, neg(is.wi));
                                float p2 = mfr_pdf(&L->bottom, neg(os.wi), wi);
                                pdfSum += (p1 + p2) / 2.0f;
                        }
                }
        }
        float t = 0.9f;
        return (1.0f - t) * (1.0f / (4.0f * PI_F)) + t * (pdfSum / (float)L->nSamples);
}

/* ---- dispatch ------------------------------------------------------------------------------- */
static void split(const bx *b, mfr *m, diel *d, layered *L) {
        if (b->type == BX_CONDUCTOR) {
                m->R = b->R;
                m->d.ax = b->ax; m->d.ay = b->ay;
                m->eta = b->eta; m->k = b->k;
        } else if (b->type == BX_DIELECTRIC) {
                d->eta = b->ieta;
                d->d.ax = b->ax; d->d.ay = b->ay;
        } else if (b->type == BX_COATED_CONDUCTOR) {
                L->top.eta = b->ieta;
                L->top.d.ax = b->ax; L->top.d.ay = b->ay;
                L->bottom.R = b->R;
                L->bottom.d.ax = b->cax; L->bottom.d.ay = b->cay;
                L->bottom.eta = b->eta; L->bottom.k = b->k;
                L->thickness = b->thickness;
                L->albedo = b->albedo;
                L->g = b->g;
                L->maxDepth = b->maxDepth;
                L->nSamples = b->nSamples;
        }
}

rgb bx_f(const bx *b, v3 wo, v3 wi, lrng *r) {
        mfr m; diel d; layered L;
        split(b, &m, &d, &L);
        switch (b->type) {
        case BX_DIFFUSE: return cdiv(b->R, PI_F);
        case BX_CONDUCTOR: return mfr_f(&m, wo, wi);
        case BX_DIELECTRIC: return diel_f(&d, wo, wi);
        case BX_COATED_CONDUCTOR: return lay_f(&L, wo, wi, r);
        }
        return C1(0.0f);
}

float bx_pdf(const bx *b, v3 wo, v3 wi, lrng *r) {
        mfr m; diel d; layered L;
        split(b, &m, &d, &L);
        switch (b->type) {
        case BX_DIFFUSE:
                if (!same_hemisphere(wo, wi)) return 0.0f;
                return abs_cos_theta(wi) / PI_F;
        case BX_CONDUCTOR: return mfr_pdf(&m, wo, wi);
        case BX_DIELECTRIC: return diel_pdf(&d, wo, wi);
        case BX_COATED_CONDUCTOR: return lay_pdf(&L, wo, wi, r);
        }
        return 0.0f;
}

bsample bx_sample(const bx *b, v3 wo, float u0, float u1, float u2, lrng *r) {
        mfr m; diel d; layered L;
        split(b, &m, &d, &L);
        switch (b->type) {
        case BX_DIFFUSE: {
                v3 wi = orc_cosine_hemisphere_v(u0, u1);
                if (wo.z < 0.0f) wi.z = -wi.z;
                bsample s = {cdiv(b->R, PI_F), wi, same_hemisphere(wo, wi) ? abs_cos_theta(wi) / PI_F : 0.0f};
                return s;
        }
        case BX_CONDUCTOR: return mfr_sample(&m, wo, u0, u1);
        case BX_DIELECTRIC: return diel_sample(&d, wo, u0, u1, u2, 1);
        case BX_COATED_CONDUCTOR: return lay_sample(&L, wo, u0, u1, u2, r);
        }
        bsample s = {C1(0.0f), V(0.0f, 0.0f, 0.0f), 0.0f};
        return s;
}

int bx_specular(const bx *b) {
        if (b->type == BX_DIELECTRIC) {
                trd t = {b->ax, b->ay};
                return tr_smooth(t);
        }
        return 0;
}

rgb bx_albedo(const bx *b) { return b->type == BX_DIFFUSE ? b->R : C1(1.0f); }

static float remap(float r, int on) { return on ? sqrtf(r) : r; }

void bx_from_material(const mb_material *m, bx *b) {
        memset(b, 0, sizeof *b);
        b->type = m->type;
        const int rm = (m->flags & 1) != 0;
        switch (m->type) {
        case BX_DIFFUSE:
                b->R = C(m->reflectance[0], m->reflectance[1], m->reflectance[2]);
                break;
        case BX_CONDUCTOR:
                b->R = C1(1.0f);
                b->ax = smax(m->roughness[0], 1e-3f);
                b->ay = smax(m->roughness[1], 1e-3f);
                b->eta = C(m->conductorEta[0], m->conductorEta[1], m->conductorEta[2]);
                b->k = C(m->conductorK[0], m->conductorK[1], m->conductorK[2]);
                break;
        case BX_DIELECTRIC:
                b->ieta = m->eta;
                b->ax = remap(m->roughness[0], rm);
                b->ay = remap(m->roughness[1], rm);
                break;
        case BX_COATED_CONDUCTOR:
                b->ieta = m->eta;
                b->R = C(m->reflectance[0], m->reflectance[1], m->reflectance[2]);
                b->ax = smax(remap(m->roughness[0], rm), 0.001f);
                b->ay = smax(remap(m->roughness[1], rm), 0.001f);
                b->cax = smax(remap(m->conductorRoughness[0], rm), 0.001f);
                b->cay = smax(remap(m->conductorRoughness[1], rm), 0.001f);
                b->eta = C(m->conductorEta[0], m->conductorEta[1], m->conductorEta[2]);
                b->k = C(m->conductorK[0], m->conductorK[1], m->conductorK[2]);
                b->thickness = smax(m->thickness, FLT_MIN);
                b->albedo = C(m->albedo[0], m->albedo[1], m->albedo[2]);
                b->g = m->g;
                b->maxDepth = m->maxDepth;
                b->nSamples = m->nSamples;
                break;
        }
}

/* ---- exports for the known-answer and consistency tests ------------------------------------- */
static void to3(v3 v, float *o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
static void rgb3(rgb c, float *o) { o[0] = c.r; o[1] = c.g; o[2] = c.b; }

void orc_bx_f(const mb_material *m, const float *wo, const float *wi, uint64_t seed, float *out) {
        bx b;
        bx_from_material(m, &b);
        lrng r = lrng_make(seed, 0, 0, 0, 0, 0);
        rgb3(bx_f(&b, V(wo[0], wo[1], wo[2]), V(wi[0], wi[1], wi[2]), &r), out);
}
float orc_bx_pdf(const mb_material *m, const float *wo, const float *wi, uint64_t seed) {
        bx b;
        bx_from_material(m, &b);
        lrng r = lrng_make(seed, 0, 0, 0, 0, 0);
        return bx_pdf(&b, V(wo[0], wo[1], wo[2]), V(wi[0], wi[1], wi[2]), &r);
}
/* out: f[3], wi[3], pdf */
void orc_bx_sample(const mb_material *m, const float *wo, const float *u, uint64_t seed, float *out) {
        bx b;
        bx_from_material(m, &b);
        lrng r = lrng_make(seed, 0, 0, 0, 0, 0);
        bsample s = bx_sample(&b, V(wo[0], wo[1], wo[2]), u[0], u[1], u[2], &r);
        rgb3(s.f, out);
        to3(s.wi, out + 3);
        out[6] = s.pdf;
}
float orc_tr_D(float ax, float ay, const float *h) { trd d = {ax, ay}; return tr_D(d, V(h[0], h[1], h[2])); }
float orc_tr_G1(float ax, float ay, const float *w) { trd d = {ax, ay}; return tr_G1(d, V(w[0], w[1], w[2])); }
float orc_tr_pdf(float ax, float ay, const float *wo, const float *h) {
        trd d = {ax, ay};
        return tr_pdf(d, V(wo[0], wo[1], wo[2]), V(h[0], h[1], h[2]));
}
void orc_tr_sample_h(float ax, float ay, const float *wo, float u0, float u1, float *out) {
        trd d = {ax, ay};
        to3(tr_sample_h(d, V(wo[0], wo[1], wo[2]), u0, u1), out);
}
int orc_refract(const float *wi, const float *n, float eta, float *out) {
        v3 wt = {0.0f, 0.0f, 0.0f};
        float etap = 0.0f;
        int ok = refract_v(V(wi[0], wi[1], wi[2]), V(n[0], n[1], n[2]), eta, &wt, &etap);
        to3(wt, out);
        out[3] = etap;
        return ok;
}
