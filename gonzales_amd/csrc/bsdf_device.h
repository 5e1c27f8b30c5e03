// Device-side vector/RGB arithmetic and the reference's BSDFs in their local shading frame, for
// the MIS bounce kernel. Every expression keeps the reference's operation order (no contraction):
//   spherical helpers, reflect, refract ........ Reflection/Reflection.swift:1-52
//   FresnelDielectric / FresnelConductor ........ Reflection/FresnelDielectric.swift:3-21,
//                                                 Reflection/FresnelConductor.swift:4-36
//   Trowbridge-Reitz D, Lambda, slope sampler ... Reflection/TrowbridgeReitzDistribution.swift:9-114
//   G1, G, visible-normal pdf ................... Reflection/MicrofacetDistribution.swift:26-39
//   diffuse + default cosine sampling ........... Bsdf/DiffuseBsdf.swift, Reflection/Bsdf.swift:23-51
//   MicrofacetReflection (conductor) ............ Bsdf/MicrofacetReflection.swift:1-40
//   DielectricBsdf, smooth and rough ............ Bsdf/DielectricBsdf.swift:19-276
//   LayeredBsdf f / sample / pdf (two-sided) .... Bsdf/LayeredBsdf.swift:55-589, 847-854
//   CoatedConductorBsdf, HenyeyGreenstein ....... Bsdf/CoatedConductorBsdf.swift, Medium/HenyeyGreenstein.swift
// The layered BSDF draws from a deterministic per-(path, bounce, call) stream where the
// reference uses system randomness.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gonzales_amd.h"

#pragma clang fp contract(off)

namespace gz {

constexpr float kPi = 3.14159265358979323846f;

struct f3 {
        float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 mk1(float v) { return f3{v, v, v}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 smul(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float absdot(f3 a, f3 b) { return fabsf(dot(a, b)); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
        return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ f3 normalized(f3 a) {  // Three.swift:108-117: v / |v|, or v if |v| = 0
        const float l = sqrtf(len2(a));
        return l == 0.0f ? a : divs(a, l);
}
// Geometry.swift:63-77
__device__ __forceinline__ void coordinate_system(f3 v1, f3 &v2, f3 &v3) {
        if (fabsf(v1.x) > fabsf(v1.y))
                v2 = divs(mk(-v1.z, 0.0f, v1.x), sqrtf(v1.x * v1.x + v1.z * v1.z));
        else
                v2 = divs(mk(0.0f, v1.z, -v1.y), sqrtf(v1.y * v1.y + v1.z * v1.z));
        v3 = cross(v1, v2);
}
// Swift max(x, y) = y >= x ? y : x; clamp as Utility.swift:28-30
__device__ __forceinline__ float smax(float x, float y) { return y >= x ? y : x; }
__device__ __forceinline__ float sclamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ float sq(float x) { return x * x; }

// RGB spectra share the 3-float type.
__device__ __forceinline__ f3 cadd(f3 a, f3 b) { return add(a, b); }
__device__ __forceinline__ f3 csub(f3 a, f3 b) { return sub(a, b); }
__device__ __forceinline__ f3 cmul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 cdivc(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ f3 csadd(f3 a, float s) { return mk(a.x + s, a.y + s, a.z + s); }
__device__ __forceinline__ f3 cssub(f3 a, float s) { return mk(a.x - s, a.y - s, a.z - s); }
__device__ __forceinline__ f3 csqrt3(f3 a) { return mk(sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)); }
__device__ __forceinline__ bool cblack(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ float cmax(f3 a) {  // max(r, max(g, b)), Spectrum.swift:357-359
        const float gb = a.z >= a.y ? a.z : a.y;
        return gb >= a.x ? gb : a.x;
}
__device__ __forceinline__ float power_heuristic(float f, float g) {  // Sampling.swift:27-30
        if (f == 0.0f || g == 0.0f) return 0.0f;
        return (f * f) / (f * f + g * g);
}

// SplitMix64 stream standing in for the reference's RandomSampler.
struct LayerRng {
        uint64_t s;
};
__device__ __forceinline__ uint64_t splitmix_mix(uint64_t z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
}
__device__ __forceinline__ LayerRng layer_rng(uint64_t seed, int px, int py, int si, int bounce, int call) {
        uint64_t z = seed ^ ((uint64_t)(uint32_t)px << 40) ^ ((uint64_t)(uint32_t)py << 20) ^ (uint64_t)(uint32_t)si;
        z += 0x9E3779B97F4A7C15ull * (uint64_t)(uint32_t)(bounce + 1);
        z ^= (uint64_t)(uint32_t)call << 58;
        return LayerRng{splitmix_mix(z)};
}
__device__ __forceinline__ float rng_next(LayerRng &r) {
        r.s += 0x9E3779B97F4A7C15ull;
        return (float)(uint32_t)(splitmix_mix(r.s) >> 40) * (1.0f / 16777216.0f);
}

// ---- local-frame geometry -------------------------------------------------------------------
__device__ __forceinline__ float cos_theta(f3 v) { return v.z; }
__device__ __forceinline__ float cos2_theta(f3 v) { return v.z * v.z; }
__device__ __forceinline__ float abs_cos_theta(f3 v) { return fabsf(v.z); }
__device__ __forceinline__ float sin2_theta(f3 v) { return smax(0.0f, 1.0f - cos2_theta(v)); }
__device__ __forceinline__ float sin_theta(f3 v) { return sqrtf(sin2_theta(v)); }
__device__ __forceinline__ float tan2_theta(f3 v) { return sin2_theta(v) / cos2_theta(v); }
__device__ __forceinline__ float tan_theta(f3 v) { return sin_theta(v) / cos_theta(v); }
__device__ __forceinline__ float cos_phi(f3 v) {
        const float st = sin_theta(v);
        return st == 0.0f ? 1.0f : sclamp(v.x / st, -1.0f, 1.0f);
}
__device__ __forceinline__ float sin_phi(f3 v) {
        const float st = sin_theta(v);
        return st == 0.0f ? 0.0f : sclamp(v.y / st, -1.0f, 1.0f);
}
__device__ __forceinline__ bool same_hemisphere(f3 a, f3 b) { return a.z * b.z > 0.0f; }
__device__ __forceinline__ f3 reflect_v(f3 v, f3 n) { return add(neg(v), smul(2.0f * dot(v, n), n)); }
__device__ __forceinline__ bool refract_v(f3 wi, f3 n, float eta, f3 &wt, float &etap) {
        float cosI = dot(n, wi);
        f3 fn = n;
        if (cosI < 0.0f) {
                eta = 1.0f / eta;
                cosI = -cosI;
                fn = neg(fn);
        }
        const float sin2I = smax(0.0f, 1.0f - sq(cosI));
        const float sin2T = sin2I / sq(eta);
        if (sin2T >= 1.0f) return false;
        const float cosT = sqrtf(1.0f - sin2T);
        wt = add(divs(neg(wi), eta), smul(cosI / eta - cosT, fn));
        etap = eta;
        return true;
}

// ---- Fresnel ----------------------------------------------------------------------------------
__device__ __forceinline__ float fresnel_dielectric(float cosI, float eta) {
        cosI = sclamp(cosI, -1.0f, 1.0f);
        if (cosI < 0.0f) {
                eta = 1.0f / eta;
                cosI = -cosI;
        }
        const float sin2I = 1.0f - sq(cosI);
        const float sin2T = sin2I / sq(eta);
        if (sin2T >= 1.0f) return 1.0f;
        const float cosT = sqrtf(1.0f - sin2T);
        const float rI = eta * cosI;
        const float rT = eta * cosT;
        const float par = (rI - cosT) / (rI + cosT);
        const float perp = (cosI - rT) / (cosI + rT);
        return (sq(par) + sq(perp)) / 2.0f;
}
__device__ __forceinline__ f3 fresnel_conductor(float cosTheta, f3 etaT, f3 kT) {
        const float c = sclamp(fabsf(cosTheta), -1.0f, 1.0f);
        const f3 eta = cdivc(etaT, mk1(1.0f));
        const f3 etak = cdivc(kT, mk1(1.0f));
        const float cos2 = c * c;
        const float sin2 = 1.0f - cos2;
        const f3 eta2 = cmul(eta, eta);
        const f3 etak2 = cmul(etak, etak);
        const f3 term0 = cssub(csub(eta2, etak2), sin2);
        const f3 a2b2 = csqrt3(cadd(cmul(term0, term0), cmul(mul(eta2, 4.0f), etak2)));
        const f3 term1 = csadd(a2b2, cos2);
        const f3 termA = csqrt3(mul(cadd(a2b2, term0), 0.5f));
        const f3 term2 = mul(termA, 2.0f * c);
        const f3 rs = cdivc(csub(term1, term2), cadd(term1, term2));
        const f3 term3 = csadd(mul(a2b2, cos2), sin2 * sin2);
        const f3 term4 = mul(term2, sin2);
        const f3 rp = cdivc(cmul(rs, csub(term3, term4)), cadd(term3, term4));
        return mul(cadd(rp, rs), 0.5f);
}

// ---- Trowbridge-Reitz ---------------------------------------------------------------------------
struct Trd {
        float ax, ay;
};
__device__ __forceinline__ float tr_D(Trd d, f3 h) {
        const float t2 = tan2_theta(h);
        if (isinf(t2)) return 0.0f;
        const float cos4 = cos2_theta(h) * cos2_theta(h);
        const float e = t2 * (sq(cos_phi(h) / d.ax) + sq(sin_phi(h) / d.ay));
        return 1.0f / (kPi * d.ax * d.ay * cos4 * sq(1.0f + e));
}
__device__ __forceinline__ float tr_lambda(Trd d, f3 v) {
        const float at = fabsf(tan_theta(v));
        if (isinf(at)) return 0.0f;
        const float cp = cos_phi(v), sp = sin_phi(v);
        const float aterm = sqrtf(cp * cp * d.ax * d.ax + sp * sp * d.ay * d.ay);
        const float a2t2 = (aterm * at) * (aterm * at);
        return (-1.0f + sqrtf(1.0f + a2t2)) / 2.0f;
}
__device__ __forceinline__ float tr_G1(Trd d, f3 v) { return 1.0f / (1.0f + tr_lambda(d, v)); }
__device__ __forceinline__ float tr_G(Trd d, f3 wo, f3 wi) {
        return 1.0f / (1.0f + tr_lambda(d, wo) + tr_lambda(d, wi));
}
__device__ __forceinline__ float tr_pdf(Trd d, f3 wo, f3 h) {
        return tr_D(d, h) * tr_G1(d, wo) * absdot(wo, h) / abs_cos_theta(wo);
}
__device__ __forceinline__ bool tr_smooth(Trd d) { return smax(d.ax, d.ay) < 0.001f; }

__device__ __forceinline__ void tr_sample11(float cosTheta, float u0, float u1, float &sx, float &sy) {
        if (cosTheta > 0.9999f) {
                const float r = sqrtf(u0 / (1.0f - u0));
                const float phi = 2.0f * kPi * u1;
                sx = r * cosf(phi);
                sy = r * sinf(phi);
                return;
        }
        const float sinTheta = sqrtf(smax(0.0f, 1.0f - cosTheta * cosTheta));
        const float tanTheta = sinTheta / cosTheta;
        const float invTan = 1.0f / tanTheta;
        const float G1 = 2.0f / (1.0f + sqrtf(1.0f + 1.0f / (invTan * invTan)));
        const float A = 2.0f * u0 / G1 - 1.0f;
        float tmp = 1.0f / (A * A - 1.0f);
        if (tmp > 1e10f) tmp = 1e10f;
        const float B = tanTheta;
        const float D = sqrtf(smax(0.0f, B * B * tmp * tmp - (A * A - B * B) * tmp));
        const float x1 = B * tmp - D;
        const float x2 = B * tmp + D;
        const float s0 = (A < 0.0f || x2 > 1.0f / tanTheta) ? x1 : x2;
        float S, u = u1;
        if (u > 0.5f) {
                S = 1.0f;
                u = 2.0f * (u - 0.5f);
        } else {
                S = -1.0f;
                u = 2.0f * (0.5f - u);
        }
        const float z = (u * (u * (u * 0.27385f - 0.73369f) + 0.46341f)) /
                        (u * (u * (u * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        sx = s0;
        sy = S * z * sqrtf(1.0f + s0 * s0);
}
__device__ __forceinline__ f3 tr_sample_h(Trd d, f3 wo, float u0, float u1) {
        const bool flip = wo.z < 0.0f;
        const f3 w = flip ? neg(wo) : wo;
        const f3 st = normalized(mk(d.ax * w.x, d.ay * w.y, w.z));
        float s0, s1;
        tr_sample11(cos_theta(st), u0, u1, s0, s1);
        const float cp = cos_phi(st), sp = sin_phi(st);
        const float tmp = cp * s0 - sp * s1;
        s1 = sp * s0 + cp * s1;
        s0 = tmp;
        s0 = d.ax * s0;
        s1 = d.ay * s1;
        const f3 h = normalized(mk(-s0, -s1, 1.0f));
        return flip ? neg(h) : h;
}

// ---- BSDF sample record -----------------------------------------------------------------------
struct BSample {
        f3 f;
        f3 wi;
        float pdf;
};
__device__ __forceinline__ BSample bs_none() { return BSample{mk1(0.0f), mk1(0.0f), 0.0f}; }
__device__ __forceinline__ bool bs_valid(const BSample &s) { return !cblack(s.f) && s.wi.z != 0.0f && s.pdf != 0.0f; }
__device__ __forceinline__ bool bs_reflection(const BSample &s, f3 wo) { return same_hemisphere(s.wi, wo); }

// ---- diffuse ---------------------------------------------------------------------------------
// Sampling.swift:3-25
__device__ __forceinline__ f3 cosine_hemisphere(float u0, float u1) {
        const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
        float x = 0.0f, y = 0.0f;
        if (!(ox == 0.0f && oy == 0.0f)) {
                float theta, r;
                if (fabsf(ox) > fabsf(oy)) {
                        r = ox;
                        theta = (kPi / 4.0f) * (oy / ox);
                } else {
                        r = oy;
                        theta = (kPi / 2.0f) - (kPi / 4.0f) * (ox / oy);
                }
                x = r * cosf(theta);
                y = r * sinf(theta);
        }
        const float z2 = 1.0f - x * x - y * y;
        return mk(x, y, sqrtf(z2 > 0.0f ? z2 : 0.0f));
}

// ---- MicrofacetReflection ---------------------------------------------------------------------
struct Mfr {
        f3 R;
        Trd d;
        f3 eta, k;
};
__device__ __forceinline__ f3 mfr_f(const Mfr &m, f3 wo, f3 wi) {
        const float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
        f3 h = add(wo, wi);
        if (co == 0.0f || ci == 0.0f) return mk1(0.0f);
        if (h.x == 0.0f && h.y == 0.0f && h.z == 0.0f) return mk1(0.0f);
        h = normalized(h);
        const f3 F = fresnel_conductor(dot(wi, h), m.eta, m.k);
        const float D = tr_D(m.d, h);
        const float G = tr_G(m.d, wo, wi);
        return divs(cmul(mul(mul(m.R, D), G), F), 4.0f * ci * co);
}
__device__ __forceinline__ float mfr_pdf(const Mfr &m, f3 wo, f3 wi) {
        if (!same_hemisphere(wo, wi)) return 0.0f;
        const f3 h = normalized(add(wo, wi));
        return tr_pdf(m.d, wo, h) / (4.0f * dot(wo, h));
}
__device__ __forceinline__ BSample mfr_sample(const Mfr &m, f3 wo, float u0, float u1) {
        if (wo.z == 0.0f) return bs_none();
        const f3 h = tr_sample_h(m.d, wo, u0, u1);
        const f3 wi = reflect_v(wo, h);
        return BSample{mfr_f(m, wo, wi), wi, mfr_pdf(m, wo, wi)};
}

// ---- DielectricBsdf ---------------------------------------------------------------------------
struct Diel {
        float eta;
        Trd d;
};
__device__ __forceinline__ bool diel_specular(const Diel &b) { return tr_smooth(b.d); }

__device__ __forceinline__ f3 diel_f(const Diel &b, f3 wo, f3 wi) {
        if (b.eta == 1.0f || tr_smooth(b.d)) return mk1(0.0f);
        const float co = cos_theta(wo), ci = cos_theta(wi);
        const bool reflect = ci * co > 0.0f;
        float etap = 1.0f;
        if (!reflect) etap = co > 0.0f ? b.eta : (1.0f / b.eta);
        f3 h = add(mul(wi, etap), wo);
        if (ci == 0.0f || co == 0.0f || len2(h) == 0.0f) return mk1(0.0f);
        h = normalized(h);
        if (h.x * 0.0f + h.y * 0.0f + h.z * 1.0f < 0.0f) h = neg(h);
        if (dot(h, wi) * cos_theta(wi) < 0.0f || dot(h, wo) * cos_theta(wo) < 0.0f) return mk1(0.0f);
        const float F = fresnel_dielectric(dot(wo, h), b.eta);
        if (reflect) {
                const float D = tr_D(b.d, h);
                const float G = tr_G(b.d, wo, wi);
                const float num = D * G * F;
                const float den = fabsf(4.0f * ci * co);
                return mk1(num / den);
        }
        const float den = sq(dot(wi, h) + dot(wo, h) / etap) * abs_cos_theta(wi) * abs_cos_theta(wo);
        const float D = tr_D(b.d, h);
        const float G = tr_G(b.d, wo, wi);
        const float Ft = 1.0f - F;
        const float adio = fabsf(dot(wi, h)) * fabsf(dot(wo, h));
        float num = D * Ft * G * adio;
        num /= sq(etap);
        return mk1(num / den);
}

__device__ __forceinline__ BSample diel_sample(const Diel &b, f3 wo, float u0, float u1, float u2, bool radiance) {
        if (b.eta == 1.0f || tr_smooth(b.d)) {
                const float R = fresnel_dielectric(cos_theta(wo), b.eta);
                const float T = 1.0f - R;
                const float pR = R / (R + T);
                const float pT = T / (R + T);
                if (u0 < pR) {
                        const f3 wi = mk(-wo.x, -wo.y, wo.z);
                        return BSample{mk1(R / abs_cos_theta(wi)), wi, pR};
                }
                f3 wi;
                float etap;
                if (!refract_v(wo, mk(0.0f, 0.0f, 1.0f), b.eta, wi, etap)) return bs_none();
                f3 f = mk1(T / abs_cos_theta(wi));
                if (radiance) f = divs(f, sq(etap));
                return BSample{f, wi, pT};
        }
        const f3 h = tr_sample_h(b.d, wo, u0, u1);
        const float R = fresnel_dielectric(dot(wo, h), b.eta);
        const float T = 1.0f - R;
        const float pR = R / (R + T);
        const float pT = T / (R + T);
        if (u2 < pR) {
                const f3 wi = reflect_v(wo, h);
                if (!same_hemisphere(wo, wi)) return bs_none();
                const float pdf = tr_pdf(b.d, wo, h) / (4.0f * absdot(wo, h)) * pR;
                const float D = tr_D(b.d, h);
                const float G = tr_G(b.d, wo, wi);
                const float num = D * G * R;
                const float den = 4.0f * cos_theta(wi) * cos_theta(wo);
                return BSample{mk1(num / den), wi, pdf};
        }
        f3 wi;
        float etap;
        if (!refract_v(wo, h, b.eta, wi, etap)) return bs_none();
        if (same_hemisphere(wo, wi) || wi.z == 0.0f) return bs_none();
        const float denom = sq(dot(wi, h) + dot(wo, h) / etap);
        const float dwm = absdot(wi, h) / denom;
        const float pdf = tr_pdf(b.d, wo, h) * dwm * pT;
        const float D = tr_D(b.d, h);
        const float G = tr_G(b.d, wo, wi);
        f3 f = mk1(T * D * G * fabsf(dot(wi, h) * dot(wo, h) / (abs_cos_theta(wi) * abs_cos_theta(wo) * denom)));
        if (radiance) f = divs(f, sq(etap));
        return BSample{f, wi, pdf};
}

__device__ __forceinline__ float diel_pdf(const Diel &b, f3 wo, f3 wi) {
        if (b.eta == 1.0f || diel_specular(b)) return 0.0f;
        const float co = cos_theta(wo), ci = cos_theta(wi);
        const bool reflect = ci * co > 0.0f;
        float etap = 1.0f;
        if (!reflect) etap = co > 0.0f ? b.eta : 1.0f / b.eta;
        f3 h = add(mul(wi, etap), wo);
        if (co == 0.0f || ci == 0.0f || len2(h) == 0.0f) return 0.0f;
        h = normalized(h);
        if (h.x * 0.0f + h.y * 0.0f + h.z * 1.0f < 0.0f) h = neg(h);
        if (dot(h, wi) * ci < 0.0f || dot(h, wo) * co < 0.0f) return 0.0f;
        const float R = fresnel_dielectric(dot(wo, h), b.eta);
        const float T = 1.0f - R;
        if (reflect) {
                const float dp = tr_pdf(b.d, wo, h);
                const float four = 4.0f * absdot(wo, h);
                const float p = R / (R + T);
                return dp / four * p;
        }
        const float denom = sq(dot(wi, h) + dot(wo, h) / etap);
        const float dwm = absdot(wi, h) / denom;
        return tr_pdf(b.d, wo, h) * dwm * T / (R + T);
}

// ---- Henyey-Greenstein ------------------------------------------------------------------------
__device__ __forceinline__ float hg_phase(float cosTheta, float g) {
        const float denom = 1.0f + g * g + 2.0f * g * cosTheta;
        return (1.0f / (4.0f * kPi)) * (1.0f - g * g) / (denom * sqrtf(denom));
}
__device__ __forceinline__ float hg_sample(float g, f3 wo, LayerRng &r, f3 &wi) {
        const float u0 = rng_next(r), u1 = rng_next(r);
        float cosTheta;
        if (fabsf(g) < 1e-3f) {
                cosTheta = 1.0f - 2.0f * u0;
        } else {
                const float sqrTerm = (1.0f - g * g) / (1.0f + g - 2.0f * g * u0);
                cosTheta = -(1.0f + g * g - sqrTerm * sqrTerm) / (2.0f * g);
        }
        const float sinTheta = sqrtf(smax(0.0f, 1.0f - cosTheta * cosTheta));
        const float phi = 2.0f * kPi * u1;
        f3 e1, e2;
        coordinate_system(wo, e1, e2);
        const f3 vx = mul(e1, sinTheta * cosf(phi));
        const f3 vy = mul(e2, sinTheta * sinf(phi));
        const f3 vz = mul(wo, cosTheta);
        wi = add(add(vx, vy), vz);
        return hg_phase(cosTheta, g);
}

// ---- the material's BSDF --------------------------------------------------------------------
struct Bx {
        int type;
        f3 R;           // diffuse reflectance; conductor: white; coated: the conductor's reflectance
        Trd d;          // conductor / dielectric / coating alpha
        f3 eta, k;      // conductor
        float ieta;     // dielectric or coating index
        Trd cd;         // coated conductor: the conductor's alpha
        float thickness;
        f3 albedo;
        float g;
        int maxDepth, nSamples;
};

__device__ __forceinline__ float remap_r(float r, bool on) { return on ? sqrtf(r) : r; }

// Conductor.swift:3-17, Dielectric.swift:44-54, CoatedConductor.swift:31-75
__device__ __forceinline__ Bx bx_from_material(const GzMaterial &m) {
        Bx b;
        b.type = m.type;
        b.R = mk(m.reflectance[0], m.reflectance[1], m.reflectance[2]);
        b.d = Trd{0.0f, 0.0f};
        b.eta = mk(m.conductorEta[0], m.conductorEta[1], m.conductorEta[2]);
        b.k = mk(m.conductorK[0], m.conductorK[1], m.conductorK[2]);
        b.ieta = m.eta;
        b.cd = Trd{0.0f, 0.0f};
        b.thickness = 0.0f;
        b.albedo = mk(m.albedo[0], m.albedo[1], m.albedo[2]);
        b.g = m.g;
        b.maxDepth = m.maxDepth;
        b.nSamples = m.nSamples;
        const bool rm = (m.flags & GZ_MAT_REMAP_ROUGHNESS) != 0;
        if (m.type == GZ_MAT_CONDUCTOR) {
                b.R = mk1(1.0f);
                b.d = Trd{smax(m.roughness[0], 1e-3f), smax(m.roughness[1], 1e-3f)};
        } else if (m.type == GZ_MAT_DIELECTRIC) {
                b.d = Trd{remap_r(m.roughness[0], rm), remap_r(m.roughness[1], rm)};
        } else if (m.type == GZ_MAT_COATED_CONDUCTOR) {
                b.d = Trd{smax(remap_r(m.roughness[0], rm), 0.001f), smax(remap_r(m.roughness[1], rm), 0.001f)};
                b.cd = Trd{smax(remap_r(m.conductorRoughness[0], rm), 0.001f),
                           smax(remap_r(m.conductorRoughness[1], rm), 0.001f)};
                b.thickness = smax(m.thickness, 1.17549435e-38f);
        }
        return b;
}

__device__ __forceinline__ Mfr bx_conductor(const Bx &b) { return Mfr{b.R, b.d, b.eta, b.k}; }
__device__ __forceinline__ Diel bx_dielectric(const Bx &b) { return Diel{b.ieta, b.d}; }
__device__ __forceinline__ Mfr bx_bottom(const Bx &b) { return Mfr{b.R, b.cd, b.eta, b.k}; }

// ---- LayeredBsdf<DielectricBsdf, MicrofacetReflection>, two-sided -------------------------------
// Interfaces: bottom (z == 0) is the conductor, top (z == thickness) the coating.
// The three interface calls stay out of line: the random walk calls each from several places, and
// inlined copies of the dielectric and conductor lobes at every call held the layered shade kernel
// at 256 VGPRs with spills, 2 waves per SIMD. Out of line, the kernel is built for 3 waves (168
// VGPRs, the rest spilled around the calls), which runs the layered class 6% faster.
__device__ __noinline__ f3 lay_if_f(const Bx &b, bool bottom, f3 wo, f3 wi) {
        return bottom ? mfr_f(bx_bottom(b), wo, wi) : diel_f(bx_dielectric(b), wo, wi);
}
__device__ __noinline__ float lay_if_pdf(const Bx &b, bool bottom, f3 wo, f3 wi) {
        return bottom ? mfr_pdf(bx_bottom(b), wo, wi) : diel_pdf(bx_dielectric(b), wo, wi);
}
__device__ __noinline__ BSample lay_if_sample(const Bx &b, bool bottom, f3 wo, float u0, float u1, float u2,
                                               bool radiance) {
        return bottom ? mfr_sample(bx_bottom(b), wo, u0, u1) : diel_sample(bx_dielectric(b), wo, u0, u1, u2, radiance);
}
__device__ __forceinline__ BSample lay_if_sample_rng(const Bx &b, bool bottom, f3 wo, LayerRng &r, bool radiance) {
        const float u0 = rng_next(r), u1 = rng_next(r), u2 = rng_next(r);
        return lay_if_sample(b, bottom, wo, u0, u1, u2, radiance);
}
__device__ __forceinline__ f3 lay_tr(float dz, f3 w) {
        if (fabsf(dz) < 1.17549435e-38f) return mk1(1.0f);
        const float val = fabsf(dz / w.z);
        return mk1(expf(-val));
}

__device__ __noinline__ f3 lay_f(const Bx &b, f3 wo, f3 wi, LayerRng &r) {
        const bool topSpec = diel_specular(bx_dielectric(b)), botSpec = false;
        f3 result = mk1(0.0f);
        if (wo.z < 0.0f) {
                wo = neg(wo);
                wi = neg(wi);
        }
        const bool enteredTop = true;
        const bool same = same_hemisphere(wo, wi);
        const float exitZ = (same != enteredTop) ? 0.0f : b.thickness;
        if (same) result = cmul(mk1((float)b.nSamples), lay_if_f(b, false, wo, wi));
        for (int s = 0; s < b.nSamples; ++s) {
                const BSample wos = lay_if_sample_rng(b, false, wo, r, true);
                if (!bs_valid(wos) || bs_reflection(wos, wo) || wos.wi.z == 0.0f) continue;
                const BSample wis = lay_if_sample_rng(b, same != enteredTop, wi, r, false);
                if (!bs_valid(wis) || bs_reflection(wis, wi) || wis.wi.z == 0.0f) continue;
                f3 beta = divs(mul(wos.f, abs_cos_theta(wos.wi)), wos.pdf);
                float z = enteredTop ? b.thickness : 0.0f;
                f3 w = wos.wi;
                for (int depth = 0; depth < b.maxDepth; ++depth) {
                        if (depth > 3 && cmax(beta) < 0.25f) {
                                const float q = smax(0.0f, 1.0f - cmax(beta));
                                if (rng_next(r) < q) break;
                                beta = divs(beta, 1.0f - q);
                        }
                        if (cblack(b.albedo)) {
                                z = (z == b.thickness) ? 0.0f : b.thickness;
                                beta = cmul(beta, lay_tr(b.thickness, w));
                        } else {
                                const float sigmaT = 1.0f;
                                const float dz = -logf(1.0f - rng_next(r)) / (sigmaT / fabsf(w.z));
                                const float zp = w.z > 0.0f ? (z + dz) : (z - dz);
                                if (z == zp) continue;
                                if (zp > 0.0f && zp < b.thickness) {
                                        float misW = 1.0f;
                                        const float ph = hg_phase(dot(neg(w), neg(wis.wi)), b.g);
                                        const bool exitSpec = (exitZ == 0.0f) ? botSpec : topSpec;
                                        if (!exitSpec) misW = power_heuristic(wis.pdf, ph);
                                        const f3 t = cmul(mul(mul(cmul(beta, b.albedo), ph), misW), lay_tr(zp - exitZ, wis.wi));
                                        result = cadd(result, divs(cmul(t, wis.f), wis.pdf));
                                        f3 nw;
                                        const float ppdf = hg_sample(b.g, neg(w), r, nw);
                                        if (ppdf == 0.0f || nw.z == 0.0f) continue;
                                        beta = cmul(beta, divs(mul(b.albedo, ppdf), ppdf));
                                        w = nw;
                                        z = zp;
                                        continue;
                                }
                                z = sclamp(zp, 0.0f, b.thickness);
                        }
                        if (z == exitZ) {
                                const BSample bs = lay_if_sample_rng(b, z == 0.0f, neg(w), r, true);
                                if (!bs_valid(bs) || bs.pdf == 0.0f || bs.wi.z == 0.0f) break;
                                if (!bs_reflection(bs, neg(w))) break;
                                beta = cmul(beta, divs(mul(bs.f, abs_cos_theta(bs.wi)), bs.pdf));
                                w = bs.wi;
                        } else {
                                const bool nonExitSpec = (z == 0.0f) ? botSpec : topSpec;
                                const bool exitSpec = (exitZ == 0.0f) ? botSpec : topSpec;
                                if (!nonExitSpec) {
                                        float misW = 1.0f;
                                        if (!exitSpec) misW = power_heuristic(wis.pdf, lay_if_pdf(b, z == 0.0f, neg(w), neg(wis.wi)));
                                        const f3 ne = lay_if_f(b, z == 0.0f, neg(w), neg(wis.wi));
                                        const f3 t = cmul(mul(mul(cmul(beta, ne), abs_cos_theta(wis.wi)), misW),
                                                          lay_tr(b.thickness, wis.wi));
                                        result = cadd(result, divs(cmul(t, wis.f), wis.pdf));
                                }
                                const BSample bs = lay_if_sample_rng(b, z == 0.0f, neg(w), r, true);
                                if (!bs_valid(bs) || bs.pdf == 0.0f || bs.wi.z == 0.0f) break;
                                if (!bs_reflection(bs, neg(w))) break;
                                beta = cmul(beta, divs(mul(bs.f, abs_cos_theta(bs.wi)), bs.pdf));
                                w = bs.wi;
                                if (!exitSpec) {
                                        const f3 fExit = lay_if_f(b, exitZ == 0.0f, neg(w), wi);
                                        if (!cblack(fExit)) {
                                                float misW = 1.0f;
                                                if (!nonExitSpec)
                                                        misW = power_heuristic(bs.pdf, lay_if_pdf(b, exitZ == 0.0f, neg(w), wi));
                                                result = cadd(result, mul(cmul(cmul(beta, lay_tr(b.thickness, bs.wi)), fExit), misW));
                                        }
                                }
                        }
                }
        }
        return divs(result, (float)b.nSamples);
}

__device__ __noinline__ BSample lay_sample(const Bx &b, f3 wo, float u0, float u1, float u2, LayerRng &r) {
        bool flip = false;
        if (wo.z < 0.0f) {
                wo = neg(wo);
                flip = true;
        }
        BSample bs = lay_if_sample(b, false, wo, u0, u1, u2, true);
        if (!bs_valid(bs) || bs.pdf == 0.0f || bs.wi.z == 0.0f) return bs_none();
        if (bs_reflection(bs, wo)) {
                if (flip) bs.wi = neg(bs.wi);
                return bs;
        }
        f3 w = bs.wi;
        f3 f = mul(bs.f, abs_cos_theta(bs.wi));
        float pdf = bs.pdf;
        float z = b.thickness;
        for (int depth = 0; depth < b.maxDepth; ++depth) {
                const float rrBeta = cmax(f) / pdf;
                if (depth > 3 && rrBeta < 0.25f) {
                        const float q = smax(0.0f, 1.0f - rrBeta);
                        if (rng_next(r) < q) return bs_none();
                        pdf *= 1.0f - q;
                }
                if (w.z == 0.0f) return bs_none();
                if (!cblack(b.albedo)) {
                        const float sigmaT = 1.0f;
                        const float dz = -logf(1.0f - rng_next(r)) / (sigmaT / abs_cos_theta(w));
                        const float zp = w.z > 0.0f ? (z + dz) : (z - dz);
                        if (zp == z) return bs_none();
                        if (zp > 0.0f && zp < b.thickness) {
                                f3 nw;
                                const float ph = hg_sample(b.g, neg(w), r, nw);
                                if (ph == 0.0f || nw.z == 0.0f) return bs_none();
                                f = cmul(f, mul(b.albedo, ph));
                                pdf *= ph;
                                w = nw;
                                z = zp;
                                continue;
                        }
                        z = sclamp(zp, 0.0f, b.thickness);
                } else {
                        z = (z == b.thickness) ? 0.0f : b.thickness;
                        f = cmul(f, mk1(1.0f));
                }
                const BSample s = lay_if_sample_rng(b, z == 0.0f, neg(w), r, true);
                if (!bs_valid(s) || s.pdf == 0.0f || s.wi.z == 0.0f) return bs_none();
                f = cmul(f, s.f);
                pdf *= s.pdf;
                w = s.wi;
                // The reference tests for transmission against the updated direction
                // (LayeredBsdf.swift:386-391), which is always true: the walk ends here.
                if (!bs_reflection(s, neg(w))) return BSample{f, flip ? neg(w) : w, pdf};
                f = mul(f, abs_cos_theta(s.wi));
        }
        return bs_none();
}

__device__ __noinline__ float lay_pdf(const Bx &b, f3 wo, f3 wi, LayerRng &r) {
        const bool topSpec = diel_specular(bx_dielectric(b)), botSpec = false;
        if (wo.z < 0.0f) {
                wo = neg(wo);
                wi = neg(wi);
        }
        float pdfSum = 0.0f;
        if (same_hemisphere(wo, wi)) pdfSum += (float)b.nSamples * diel_pdf(bx_dielectric(b), wo, wi);
        for (int s = 0; s < b.nSamples; ++s) {
                if (same_hemisphere(wo, wi)) {
                        const BSample os = lay_if_sample_rng(b, false, wo, r, true);
                        const BSample is = lay_if_sample_rng(b, false, wi, r, true);
                        if (bs_valid(os) && !bs_reflection(os, wo) && bs_valid(is) && !bs_reflection(is, wi)) {
                                if (topSpec) {
                                        pdfSum += mfr_pdf(bx_bottom(b), neg(os.wi), neg(is.wi));
                                } else {
                                        const BSample rs = lay_if_sample_rng(b, true, neg(os.wi), r, true);
                                        if (bs_valid(rs) && rs.pdf > 0.0f) {
                                                if (botSpec) {
                                                        pdfSum += diel_pdf(bx_dielectric(b), neg(rs.wi), wi);
                                                } else {
                                                        const float rPDF = mfr_pdf(bx_bottom(b), neg(os.wi), neg(is.wi));
                                                        float w = power_heuristic(is.pdf, rPDF);
                                                        pdfSum += w * rPDF;
                                                        const float tPDF = diel_pdf(bx_dielectric(b), neg(rs.wi), wi);
                                                        w = power_heuristic(rs.pdf, tPDF);
                                                        pdfSum += w * tPDF;
                                                }
                                        }
                                }
                        }
                } else {
                        const BSample os = lay_if_sample_rng(b, false, wo, r, true);
                        const BSample is = lay_if_sample_rng(b, true, wi, r, true);
                        if (bs_valid(os) && !bs_reflection(os, wo) && bs_valid(is) && !bs_reflection(is, wi)) {
                                const float p1 = diel_pdf(bx_dielectric(b), wo, neg(is.wi));
                                const float p2 = mfr_pdf(bx_bottom(b), neg(os.wi), wi);
                                pdfSum += (p1 + p2) / 2.0f;
                        }
                }
        }
        const float t = 0.9f;
        return (1.0f - t) * (1.0f / (4.0f * kPi)) + t * (pdfSum / (float)b.nSamples);
}

// ---- dispatch -----------------------------------------------------------------------------------
// The BSDF set a kernel is compiled with (its kAll parameter): kBsdfDiffuse only the diffuse BSDF,
// for scenes or class queues without other materials; kBsdfAll every BSDF; kBsdfNoLayer every
// BSDF but the layered one, whose noinline random walk sets the register count of any kernel that
// can call it (256 against 143 for the wavefront's shade kernel). A kernel only sees materials of
// its set; the others evaluate to black.
enum : int { kBsdfDiffuse = 0, kBsdfAll = 1, kBsdfNoLayer = 2 };
template <int kAll>
__device__ __forceinline__ f3 bx_f(const Bx &b, f3 wo, f3 wi, LayerRng &r) {
        if (!kAll || b.type == GZ_MAT_DIFFUSE) return divs(b.R, kPi);
        switch (b.type) {
        case GZ_MAT_CONDUCTOR: return mfr_f(bx_conductor(b), wo, wi);
        case GZ_MAT_DIELECTRIC: return diel_f(bx_dielectric(b), wo, wi);
        case GZ_MAT_COATED_CONDUCTOR:
                if constexpr (kAll == kBsdfAll) return lay_f(b, wo, wi, r);
                break;
        }
        return mk1(0.0f);
}
template <int kAll>
__device__ __forceinline__ float bx_pdf(const Bx &b, f3 wo, f3 wi, LayerRng &r) {
        if (!kAll || b.type == GZ_MAT_DIFFUSE) {
                if (!same_hemisphere(wo, wi)) return 0.0f;
                return abs_cos_theta(wi) / kPi;
        }
        switch (b.type) {
        case GZ_MAT_CONDUCTOR: return mfr_pdf(bx_conductor(b), wo, wi);
        case GZ_MAT_DIELECTRIC: return diel_pdf(bx_dielectric(b), wo, wi);
        case GZ_MAT_COATED_CONDUCTOR:
                if constexpr (kAll == kBsdfAll) return lay_pdf(b, wo, wi, r);
                break;
        }
        return 0.0f;
}
template <int kAll>
__device__ __forceinline__ BSample bx_sample(const Bx &b, f3 wo, float u0, float u1, float u2, LayerRng &r) {
        if (!kAll || b.type == GZ_MAT_DIFFUSE) {
                f3 wi = cosine_hemisphere(u0, u1);
                if (wo.z < 0.0f) wi.z = -wi.z;
                return BSample{divs(b.R, kPi), wi, same_hemisphere(wo, wi) ? abs_cos_theta(wi) / kPi : 0.0f};
        }
        switch (b.type) {
        case GZ_MAT_CONDUCTOR: return mfr_sample(bx_conductor(b), wo, u0, u1);
        case GZ_MAT_DIELECTRIC: return diel_sample(bx_dielectric(b), wo, u0, u1, u2, true);
        case GZ_MAT_COATED_CONDUCTOR:
                if constexpr (kAll == kBsdfAll) return lay_sample(b, wo, u0, u1, u2, r);
                break;
        }
        return bs_none();
}
template <int kAll>
__device__ __forceinline__ bool bx_specular(const Bx &b) {
        return kAll && b.type == GZ_MAT_DIELECTRIC && tr_smooth(b.d);
}
template <int kAll>
__device__ __forceinline__ f3 bx_albedo(const Bx &b) { return (!kAll || b.type == GZ_MAT_DIFFUSE) ? b.R : mk1(1.0f); }

}  // namespace gz
