/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's camera-sample, primary
 * ray and film steps around the path (the checker of the GPU camera and film kernels); never
 * shipped, never on the product path.
 *
 *   Tile.render phase 1: startPixelSample, getCameraSample, generateRay, filter weight
 *                                            Core/Tile.swift:27-59, Sampler/Sampler.swift:69-83
 *   GaussianFilter evaluate / sample / erfinv Filter/GaussianFilter.swift:7-121
 *   PerspectiveCamera.generateRay (pinhole) Camera/PerspectiveCamera.swift:37-48
 *   Transform * Point / Vector               Geometry/Transform.swift:31-65
 *   Film.add(sample:) into beauty, albedo, normal images; Pixel sums and normalisation
 *                                            Core/Film.swift:107-136, Image/Image.swift:21-30,
 *                                            Core/Pixel.swift:12-39
 * The camera matrices (rasterToCamera, cameraToWorld) are inputs, built by the caller. erff,
 * expf, logf come from the C library (the reference uses Glibc's).
 *
 * Pinning (tests/test_oracle_film.py): erfinv(erf(x)) = x, the filter's samples follow the
 * truncated Gaussian, its weight evaluate/pdf = (2 norm - 1)^2 inside the support, pinhole
 * geometry (the centre ray is the view direction, edge rays make fov/2 with it), and the film
 * sums equal a float64 accumulation.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float ox, oy, oz, dx, dy, dz; } fm_ray;
typedef struct {
        fm_ray ray;
        float throughput[3], estimate[3], albedo[3], firstNormal[3];
        float filterWeight;
        int32_t pixelX, pixelY, sampleIndex, dim, lightDim, bounce, flags, pad0, pad1;
} fm_path;
_Static_assert(sizeof(fm_path) == 112, "GzPath");

/* GzCamera (include/gonzales_amd.h) */
typedef struct {
        float rasterToCamera[16];   /* row-major 4x4 */
        float cameraToWorld[16];
        int32_t resX, resY;         /* film resolution */
        int32_t x0, y0, x1, y1;     /* pixel bounds rendered (crop), [x0, x1) x [y0, y1) */
        int32_t spp;
        int32_t filterType;         /* 0 Gaussian, 1 box */
        float filterRadius[2];
        float filterSigma;
        float lensRadius, focalDistance;
        int32_t pad;
        int64_t zsSeed;
        int32_t zsLog2Spp, zsBase4Digits;
} fm_camera;
_Static_assert(sizeof(fm_camera) == 200, "GzCamera");

typedef struct { int64_t seed; int32_t log2spp, nBase4Digits; uint64_t morton; int32_t dim, pad; } fm_zs;
void orc_zsobol_get2d(fm_zs *z, const uint32_t *m, float *out);
uint64_t orc_encode_morton2(uint32_t x, uint32_t y);

static const float PI_F = 3.14159265358979323846f;
static inline float sq(float x) { return x * x; }

/* ---- Gaussian filter ---------------------------------------------------------------------- */
static float gaussian(float x, float sigma) {
        return 1.0f / sqrtf(2.0f * PI_F * sigma * sigma) * expf(-sq(x - 0.0f) / (2.0f * sigma * sigma));
}
static float smax(float x, float y) { return y >= x ? y : x; }
static float smin(float x, float y) { return y < x ? y : x; }

float orc_gaussian_evaluate(float px, float py, float sigma) {
        float x = smax(0.0f, gaussian(px, sigma));
        float y = smax(0.0f, gaussian(py, sigma));
        return x * y;
}

static float newton(float x, float y) {
        const float twoOverSqrtPi = 2.0f / sqrtf(PI_F);
        return (erff(x) - y) / (twoOverSqrtPi * expf(-(x * x)));
}

float orc_erfinv(float y) {
        const float a0 = 0.886226899f, a1 = -1.645349621f, a2 = 0.914624893f, a3 = -0.140543331f;
        const float b0 = -2.118377725f, b1 = 1.442710462f, b2 = -0.329097515f, b3 = 0.012229801f;
        const float c0 = -1.970840454f, c1 = -1.624906493f, c2 = 3.429567803f, c3 = 1.641345311f;
        const float d0 = 3.543889200f, d1 = 1.637067800f;
        float absY = fabsf(y);
        if (absY <= 0.7f) {
                float z = y * y;
                float np = a3 * z + a2;
                np = np * z + a1;
                float num = np * z + a0;
                float dp = b3 * z + b2;
                dp = dp * z + b1;
                dp = dp * z + b0;
                float den = dp * z + 1.0f;
                float x = y * num / den;
                x -= newton(x, y);
                x -= newton(x, y);
                return x;
        } else if (absY < 1.0f) {
                float zBase = (1.0f - absY) / 2.0f;
                float z = sqrtf(-logf(zBase));
                float np = c3 * z + c2;
                np = np * z + c1;
                float num = np * z + c0;
                float dp = d1 * z + d0;
                float den = dp * z + 1.0f;
                float signY = signbit(y) ? -1.0f : 1.0f;
                float x = signY * num / den;
                x -= newton(x, y);
                x -= newton(x, y);
                return x;
        } else if (absY == 1.0f) {
                return signbit(y) ? -INFINITY : INFINITY;
        }
        return NAN;
}

static void sample1d(float u, float radius, float sigma, float *xs, float *density) {
        float norm = 0.5f * (1.0f + erff(radius / (sigma * sqrtf(2.0f))));
        float uScaled = (1.0f - norm) + u * (2.0f * norm - 1.0f);
        float x = sigma * sqrtf(2.0f) * orc_erfinv(2.0f * uScaled - 1.0f);
        float pdfU = expf(-x * x / (2.0f * sigma * sigma)) / (sigma * sqrtf(2.0f * PI_F));
        *density = pdfU / (2.0f * norm - 1.0f);
        *xs = smax(-radius, smin(radius, x));
}

/* out: location x, y, probability density */
void orc_gaussian_sample(float u0, float u1, float rx, float ry, float sigma, float *out) {
        float dx, px, dy, py;
        sample1d(u0, rx, sigma, &dx, &px);
        sample1d(u1, ry, sigma, &dy, &py);
        out[0] = dx;
        out[1] = dy;
        out[2] = px * py;
}

/* BoxFilter (Filter/BoxFilter.swift:3-14): evaluate 1; location (2u - 1) r, pdf 1 */
static void filter_sample(const fm_camera *c, float u0, float u1, float *dx, float *dy, float *pdf) {
        if (c->filterType == 1) {
                *dx = (2.0f * u0 - 1.0f) * c->filterRadius[0];
                *dy = (2.0f * u1 - 1.0f) * c->filterRadius[1];
                *pdf = 1.0f;
                return;
        }
        float o[3];
        orc_gaussian_sample(u0, u1, c->filterRadius[0], c->filterRadius[1], c->filterSigma, o);
        *dx = o[0];
        *dy = o[1];
        *pdf = o[2];
}
static float filter_eval(const fm_camera *c, float x, float y) {
        if (c->filterType == 1) return 1.0f;
        return orc_gaussian_evaluate(x, y, c->filterSigma);
}

/* ---- camera ------------------------------------------------------------------------------- */
static void xform_point(const float *m, float x, float y, float z, float *o) {
        float px = m[0] * x + m[1] * y + m[2] * z + m[3];
        float py = m[4] * x + m[5] * y + m[6] * z + m[7];
        float pz = m[8] * x + m[9] * y + m[10] * z + m[11];
        float w = m[12] * x + m[13] * y + m[14] * z + m[15];
        o[0] = px / w;
        o[1] = py / w;
        o[2] = pz / w;
}
static void xform_vector(const float *m, float x, float y, float z, float *o) {
        o[0] = m[0] * x + m[1] * y + m[2] * z;
        o[1] = m[4] * x + m[5] * y + m[6] * z;
        o[2] = m[8] * x + m[9] * y + m[10] * z;
}

static void generate_ray(const fm_camera *c, float fx, float fy, fm_ray *r) {
        float cp[3];
        xform_point(c->rasterToCamera, fx, fy, 0.0f, cp);
        float l = sqrtf(cp[0] * cp[0] + cp[1] * cp[1] + cp[2] * cp[2]);
        float d[3] = {cp[0], cp[1], cp[2]};
        if (l != 0.0f) {
                d[0] = cp[0] / l;
                d[1] = cp[1] / l;
                d[2] = cp[2] / l;
        }
        float o[3], wd[3];
        xform_point(c->cameraToWorld, 0.0f, 0.0f, 0.0f, o);
        xform_vector(c->cameraToWorld, d[0], d[1], d[2], wd);
        r->ox = o[0]; r->oy = o[1]; r->oz = o[2];
        r->dx = wd[0]; r->dy = wd[1]; r->dz = wd[2];
}

/* Paths first..first+count-1 of the crop, pixel-major (index = pixel * spp + sample). */
int orc_camera_paths(const fm_camera *c, const uint32_t *table, int64_t first, int64_t count, fm_path *out) {
        if (c->lensRadius > 0.0f) return -1; /* depth of field is not restated */
        const int64_t w = c->x1 - c->x0;
        for (int64_t k = 0; k < count; ++k) {
                int64_t i = first + k;
                int64_t pix = i / c->spp;
                int32_t s = (int32_t)(i % c->spp);
                int32_t px = c->x0 + (int32_t)(pix % w), py = c->y0 + (int32_t)(pix / w);
                fm_zs z;
                z.seed = c->zsSeed;
                z.log2spp = c->zsLog2Spp;
                z.nBase4Digits = c->zsBase4Digits;
                uint64_t m = orc_encode_morton2((uint32_t)px, (uint32_t)py);
                z.morton = (z.log2spp >= 64 ? 0 : m << z.log2spp) | (uint64_t)(uint32_t)s;
                z.dim = 0;
                float u[2], lens[2];
                orc_zsobol_get2d(&z, table, u);
                float dx, dy, fpdf;
                filter_sample(c, u[0], u[1], &dx, &dy, &fpdf);
                float filmX = (float)px + 0.5f + dx, filmY = (float)py + 0.5f + dy;
                orc_zsobol_get2d(&z, table, lens);
                fm_path *p = &out[k];
                memset(p, 0, sizeof *p);
                generate_ray(c, filmX, filmY, &p->ray);
                float deltaX = filmX - ((float)px + 0.5f);
                float deltaY = filmY - ((float)py + 0.5f);
                float fv = filter_eval(c, deltaX, deltaY);
                p->filterWeight = fv / fpdf;
                p->throughput[0] = p->throughput[1] = p->throughput[2] = 1.0f;
                p->pixelX = px;
                p->pixelY = py;
                p->sampleIndex = s;
                p->dim = z.dim;
                p->lightDim = 0;
                p->bounce = 0;
                p->flags = 1;
        }
        return 0;
}

/* Film.add for the samples of paths first..first+count-1 (pixel-major, spp per pixel, whole
 * pixels), in sample order. Images are resX*resY RGB (light, albedo, normal) + weight. */
int orc_film_add(const fm_camera *c, const fm_path *paths, int64_t first, int64_t count, float iso,
                 float maxComponentValue, float *light, float *weight, float *albedo, float *normal) {
        for (int64_t k = 0; k < count; ++k) {
                const fm_path *p = &paths[k];
                int32_t x = p->pixelX, y = p->pixelY;
                if (!(x >= 0 && y >= 0 && x < c->resX && y < c->resY)) continue;
                int64_t idx = (int64_t)y * c->resX + x;
                float k100 = iso / 100.0f;
                float s[3] = {p->estimate[0] * k100, p->estimate[1] * k100, p->estimate[2] * k100};
                if (maxComponentValue > 0.0f) {
                        float gb = s[2] >= s[1] ? s[2] : s[1];
                        float mv = gb >= s[0] ? gb : s[0];
                        if (mv > maxComponentValue) {
                                float f = maxComponentValue / mv;
                                s[0] *= f; s[1] *= f; s[2] *= f;
                        }
                }
                for (int ch = 0; ch < 3; ++ch) light[3 * idx + ch] = light[3 * idx + ch] + s[ch];
                weight[idx] = weight[idx] + p->filterWeight;
                for (int ch = 0; ch < 3; ++ch) albedo[3 * idx + ch] = albedo[3 * idx + ch] + p->albedo[ch];
                const float *n = p->firstNormal;
                float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                float nn[3] = {n[0], n[1], n[2]};
                if (l != 0.0f) { nn[0] = n[0] / l; nn[1] = n[1] / l; nn[2] = n[2] / l; }
                for (int ch = 0; ch < 3; ++ch) normal[3 * idx + ch] = normal[3 * idx + ch] + fabsf(nn[ch]);
        }
        (void)first;
        return 0;
}

/* Pixel.normalized: black stays, otherwise light / weight; returns the number of pixels with
 * light but zero weight (the reference throws PixelError.zeroWeight). */
int64_t orc_film_resolve(const float *light, const float *weight, int64_t pixels, float *out) {
        int64_t bad = 0;
        for (int64_t i = 0; i < pixels; ++i) {
                const float *l = &light[3 * i];
                if (l[0] == 0.0f && l[1] == 0.0f && l[2] == 0.0f) {
                        out[3 * i] = l[0]; out[3 * i + 1] = l[1]; out[3 * i + 2] = l[2];
                        continue;
                }
                if (weight[i] == 0.0f) {
                        ++bad;
                        out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = NAN;
                        continue;
                }
                for (int ch = 0; ch < 3; ++ch) out[3 * i + ch] = l[ch] / weight[i];
        }
        return bad;
}
