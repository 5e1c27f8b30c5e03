// Camera samples, primary rays and film accumulation on gfx950: the steps on either side of
// the path (Core/Tile.swift:27-59 and :170-219, Core/Film.swift:107-136).
//
//   k_camera_paths: one lane per path, paths pixel-major over the crop (index = pixel * spp +
//     sample): startPixelSample (dim 0), getCameraSample (filter get2D, lens get2D;
//     Sampler.swift:69-83), the filter's sample and weight (GaussianFilter.swift:7-121,
//     BoxFilter.swift:3-14), PerspectiveCamera.generateRay as a pinhole
//     (PerspectiveCamera.swift:37-48, Transform.swift:31-65); writes fresh GzPath records.
//   k_film_add: one lane per pixel, adding its spp samples in sample order to the beauty
//     (estimate x iso/100, clamped to maxComponentValue when > 0), weight, albedo and
//     normal (|normalized n|) sums (Film.swift:107-136, Image.swift:21-30, Pixel.swift:33-39).
//   k_film_resolve: Pixel.normalized (Pixel.swift:12-16): black stays, else light / weight.
#include <cmath>

#include "bsdf_device.h"
#include "gz_internal.h"
#include "zsobol_device.h"

#pragma clang fp contract(off)

namespace gz {

const uint32_t *device_sobol_table();  // gz_zsobol.hip

constexpr int kFilmBlock = 256;

__device__ __forceinline__ float gaussian(float x, float sigma) {
        return 1.0f / sqrtf(2.0f * kPi * sigma * sigma) * expf(-sq(x - 0.0f) / (2.0f * sigma * sigma));
}
__device__ __forceinline__ float smin(float x, float y) { return y < x ? y : x; }

__device__ __forceinline__ float newton_step(float x, float y) {
        const float twoOverSqrtPi = 2.0f / sqrtf(kPi);
        return (erff(x) - y) / (twoOverSqrtPi * expf(-(x * x)));
}

// GaussianFilter.erfinv: rational first guess, two Newton steps on erf (GaussianFilter.swift:27-95)
__device__ float erfinv_ref(float y) {
        const float a0 = 0.886226899f, a1 = -1.645349621f, a2 = 0.914624893f, a3 = -0.140543331f;
        const float b0 = -2.118377725f, b1 = 1.442710462f, b2 = -0.329097515f, b3 = 0.012229801f;
        const float c0 = -1.970840454f, c1 = -1.624906493f, c2 = 3.429567803f, c3 = 1.641345311f;
        const float d0 = 3.543889200f, d1 = 1.637067800f;
        const float absY = fabsf(y);
        if (absY <= 0.7f) {
                const float z = y * y;
                float np = a3 * z + a2;
                np = np * z + a1;
                const float num = np * z + a0;
                float dp = b3 * z + b2;
                dp = dp * z + b1;
                dp = dp * z + b0;
                const float den = dp * z + 1.0f;
                float x = y * num / den;
                x -= newton_step(x, y);
                x -= newton_step(x, y);
                return x;
        } else if (absY < 1.0f) {
                const float zBase = (1.0f - absY) / 2.0f;
                const float z = sqrtf(-logf(zBase));
                float np = c3 * z + c2;
                np = np * z + c1;
                const float num = np * z + c0;
                const float dp = d1 * z + d0;
                const float den = dp * z + 1.0f;
                const float signY = signbit(y) ? -1.0f : 1.0f;
                float x = signY * num / den;
                x -= newton_step(x, y);
                x -= newton_step(x, y);
                return x;
        } else if (absY == 1.0f) {
                return signbit(y) ? -INFINITY : INFINITY;
        }
        return NAN;
}

__device__ __forceinline__ void sample1d(float u, float radius, float sigma, float &xs, float &density) {
        const float norm = 0.5f * (1.0f + erff(radius / (sigma * sqrtf(2.0f))));
        const float uScaled = (1.0f - norm) + u * (2.0f * norm - 1.0f);
        const float x = sigma * sqrtf(2.0f) * erfinv_ref(2.0f * uScaled - 1.0f);
        const float pdfU = expf(-x * x / (2.0f * sigma * sigma)) / (sigma * sqrtf(2.0f * kPi));
        density = pdfU / (2.0f * norm - 1.0f);
        xs = smax(-radius, smin(radius, x));
}

__device__ __forceinline__ void filter_sample(const GzCamera &c, float u0, float u1, float &dx, float &dy, float &pdf) {
        if (c.filterType == GZ_FILTER_BOX) {
                dx = (2.0f * u0 - 1.0f) * c.filterRadius[0];
                dy = (2.0f * u1 - 1.0f) * c.filterRadius[1];
                pdf = 1.0f;
                return;
        }
        float px, py;
        sample1d(u0, c.filterRadius[0], c.filterSigma, dx, px);
        sample1d(u1, c.filterRadius[1], c.filterSigma, dy, py);
        pdf = px * py;
}
__device__ __forceinline__ float filter_eval(const GzCamera &c, float x, float y) {
        if (c.filterType == GZ_FILTER_BOX) return 1.0f;
        const float gx = smax(0.0f, gaussian(x, c.filterSigma));
        const float gy = smax(0.0f, gaussian(y, c.filterSigma));
        return gx * gy;
}

__device__ __forceinline__ f3 xform_point(const float *m, float x, float y, float z) {
        const float px = m[0] * x + m[1] * y + m[2] * z + m[3];
        const float py = m[4] * x + m[5] * y + m[6] * z + m[7];
        const float pz = m[8] * x + m[9] * y + m[10] * z + m[11];
        const float w = m[12] * x + m[13] * y + m[14] * z + m[15];
        return mk(px / w, py / w, pz / w);
}
__device__ __forceinline__ f3 xform_vector(const float *m, f3 v) {
        return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
                  m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

__global__ __launch_bounds__(kFilmBlock) void k_camera_paths(GzCamera c, const uint32_t *__restrict__ table,
                                                             int64_t first, int64_t count, GzPath *__restrict__ out) {
        const int64_t k = (int64_t)blockIdx.x * kFilmBlock + threadIdx.x;
        if (k >= count) return;
        const int64_t i = first + k;
        const int64_t w = c.x1 - c.x0;
        const int64_t pix = i / c.spp;
        const int s = (int)(i % c.spp);
        const int px = c.x0 + (int)(pix % w), py = c.y0 + (int)(pix / w);
        const ZSobolParams zp{c.zsSeed, c.zsLog2Spp, c.zsBase4Digits};
        ZSobolState z = zs_start(zp, px, py, s, 0);
        float u0, u1, l0, l1;
        zs_get2d(zp, z, table, u0, u1);
        float dx, dy, fpdf;
        filter_sample(c, u0, u1, dx, dy, fpdf);
        const float filmX = (float)px + 0.5f + dx, filmY = (float)py + 0.5f + dy;
        zs_get2d(zp, z, table, l0, l1);
        const f3 cp = xform_point(c.rasterToCamera, filmX, filmY, 0.0f);
        const f3 d = normalized(cp);
        const f3 o = xform_point(c.cameraToWorld, 0.0f, 0.0f, 0.0f);
        const f3 wd = xform_vector(c.cameraToWorld, d);
        const float fv = filter_eval(c, filmX - ((float)px + 0.5f), filmY - ((float)py + 0.5f));
        GzPath p;
        p.ray = Ray_C{o.x, o.y, o.z, wd.x, wd.y, wd.z};
        for (int ch = 0; ch < 3; ++ch) {
                p.throughput[ch] = 1.0f;
                p.estimate[ch] = 0.0f;
                p.albedo[ch] = 0.0f;
                p.firstNormal[ch] = 0.0f;
        }
        p.filterWeight = fv / fpdf;
        p.pixelX = px;
        p.pixelY = py;
        p.sampleIndex = s;
        p.dim = z.dim;
        p.lightDim = 0;
        p.bounce = 0;
        p.flags = GZ_PATH_ACTIVE;
        p.pad[0] = p.pad[1] = 0;
        out[k] = p;
}

__global__ __launch_bounds__(kFilmBlock) void k_film_add(int resX, int resY, int spp, const GzPath *__restrict__ paths,
                                                         int64_t pixels, float iso, float maxComponent,
                                                         float *__restrict__ light, float *__restrict__ weight,
                                                         float *__restrict__ albedo, float *__restrict__ normal) {
        const int64_t j = (int64_t)blockIdx.x * kFilmBlock + threadIdx.x;
        if (j >= pixels) return;
        const GzPath *ps = paths + j * spp;
        for (int s = 0; s < spp; ++s) {
                const GzPath &p = ps[s];
                const int x = p.pixelX, y = p.pixelY;
                if (!(x >= 0 && y >= 0 && x < resX && y < resY)) continue;
                const int64_t idx = (int64_t)y * resX + x;
                const float k100 = iso / 100.0f;
                f3 sc = mk(p.estimate[0] * k100, p.estimate[1] * k100, p.estimate[2] * k100);
                if (maxComponent > 0.0f) {
                        const float mv = cmax(sc);
                        if (mv > maxComponent) sc = mul(sc, maxComponent / mv);
                }
                light[3 * idx] = light[3 * idx] + sc.x;
                light[3 * idx + 1] = light[3 * idx + 1] + sc.y;
                light[3 * idx + 2] = light[3 * idx + 2] + sc.z;
                weight[idx] = weight[idx] + p.filterWeight;
                for (int ch = 0; ch < 3; ++ch) albedo[3 * idx + ch] = albedo[3 * idx + ch] + p.albedo[ch];
                const f3 n = normalized(mk(p.firstNormal[0], p.firstNormal[1], p.firstNormal[2]));
                normal[3 * idx] = normal[3 * idx] + fabsf(n.x);
                normal[3 * idx + 1] = normal[3 * idx + 1] + fabsf(n.y);
                normal[3 * idx + 2] = normal[3 * idx + 2] + fabsf(n.z);
        }
}

__global__ __launch_bounds__(kFilmBlock) void k_film_resolve(const float *__restrict__ light,
                                                             const float *__restrict__ weight, int64_t pixels,
                                                             float *__restrict__ out, int *__restrict__ zeroWeight) {
        const int64_t i = (int64_t)blockIdx.x * kFilmBlock + threadIdx.x;
        if (i >= pixels) return;
        const float l0 = light[3 * i], l1 = light[3 * i + 1], l2 = light[3 * i + 2];
        if (l0 == 0.0f && l1 == 0.0f && l2 == 0.0f) {
                out[3 * i] = l0;
                out[3 * i + 1] = l1;
                out[3 * i + 2] = l2;
                return;
        }
        const float w = weight[i];
        if (w == 0.0f) {
                atomicAdd(zeroWeight, 1);
                out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = NAN;
                return;
        }
        out[3 * i] = l0 / w;
        out[3 * i + 1] = l1 / w;
        out[3 * i + 2] = l2 / w;
}

static bool camera_ok(const GzCamera *c, const char *who) {
        if (!c) {
                set_error("%s: null camera", who);
                return false;
        }
        if (c->resX <= 0 || c->resY <= 0 || c->x0 < 0 || c->y0 < 0 || c->x1 > c->resX || c->y1 > c->resY ||
            c->x0 >= c->x1 || c->y0 >= c->y1 || c->spp <= 0) {
                set_error("%s: bad resolution, crop [%d,%d)x[%d,%d) or spp %d", who, c->x0, c->x1, c->y0, c->y1,
                          c->spp);
                return false;
        }
        if (c->filterType != GZ_FILTER_GAUSSIAN && c->filterType != GZ_FILTER_BOX) {
                set_error("%s: unknown filter type %d", who, c->filterType);
                return false;
        }
        if (c->lensRadius != 0.0f) {
                set_error("%s: thin-lens depth of field is not supported (lensRadius %g)", who, (double)c->lensRadius);
                return false;
        }
        if (c->zsLog2Spp < 0 || c->zsLog2Spp > 64 || c->zsBase4Digits < 0 || c->zsBase4Digits > 64 || c->zsSeed < 0) {
                set_error("%s: bad Z-Sobol parameters", who);
                return false;
        }
        return true;
}

static int64_t crop_paths(const GzCamera *c) {
        return (int64_t)(c->x1 - c->x0) * (int64_t)(c->y1 - c->y0) * (int64_t)c->spp;
}

}  // namespace gz

using namespace gz;

extern "C" int gz_gpu_camera_paths_device(const GzCamera *camera, int64_t first, int64_t count, GzPath *d_paths,
                                          void *stream) {
        clear_error();
        if (!camera_ok(camera, "gz_gpu_camera_paths_device")) return GZ_E_INVALID;
        if (first < 0 || count < 0 || first + count > crop_paths(camera) || (count > 0 && !d_paths)) {
                set_error("gz_gpu_camera_paths_device: paths [%lld, %lld) outside the crop's %lld", (long long)first,
                          (long long)(first + count), (long long)crop_paths(camera));
                return GZ_E_INVALID;
        }
        if (count == 0) return GZ_OK;
        const uint32_t *table = device_sobol_table();
        if (!table) return GZ_E_HIP;
        const int64_t blocks = (count + kFilmBlock - 1) / kFilmBlock;
        if (blocks > 0x7fffffffLL) {
                set_error("gz_gpu_camera_paths_device: too many paths for one launch");
                return GZ_E_INVALID;
        }
        hipLaunchKernelGGL(k_camera_paths, dim3((unsigned)blocks), dim3(kFilmBlock), 0, static_cast<hipStream_t>(stream),
                           *camera, table, first, count, d_paths);
        return hip_ok(hipGetLastError(), "camera launch") ? GZ_OK : GZ_E_HIP;
}

extern "C" int gz_gpu_film_add_device(const GzCamera *camera, const GzPath *d_paths, int64_t count, float iso,
                                      float maxComponentValue, float *d_light, float *d_weight, float *d_albedo,
                                      float *d_normal, void *stream) {
        clear_error();
        if (!camera_ok(camera, "gz_gpu_film_add_device")) return GZ_E_INVALID;
        if (count < 0 || count % camera->spp != 0 || (count > 0 && (!d_paths || !d_light || !d_weight || !d_albedo ||
                                                                   !d_normal))) {
                set_error("gz_gpu_film_add_device: need whole pixels (count %lld, spp %d) and all images",
                          (long long)count, camera->spp);
                return GZ_E_INVALID;
        }
        if (count == 0) return GZ_OK;
        const int64_t pixels = count / camera->spp;
        const int64_t blocks = (pixels + kFilmBlock - 1) / kFilmBlock;
        hipLaunchKernelGGL(k_film_add, dim3((unsigned)blocks), dim3(kFilmBlock), 0, static_cast<hipStream_t>(stream),
                           camera->resX, camera->resY, camera->spp, d_paths, pixels, iso, maxComponentValue, d_light,
                           d_weight, d_albedo, d_normal);
        return hip_ok(hipGetLastError(), "film launch") ? GZ_OK : GZ_E_HIP;
}

extern "C" int gz_gpu_film_resolve_device(const float *d_light, const float *d_weight, int64_t pixels, float *d_out,
                                          int32_t *d_zeroWeight, void *stream) {
        clear_error();
        if (pixels < 0 || (pixels > 0 && (!d_light || !d_weight || !d_out || !d_zeroWeight))) {
                set_error("gz_gpu_film_resolve_device: bad arguments");
                return GZ_E_INVALID;
        }
        if (pixels == 0) return GZ_OK;
        const int64_t blocks = (pixels + kFilmBlock - 1) / kFilmBlock;
        hipLaunchKernelGGL(k_film_resolve, dim3((unsigned)blocks), dim3(kFilmBlock), 0, static_cast<hipStream_t>(stream),
                           d_light, d_weight, pixels, d_out, d_zeroWeight);
        return hip_ok(hipGetLastError(), "film resolve launch") ? GZ_OK : GZ_E_HIP;
}
