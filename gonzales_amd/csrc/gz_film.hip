// Camera samples, primary rays and film accumulation on gfx950: the steps on either side of
// the path (Core/Tile.swift:27-59 and :170-219, Core/Film.swift:107-136).
//
//   k_camera_paths: one lane per path, paths pixel-major over the crop (index = pixel * spp +
//     sample): startPixelSample (dim 0), getCameraSample (filter get2D, lens get2D;
//     Sampler.swift:69-83), the filter's sample and weight (GaussianFilter.swift:7-121,
//     BoxFilter.swift:3-14), PerspectiveCamera.generateRay as a pinhole
//     (PerspectiveCamera.swift:37-48, Transform.swift:31-65); writes fresh GzPath records.
//   k_film_add_staged: one lane per pixel, adding its spp samples in sample order to the beauty
//     (estimate x iso/100, clamped to maxComponentValue when > 0), weight, albedo and
//     normal (|normalized n|) sums (Film.swift:107-136, Image.swift:21-30, Pixel.swift:33-39);
//     the records staged through LDS.
//   k_film_resolve: Pixel.normalized (Pixel.swift:12-16): black stays, else light / weight.
#include "film_device.h"

#pragma clang fp contract(off)

namespace gz {

constexpr int kFilmBlock = 256;

__global__ __launch_bounds__(kFilmBlock) void k_camera_paths(GzCamera c, const uint32_t *__restrict__ table,
                                                             int64_t first, int64_t count, GzPath *__restrict__ out) {
        const int64_t k = (int64_t)blockIdx.x * kFilmBlock + threadIdx.x;
        if (k >= count) return;
        out[k] = camera_path(c, table, first + k);
}

// Film accumulation with the records staged through LDS. One lane per pixel reading its own
// records straight from memory (7 KB from its neighbours' at 64 spp) fetched every line of the
// records 2.5 times on the 5M-triangle frame. Here a 64-lane block takes 64 pixels; per round of kFilmStage samples its
// lanes load those samples of all 64 pixels with loads that run over each pixel's consecutive
// records, then each lane adds its own pixel's staged samples in sample order, holding the
// pixel's sums in registers while consecutive samples land on that pixel (the same additions, in
// the same order, as adding each sample to the sums in memory).
constexpr int kFilmStage = 4;
constexpr int kFilmFields = 12;                           // estimate, albedo, normal, weight, x, y
constexpr int kFilmRow = kFilmStage * kFilmFields + 1;  // odd: lanes reading their rows use distinct banks

__global__ __launch_bounds__(64) void k_film_add_staged(int resX, int resY, int spp, const GzPath *__restrict__ paths,
                                                       int64_t pixels, float iso, float maxComponent,
                                                       float *__restrict__ light, float *__restrict__ weight,
                                                       float *__restrict__ albedo, float *__restrict__ normal) {
        __shared__ float st[64 * kFilmRow];
        const int lane = (int)threadIdx.x;
        const int64_t p0 = (int64_t)blockIdx.x * 64;
        const bool mine = p0 + lane < pixels;
        int64_t cur = -1;
        float L[3] = {0.0f, 0.0f, 0.0f}, W = 0.0f, A[3] = {0.0f, 0.0f, 0.0f}, N[3] = {0.0f, 0.0f, 0.0f};
        auto flush = [&]() {
                if (cur < 0) return;
                for (int ch = 0; ch < 3; ++ch) {
                        light[3 * cur + ch] = L[ch];
                        albedo[3 * cur + ch] = A[ch];
                        normal[3 * cur + ch] = N[ch];
                }
                weight[cur] = W;
        };
        const float k100 = iso / 100.0f;
        for (int s0 = 0; s0 < spp; s0 += kFilmStage) {
                for (int f = lane; f < 64 * kFilmStage; f += 64) {
                        const int q = f / kFilmStage, k = f % kFilmStage;
                        if (p0 + q >= pixels || s0 + k >= spp) continue;
                        const GzPath &p = paths[(p0 + q) * spp + s0 + k];
                        float *row = st + q * kFilmRow + k * kFilmFields;
                        for (int c = 0; c < 3; ++c) {
                                row[c] = p.estimate[c];
                                row[3 + c] = p.albedo[c];
                                row[6 + c] = p.firstNormal[c];
                        }
                        row[9] = p.filterWeight;
                        row[10] = __int_as_float(p.pixelX);
                        row[11] = __int_as_float(p.pixelY);
                }
                __syncthreads();
                if (mine) {
                        const int kn = spp - s0 < kFilmStage ? spp - s0 : kFilmStage;
                        for (int k = 0; k < kn; ++k) {
                                const float *row = st + lane * kFilmRow + k * kFilmFields;
                                const int x = __float_as_int(row[10]), y = __float_as_int(row[11]);
                                if (!(x >= 0 && y >= 0 && x < resX && y < resY)) continue;
                                const int64_t idx = (int64_t)y * resX + x;
                                if (idx != cur) {
                                        flush();
                                        cur = idx;
                                        for (int ch = 0; ch < 3; ++ch) {
                                                L[ch] = light[3 * idx + ch];
                                                A[ch] = albedo[3 * idx + ch];
                                                N[ch] = normal[3 * idx + ch];
                                        }
                                        W = weight[idx];
                                }
                                f3 sc = mk(row[0] * k100, row[1] * k100, row[2] * k100);
                                if (maxComponent > 0.0f) {
                                        const float mv = cmax(sc);
                                        if (mv > maxComponent) sc = mul(sc, maxComponent / mv);
                                }
                                L[0] = L[0] + sc.x;
                                L[1] = L[1] + sc.y;
                                L[2] = L[2] + sc.z;
                                W = W + row[9];
                                for (int ch = 0; ch < 3; ++ch) A[ch] = A[ch] + row[3 + ch];
                                const f3 n = normalized(mk(row[6], row[7], row[8]));
                                N[0] = N[0] + fabsf(n.x);
                                N[1] = N[1] + fabsf(n.y);
                                N[2] = N[2] + fabsf(n.z);
                        }
                }
                __syncthreads();
        }
        if (mine) flush();
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

bool camera_ok(const GzCamera *c, const char *who) {
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

int64_t crop_paths(const GzCamera *c) {
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
        const int64_t blocks = (pixels + 63) / 64;
        if (blocks > 0x7fffffffLL) {
                set_error("gz_gpu_film_add_device: too many pixels for one launch");
                return GZ_E_INVALID;
        }
        hipLaunchKernelGGL(k_film_add_staged, dim3((unsigned)blocks), dim3(64), 0, static_cast<hipStream_t>(stream),
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
