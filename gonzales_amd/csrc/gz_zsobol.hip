// Z-Sobol sample generation on the device (the sampler of Sources/libgonzales/Sampler/
// ZSobolSampler.swift, driven per path as Core/Tile.swift:28-33 does: startPixelSample, then a
// fixed sequence of get1D/get2D/get3D calls). One lane per (pixel, sample index).
#include <cstring>
#include <mutex>
#include <vector>

#include "gz_internal.h"
#include "zsobol_device.h"

namespace gz {

const uint32_t *sobol_table();  // sobol_table.cpp

constexpr int kZsBlock = 256;
constexpr int kMaxCalls = 64;

struct CallPattern {
        int8_t calls[kMaxCalls];
        int32_t n;
        int32_t stride;
};

// mode 0: explicit (px, py, index) lists; mode 1: a tile x0..x0+w, y0..y0+h, spp samples per
// pixel, sample i = pixel (i / spp) in row-major order, index i % spp (Tile.swift:28-29 order).
template <int kMode>
__global__ __launch_bounds__(kZsBlock) void k_zsobol(ZSobolParams p, const uint32_t *__restrict__ table,
                                                     const int32_t *__restrict__ pxs, const int32_t *__restrict__ pys,
                                                     const int64_t *__restrict__ idxs, int32_t x0, int32_t y0,
                                                     int32_t w, int32_t spp, int64_t n, int32_t startDim,
                                                     CallPattern pat, float *__restrict__ out, int32_t soa) {
        const int64_t i = (int64_t)blockIdx.x * kZsBlock + threadIdx.x;
        if (i >= n) return;
        int32_t px, py;
        int64_t index;
        if (kMode == 0) {
                px = pxs[i];
                py = pys[i];
                index = idxs[i];
        } else {
                const int64_t pix = i / spp;
                index = i % spp;
                px = x0 + (int32_t)(pix % w);
                py = y0 + (int32_t)(pix / w);
        }
        ZSobolState z = zs_start(p, px, py, index, startDim);
        int k = 0;
        auto put = [&](float v) {
                if (soa)
                        out[(int64_t)k * n + i] = v;
                else
                        out[i * pat.stride + k] = v;
                ++k;
        };
        for (int c = 0; c < pat.n; ++c) {
                const int8_t kind = pat.calls[c];
                if (kind == 1) {
                        put(zs_get1d(p, z, table));
                } else if (kind == 2) {
                        float a, b;
                        zs_get2d(p, z, table, a, b);
                        put(a);
                        put(b);
                } else {
                        float a, b, cc;
                        zs_get3d(p, z, table, a, b, cc);
                        put(a);
                        put(b);
                        put(cc);
                }
        }
}

static std::mutex g_table_mu;
static uint32_t *g_dev_table[64] = {nullptr};

// The byte table of the 1024 x 52 matrix table (zsobol_device.h), on the host.
static std::vector<uint32_t> byte_table() {
        const uint32_t *cols = sobol_table();
        std::vector<uint32_t> t(kSobolByteTableWords, 0u);
        for (int d = 0; d < kSobolDims; ++d)
                for (int k = 0; k < kSobolBytes; ++k)
                        for (int v = 0; v < 256; ++v) {
                                uint32_t acc = 0;
                                for (int j = 0; j < 8; ++j) {
                                        const int c = 8 * k + j;
                                        if (((v >> j) & 1) && c < kSobolCols) acc ^= cols[d * kSobolCols + c];
                                }
                                t[((size_t)d * kSobolBytes + k) * 256 + v] = acc;
                        }
        return t;
}

// Device copy of the matrices in their byte-table form (7 MiB), made once per device.
const uint32_t *device_sobol_table() {
        int dev = 0;
        if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
        if (dev < 0 || dev >= 64) {
                set_error("device ordinal %d out of range", dev);
                return nullptr;
        }
        std::lock_guard<std::mutex> lock(g_table_mu);
        if (!g_dev_table[dev]) {
                uint32_t *d = nullptr;
                const size_t bytes = sizeof(uint32_t) * kSobolByteTableWords;
                if (!hip_ok(hipMalloc((void **)&d, bytes), "hipMalloc(sobol table)")) return nullptr;
                const std::vector<uint32_t> host = byte_table();
                if (!hip_ok(hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice), "upload sobol table")) {
                        (void)hipFree(d);
                        return nullptr;
                }
                g_dev_table[dev] = d;
        }
        return g_dev_table[dev];
}

static bool make_pattern(const int8_t *pattern, int32_t nCalls, CallPattern &pat) {
        if (nCalls < 0 || nCalls > kMaxCalls || (nCalls > 0 && !pattern)) {
                set_error("call pattern: %d calls (0..%d allowed)", nCalls, kMaxCalls);
                return false;
        }
        std::memset(&pat, 0, sizeof pat);
        pat.n = nCalls;
        for (int c = 0; c < nCalls; ++c) {
                if (pattern[c] < 1 || pattern[c] > 3) {
                        set_error("call pattern entry %d is %d (must be 1, 2 or 3)", c, (int)pattern[c]);
                        return false;
                }
                pat.calls[c] = pattern[c];
                pat.stride += pattern[c];
        }
        return true;
}

static bool check_params(const GzZSobolParams *params, int32_t startDim) {
        if (!params) {
                set_error("null sampler parameters");
                return false;
        }
        if (params->log2SamplesPerPixel < 0 || params->log2SamplesPerPixel > 64 || params->nBase4Digits < 0 ||
            params->nBase4Digits > 64 || params->seed < 0) {
                set_error("sampler parameters out of range");
                return false;
        }
        if (startDim < 0 || startDim >= kSobolDims) {
                set_error("start dimension %d outside [0, %d)", startDim, kSobolDims);
                return false;
        }
        return true;
}

}  // namespace gz

using namespace gz;

// ZSobolSampler.init (ZSobolSampler.swift:114-122).
extern "C" int gz_zsobol_params(int64_t samplesPerPixel, int64_t resX, int64_t resY, int64_t seed,
                                GzZSobolParams *out) {
        clear_error();
        if (!out || seed < 0 || resX < 0 || resY < 0) {
                set_error("gz_zsobol_params: bad arguments");
                return GZ_E_INVALID;
        }
        const int tzSpp = samplesPerPixel == 0 ? 64 : __builtin_ctzll((unsigned long long)samplesPerPixel);
        int64_t r = resX > resY ? resX : resY;
        r -= 1;
        r |= r >> 1;
        r |= r >> 2;
        r |= r >> 4;
        r |= r >> 8;
        r |= r >> 16;
        r += 1;
        const int tzRes = r == 0 ? 64 : __builtin_ctzll((unsigned long long)r);
        out->seed = seed;
        out->log2SamplesPerPixel = tzSpp;
        out->nBase4Digits = tzRes + (tzSpp + 1) / 2;
        return GZ_OK;
}

extern "C" int gz_zsobol_samples_device(const GzZSobolParams *params, const int32_t *d_px, const int32_t *d_py,
                                        const int64_t *d_sampleIndex, int64_t count, int32_t startDim,
                                        const int8_t *pattern, int32_t nCalls, float *d_out, int32_t layout,
                                        void *stream) {
        clear_error();
        if (!check_params(params, startDim)) return GZ_E_INVALID;
        CallPattern pat;
        if (!make_pattern(pattern, nCalls, pat)) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_px || !d_py || !d_sampleIndex || !d_out))) {
                set_error("gz_zsobol_samples_device: bad buffers");
                return GZ_E_INVALID;
        }
        if (count == 0 || pat.n == 0) return GZ_OK;
        const uint32_t *table = device_sobol_table();
        if (!table) return GZ_E_HIP;
        ZSobolParams p{params->seed, params->log2SamplesPerPixel, params->nBase4Digits};
        hipLaunchKernelGGL(k_zsobol<0>, dim3((unsigned)((count + kZsBlock - 1) / kZsBlock)), dim3(kZsBlock), 0,
                           static_cast<hipStream_t>(stream), p, table, d_px, d_py, d_sampleIndex, 0, 0, 1, 1, count,
                           startDim, pat, d_out, layout ? 1 : 0);
        return hip_ok(hipGetLastError(), "zsobol launch") ? GZ_OK : GZ_E_HIP;
}

extern "C" int gz_zsobol_tile_device(const GzZSobolParams *params, int32_t x0, int32_t y0, int32_t width,
                                     int32_t height, int32_t spp, int32_t startDim, const int8_t *pattern,
                                     int32_t nCalls, float *d_out, int32_t layout, void *stream) {
        clear_error();
        if (!check_params(params, startDim)) return GZ_E_INVALID;
        CallPattern pat;
        if (!make_pattern(pattern, nCalls, pat)) return GZ_E_INVALID;
        if (width < 0 || height < 0 || spp < 0 || x0 < 0 || y0 < 0) {
                set_error("gz_zsobol_tile_device: negative tile size");
                return GZ_E_INVALID;
        }
        const int64_t count = (int64_t)width * height * spp;
        if (count == 0 || pat.n == 0) return GZ_OK;
        if (!d_out) {
                set_error("gz_zsobol_tile_device: null output");
                return GZ_E_INVALID;
        }
        const uint32_t *table = device_sobol_table();
        if (!table) return GZ_E_HIP;
        ZSobolParams p{params->seed, params->log2SamplesPerPixel, params->nBase4Digits};
        hipLaunchKernelGGL(k_zsobol<1>, dim3((unsigned)((count + kZsBlock - 1) / kZsBlock)), dim3(kZsBlock), 0,
                           static_cast<hipStream_t>(stream), p, table, nullptr, nullptr, nullptr, x0, y0, width, spp,
                           count, startDim, pat, d_out, layout ? 1 : 0);
        return hip_ok(hipGetLastError(), "zsobol launch") ? GZ_OK : GZ_E_HIP;
}
