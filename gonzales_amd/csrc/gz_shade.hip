// Diffuse wavefront shade (the shade step of the reference's Tile.render bounce loop,
// Core/Tile.swift:100-155), restating shade_core (Sources/mojoKernel/kernel.mojo:676-773):
//   - inactive path or miss -> deactivate;
//   - material type 2 (area light) -> estimate += throughput * emission, deactivate (two-sided,
//     every bounce);
//   - material type 1 -> geometric normal normalize((p1-p0) x (p2-p0)) flipped against the ray,
//     PCG32 re-seeded from (pcgState, pcgInc) with pcg32_srandom semantics (kernel.mojo:99-118),
//     cosine-hemisphere direction in the Duff et al. frame, origin o + d*t + n*1e-4,
//     throughput *= albedo, pcgState advanced;
//   - any other material type -> deactivate.
// Arithmetic is written in the reference's order with no contraction; sqrt and division are
// IEEE-correct; cos/sin come from the device math library (the reference's come from the Mojo
// standard library), so the direction agrees to a few ulp, not bit for bit.
#include "gz_internal.h"

#pragma clang fp contract(off)

namespace gz {

constexpr int kShadeBlock = 256;

struct Pcg {
        uint64_t state, inc;
        __device__ __forceinline__ uint32_t next() {
                const uint64_t old = state;
                state = old * 6364136223846793005ull + inc;
                const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
                const uint32_t rot = (uint32_t)(old >> 59);
                return (xs >> rot) | (xs << ((0u - rot) & 31u));
        }
        __device__ __forceinline__ float next_float() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

__global__ __launch_bounds__(kShadeBlock) void k_shade(PathState_C *__restrict__ paths,
                                                       const Intersection_C *__restrict__ isects,
                                                       const DevMesh *__restrict__ meshes, int64_t meshCount,
                                                       const Material_C *__restrict__ mats, int64_t matCount, int64_t n,
                                                       int *__restrict__ err) {
        const int64_t i = (int64_t)blockIdx.x * kShadeBlock + threadIdx.x;
        if (i >= n) return;
        PathState_C *p = paths + i;
        if (p->active == 0) return;
        const Intersection_C is = isects[i];
        if (is.hit == 0) {
                p->active = 0;
                return;
        }
        const int64_t mi = is.primId.materialIndex;
        if (mi < 0 || mi >= matCount) {  // the reference reads out of bounds here
                atomicOr(err, 2);
                p->active = 0;
                return;
        }
        const Material_C m = mats[mi];
        if (m.type == 2) {
                p->estimate[0] += p->throughput[0] * m.emission[0];
                p->estimate[1] += p->throughput[1] * m.emission[1];
                p->estimate[2] += p->throughput[2] * m.emission[2];
                p->active = 0;
                return;
        }
        if (m.type != 1) {
                p->active = 0;
                return;
        }
        int64_t mesh, base;
        if (is.primId.type == 0) {
                mesh = is.primId.id1;
                base = is.primId.id2;
        } else if (is.primId.type == 1 || is.primId.type == 2 || is.primId.type == 3) {
                mesh = is.primId.id2 >> 32;
                base = (is.primId.id2 & 0xFFFFFFFFLL) * 3;
        } else {
                p->active = 0;
                return;
        }
        if (mesh < 0 || mesh >= meshCount) {
                atomicOr(err, 2);
                p->active = 0;
                return;
        }
        const DevMesh me = meshes[mesh];
        if (base < 0 || base + 2 >= me.vertexIndexCount) {
                atomicOr(err, 2);
                p->active = 0;
                return;
        }
        const int64_t i0 = me.vertexIndices[base], i1 = me.vertexIndices[base + 1], i2 = me.vertexIndices[base + 2];
        const int64_t lim = me.pointFloats;
        if (i0 < 0 || i1 < 0 || i2 < 0 || i0 * 4 + 2 >= lim || i1 * 4 + 2 >= lim || i2 * 4 + 2 >= lim) {
                atomicOr(err, 2);
                p->active = 0;
                return;
        }
        const float p0x = me.points[i0 * 4], p0y = me.points[i0 * 4 + 1], p0z = me.points[i0 * 4 + 2];
        const float p1x = me.points[i1 * 4], p1y = me.points[i1 * 4 + 1], p1z = me.points[i1 * 4 + 2];
        const float p2x = me.points[i2 * 4], p2y = me.points[i2 * 4 + 1], p2z = me.points[i2 * 4 + 2];
        const float e1x = p1x - p0x, e1y = p1y - p0y, e1z = p1z - p0z;
        const float e2x = p2x - p0x, e2y = p2y - p0y, e2z = p2z - p0z;
        float nx = e1y * e2z - e1z * e2y;
        float ny = e1z * e2x - e1x * e2z;
        float nz = e1x * e2y - e1y * e2x;
        const float nlen = (nx * nx + ny * ny) + nz * nz;
        if (nlen > 0.0f) {
                const float s = 1.0f / sqrtf(nlen);
                nx = nx * s;
                ny = ny * s;
                nz = nz * s;
        }
        const float dx = p->ray.dirX, dy = p->ray.dirY, dz = p->ray.dirZ;
        if ((nx * dx + ny * dy) + nz * dz > 0.0f) {
                nx = nx * -1.0f;
                ny = ny * -1.0f;
                nz = nz * -1.0f;
        }
        Pcg g;
        g.state = 0;
        g.inc = (p->pcgInc << 1) | 1u;
        (void)g.next();
        g.state += p->pcgState;
        (void)g.next();
        const float u1 = g.next_float();
        const float u2 = g.next_float();
        p->pcgState = g.state;
        const float r = sqrtf(u1);
        const float theta = 2.0f * 3.14159265359f * u2;
        const float x = r * cosf(theta);
        const float y = r * sinf(theta);
        const float z2 = 1.0f - u1;
        const float z = sqrtf(z2 > 0.0f ? z2 : 0.0f);
        const float sign = nz >= 0.0f ? 1.0f : -1.0f;
        const float a = -1.0f / (sign + nz);
        const float b = nx * ny * a;
        const float tx = 1.0f + sign * nx * nx * a, ty = sign * b, tz = -sign * nx;
        const float bx = b, by = sign + ny * ny * a, bz = -ny;
        float ox = tx * x + bx * y + nx * z;
        float oy = ty * x + by * y + ny * z;
        float oz = tz * x + bz * y + nz * z;
        const float dlen = (ox * ox + oy * oy) + oz * oz;
        if (dlen > 0.0f) {
                const float s = 1.0f / sqrtf(dlen);
                ox = ox * s;
                oy = oy * s;
                oz = oz * s;
        }
        const float t = is.tHit;
        p->ray.orgX = p->ray.orgX + dx * t + nx * 0.0001f;
        p->ray.orgY = p->ray.orgY + dy * t + ny * 0.0001f;
        p->ray.orgZ = p->ray.orgZ + dz * t + nz * 0.0001f;
        p->ray.dirX = ox;
        p->ray.dirY = oy;
        p->ray.dirZ = oz;
        p->throughput[0] *= m.albedo[0];
        p->throughput[1] *= m.albedo[1];
        p->throughput[2] *= m.albedo[2];
}

}  // namespace gz

using namespace gz;

static int shade_launch(Scene *s, PathState_C *d_paths, int64_t n, const Intersection_C *d_isects, hipStream_t st) {
        if (n <= 0) return GZ_OK;
        const StreamSlot *slot = stream_slot(s, st);
        if (!slot) return GZ_E_ALLOC;
        hipLaunchKernelGGL(k_shade, dim3((unsigned)((n + kShadeBlock - 1) / kShadeBlock)), dim3(kShadeBlock), 0, st,
                           d_paths, d_isects, s->dMeshes, s->meshCount, s->dMaterials, s->materialCount, n, slot->dErr);
        return hip_ok(hipGetLastError(), "shade launch") ? GZ_OK : GZ_E_HIP;
}

extern "C" int gz_gpu_shade_device(void *handle, PathState_C *d_paths, int64_t count, const Intersection_C *d_isects,
                                   void *stream) {
        clear_error();
        if (!handle) {
                set_error("gz_gpu_shade_device: null scene handle");
                return GZ_E_INVALID;
        }
        if (count < 0 || (count > 0 && (!d_paths || !d_isects))) {
                set_error("gz_gpu_shade_device: bad arguments");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return shade_launch(s, d_paths, count, d_isects, static_cast<hipStream_t>(stream));
}

// mojo_gpu_shade_batch (kernel.mojo:788-846): host paths updated in place.
extern "C" void gz_gpu_shade_batch(void *handle, PathState_C *paths, int64_t count, const Intersection_C *intersections) {
        clear_error();
        if (!handle || count <= 0) return;  // kernel.mojo:795-800
        if (!paths || !intersections) {
                set_error("gz_gpu_shade_batch: null buffer");
                return;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return;
        const size_t n = (size_t)count;
        const size_t offI = (n * sizeof(PathState_C) + 255) & ~(size_t)255;
        if (!ensure_stage(s, offI + n * sizeof(Intersection_C))) return;
        char *base = static_cast<char *>(s->dStage);
        PathState_C *dP = reinterpret_cast<PathState_C *>(base);
        Intersection_C *dI = reinterpret_cast<Intersection_C *>(base + offI);
        hipStream_t st = s->stream;
        if (!hip_ok(hipMemcpyAsync(dP, paths, n * sizeof(PathState_C), hipMemcpyHostToDevice, st), "copy paths") ||
            !hip_ok(hipMemcpyAsync(dI, intersections, n * sizeof(Intersection_C), hipMemcpyHostToDevice, st),
                    "copy intersections"))
                return;
        if (shade_launch(s, dP, count, dI, st) != GZ_OK) return;
        if (!hip_ok(hipMemcpyAsync(paths, dP, n * sizeof(PathState_C), hipMemcpyDeviceToHost, st), "copy paths back"))
                return;
        (void)hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
}
