// Scene upload, lifetime, errors and device discovery for the C-ABI.
//
// gz_gpu_upload_scene replaces mojo_gpu_upload_scene (kernel.mojo:430-587). The reference copies
// the Swift arrays byte by byte into device buffers of the same layout; this library instead
// re-lays the scene out for the traversal kernel while copying:
//   - every reachable interior BVH2 node becomes one 64-byte child-pair record (both children's
//     boxes + descriptors), numbered depth-first from the root;
//   - every reachable leaf becomes a block of 48-byte triangle records, the vertices gathered
//     through PrimId -> mesh -> vertexIndices -> points exactly as kernel.mojo:266-300 does;
//   - primIds, per-mesh points/indices and the material table are copied unchanged for the
//     Intersection_C expansion and the shading kernel.
// The upload validates what the reference would read out of bounds (child indices, leaf ranges,
// mesh and vertex indices) and rejects cyclic node graphs, so no kernel can fault or loop.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "gz_internal.h"

namespace gz {

bool hip_ok(hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        set_error("%s: %s", what, hipGetErrorString(e));
        return false;
}

bool enter_device(Scene *s) {
        int cur = -1;
        if (!hip_ok(hipGetDevice(&cur), "hipGetDevice")) return false;
        if (cur != s->device && !hip_ok(hipSetDevice(s->device), "hipSetDevice")) return false;
        return true;
}

bool ensure_stage(Scene *s, size_t bytes) {
        if (bytes <= s->stageBytes) return true;
        if (s->dStage) {
                (void)hipFree(s->dStage);
                s->dStage = nullptr;
                s->stageBytes = 0;
        }
        size_t want = bytes + bytes / 4;
        if (!hip_ok(hipMalloc(&s->dStage, want), "hipMalloc(staging)")) return false;
        s->stageBytes = want;
        return true;
}

StreamSlot *stream_slot(Scene *s, hipStream_t stream) {
        auto it = s->slots.find(stream);
        if (it != s->slots.end()) return &it->second;
        // 128 bytes: the flag at 0, the eight ray counters at 64 (zeroed before any launch can use them)
        void *p = nullptr;
        if (alloc_fault("stream slot") || !hip_ok(hipMalloc(&p, 128), "hipMalloc(stream slot)")) return nullptr;
        if (!hip_ok(hipMemset(p, 0, 128), "hipMemset(stream slot)")) {
                (void)hipFree(p);
                return nullptr;
        }
        s->deviceBytes += 128;
        StreamSlot slot{static_cast<int *>(p), reinterpret_cast<unsigned long long *>(static_cast<char *>(p) + 64)};
        return &s->slots.emplace(stream, slot).first->second;
}

int check_stream(Scene *s, hipStream_t st) {
        auto it = s->slots.find(st);
        if (it == s->slots.end()) return GZ_OK;  // nothing ran on this stream
        int flag = 0;
        if (!hip_ok(hipMemcpyAsync(&flag, it->second.dErr, sizeof(int), hipMemcpyDeviceToHost, st), "read error flag") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        if (!flag) return GZ_OK;
        if (!hip_ok(hipMemsetAsync(it->second.dErr, 0, sizeof(int), st), "clear error flag") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        if (flag & 1) {
                set_error("traversal stack overflow: a ray needed more than %d entries", GZ_STACK_CAPACITY);
                return GZ_E_OVERFLOW;
        }
        set_error("a shading kernel met a primitive, mesh or vertex index out of range");
        return GZ_E_INVALID;
}

static void free_scene(Scene *s) {
        if (!s) return;
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != s->device) (void)hipSetDevice(s->device);
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        release_mis(s->mis);
        release_mis_queues(s->misQueues);
        for (auto &kv : s->slots) {
                if (kv.second.dErr) (void)hipFree(kv.second.dErr);
                if (kv.second.dBin) (void)hipFree(kv.second.dBin);
                for (hipEvent_t e : kv.second.ev)
                        if (e) (void)hipEventDestroy(e);
        }
        void *ptrs[] = {s->dNodes, s->dTris, s->dPrims, s->dMeshes, s->dMaterials, s->dStage, s->dBuildOrder,
                        s->dPrimTris};
        for (void *p : ptrs)
                if (p) (void)hipFree(p);
        for (int64_t i = 0; i < s->meshBufferCount; ++i)
                if (s->meshBuffers[i]) (void)hipFree(s->meshBuffers[i]);
        delete[] s->meshBuffers;
        if (s->stream) (void)hipStreamDestroy(s->stream);
        if (cur >= 0 && cur != s->device) (void)hipSetDevice(cur);
        delete s;
}

template <class T>
static bool upload(T **dst, const T *src, int64_t count, Scene *s, const char *what) {
        size_t bytes = sizeof(T) * (size_t)(count > 0 ? count : 1);
        if (alloc_fault(what) || !hip_ok(hipMalloc((void **)dst, bytes), what)) return false;
        s->deviceBytes += (int64_t)bytes;
        if (count > 0 && !hip_ok(hipMemcpy(*dst, src, sizeof(T) * (size_t)count, hipMemcpyHostToDevice), what))
                return false;
        return true;
}

bool scene_upload_common(Scene *s, const PrimId_C *prims, int64_t nPrims, const TriangleMesh_C *meshes,
                         int64_t nMeshes, const int64_t *ptsCounts, const int64_t *faceCounts,
                         const int64_t *vtxCounts, const Material_C *materials, int64_t nMaterials) {
        if (!s->stream && !hip_ok(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate"))
                return false;
        // surfaces whose material is an interface: triangles (type 0) and geometric or transformed
        // primitives (types 1, 2) the traversal does not skip (id2 != -1); area lights (type 3)
        // shade with material 0
        s->interfacePrims = 0;
        for (int64_t i = 0; i < nPrims; ++i) {
                const PrimId_C &p = prims[i];
                const bool surface = p.type == 0 || ((p.type == 1 || p.type == 2) && p.id2 != -1);
                if (surface && p.materialIndex < 0) ++s->interfacePrims;
        }
        bool ok = upload(&s->dPrims, prims, nPrims, s, "upload primIds") &&
                  upload(&s->dMaterials, materials, nMaterials, s, "upload materials") &&
                  hip_ok(hipDeviceGetAttribute(&s->cuCount, hipDeviceAttributeMultiprocessorCount, s->device),
                         "hipDeviceGetAttribute") &&
                  stream_slot(s, s->stream) != nullptr;
        if (ok && nMeshes > 0) {
                // Per-mesh arrays for the shading kernels (kernel.mojo:488-531 uploads the same).
                std::vector<DevMesh> dm((size_t)nMeshes);
                s->meshBuffers = new (std::nothrow) void *[(size_t)(3 * nMeshes)]();
                if (!s->meshBuffers) {
                        set_error("scene upload: out of host memory");
                        ok = false;
                }
                for (int64_t i = 0; ok && i < nMeshes; ++i) {
                        float *pts = nullptr;
                        int64_t *fi = nullptr, *vi = nullptr;
                        const int64_t nf = faceCounts ? faceCounts[i] : 0;
                        ok = upload(&pts, meshes[i].points, ptsCounts[i], s, "upload points") &&
                             upload(&vi, meshes[i].vertexIndices, vtxCounts[i], s, "upload vertexIndices") &&
                             upload(&fi, nf > 0 ? meshes[i].faceIndices : nullptr, nf, s, "upload faceIndices");
                        s->meshBuffers[3 * i] = pts;
                        s->meshBuffers[3 * i + 1] = vi;
                        s->meshBuffers[3 * i + 2] = fi;
                        s->meshBufferCount = 3 * (i + 1);
                        dm[(size_t)i] = DevMesh{pts, fi, vi, ptsCounts[i], vtxCounts[i]};
                }
                if (ok) ok = upload(&s->dMeshes, dm.data(), nMeshes, s, "upload meshes");
        }
        return ok;
}

void scene_release(Scene *s) { free_scene(s); }

static inline int32_t as_int(float f) {
        int32_t i;
        std::memcpy(&i, &f, 4);
        return i;
}
static inline float as_float(int32_t i) {
        float f;
        std::memcpy(&f, &i, 4);
        return f;
}

}  // namespace gz

using namespace gz;

extern "C" bool gz_gpu_available(void) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return false;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return false;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
        return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

extern "C" void *gz_gpu_upload_scene(const BVH2Node *nodes, int64_t nNodes, const PrimId_C *prims, int64_t nPrims,
                                     const TriangleMesh_C *meshes, int64_t nMeshes, const int64_t *ptsCounts,
                                     const int64_t *faceCounts, const int64_t *vtxCounts, const Material_C *materials,
                                     int64_t nMaterials) {
        clear_error();
        alloc_fault_arm();
        if (nNodes < 0 || nPrims < 0 || nMeshes < 0 || nMaterials < 0) {
                set_error("gz_gpu_upload_scene: negative count");
                return nullptr;
        }
        if ((nNodes > 0 && !nodes) || (nPrims > 0 && !prims) || (nMeshes > 0 && (!meshes || !ptsCounts || !vtxCounts)) ||
            (nMaterials > 0 && !materials)) {
                set_error("gz_gpu_upload_scene: null array with a non-zero count");
                return nullptr;
        }
        if (!gz_gpu_available()) {
                set_error("gz_gpu_upload_scene: no gfx950 device visible");
                return nullptr;
        }
        if (nNodes > INT32_MAX || nPrims > INT32_MAX) {
                set_error("gz_gpu_upload_scene: more than 2^31 nodes or primitives");
                return nullptr;
        }
        Scene *s = nullptr;
        try {
                s = new Scene();
        } catch (const std::bad_alloc &) {
                set_error("gz_gpu_upload_scene: out of host memory");
                return nullptr;
        }
        if (!hip_ok(hipGetDevice(&s->device), "hipGetDevice")) {
                delete s;
                return nullptr;
        }
        s->nodeCount = nNodes;
        s->primCount = nPrims;
        s->meshCount = nMeshes;
        s->materialCount = nMaterials;
        s->empty = nNodes == 0;

        std::vector<NodePair> pairs;
        std::vector<TriRec> tris;
        try {
                if (alloc_fault("host node and triangle records")) throw std::bad_alloc();
                if (nNodes > 0) {
                        // Depth-first walk from node 0 (the reference's traversal start). Record ids are
                        // assigned on first visit; a node reachable twice (a DAG) keeps one record.
                        std::vector<int32_t> recOf((size_t)nNodes, INT32_MIN);
                        struct Item { int64_t node; int32_t depth; };
                        std::vector<Item> todo;
                        std::vector<int64_t> order;  // interior nodes in record order
                        todo.push_back({0, 0});
                        int64_t triNext = 0;
                        std::vector<int64_t> leafStartOf;  // parallel to leaves
                        while (!todo.empty()) {
                                Item it = todo.back();
                                todo.pop_back();
                                const int64_t i = it.node;
                                if (recOf[(size_t)i] != INT32_MIN) continue;
                                if (it.depth > s->maxDepth) s->maxDepth = it.depth;
                                const BVH2Node &nd = nodes[i];
                                if (nd.count > 0) {
                                        if (nd.offset < 0 || (int64_t)nd.offset + nd.count > nPrims) {
                                                set_error("gz_gpu_upload_scene: leaf %lld range [%d, %d+%d) outside %lld primIds",
                                                          (long long)i, nd.offset, nd.offset, nd.count, (long long)nPrims);
                                                free_scene(s);
                                                return nullptr;
                                        }
                                        if (triNext + nd.count > INT32_MAX) {
                                                set_error("gz_gpu_upload_scene: too many triangle records");
                                                free_scene(s);
                                                return nullptr;
                                        }
                                        recOf[(size_t)i] = ~(int32_t)triNext;
                                        triNext += nd.count;
                                } else {
                                        const int64_t l = i + 1, r = nd.offset;
                                        if (l >= nNodes || r <= i || r >= nNodes) {
                                                set_error("gz_gpu_upload_scene: interior node %lld has children (%lld, %lld) "
                                                          "outside (node, %lld)", (long long)i, (long long)l, (long long)r,
                                                          (long long)nNodes);
                                                free_scene(s);
                                                return nullptr;
                                        }
                                        recOf[(size_t)i] = (int32_t)order.size();
                                        order.push_back(i);
                                        // Children have larger indices than their parent, so the graph is acyclic.
                                        todo.push_back({r, it.depth + 1});
                                        todo.push_back({l, it.depth + 1});
                                }
                        }
                        s->interiorCount = (int64_t)order.size();
                        // Every box the kernel tests is a child of a reachable interior node.
                        bool safe = true;
                        for (size_t k = 0; k < order.size() && safe; ++k) {
                                const BVH2Node *ch[2] = {&nodes[order[k] + 1], &nodes[nodes[order[k]].offset]};
                                for (const BVH2Node *c : ch) {
                                        const float v[6] = {c->boundsMinX, c->boundsMinY, c->boundsMinZ,
                                                            c->boundsMaxX, c->boundsMaxY, c->boundsMaxZ};
                                        for (float x : v)
                                                if (!(x >= -0x1p40f && x <= 0x1p40f)) safe = false;
                                }
                        }
                        s->boundsSafe = safe;
                        s->triCount = triNext;
                        s->rootDesc = recOf[0];
                        pairs.resize(order.size());
                        for (size_t k = 0; k < order.size(); ++k) {
                                const int64_t i = order[k];
                                const BVH2Node &L = nodes[i + 1];
                                const BVH2Node &R = nodes[nodes[i].offset];
                                NodePair &p = pairs[k];
                                p.a = make_float4(L.boundsMinX, L.boundsMinY, L.boundsMinZ, L.boundsMaxX);
                                p.b = make_float4(L.boundsMaxY, L.boundsMaxZ, R.boundsMinX, R.boundsMinY);
                                p.c = make_float4(R.boundsMinZ, R.boundsMaxX, R.boundsMaxY, R.boundsMaxZ);
                                p.d = make_int4(recOf[(size_t)(i + 1)], recOf[(size_t)nodes[i].offset], 0, 0);
                        }
                        // Triangle records, leaf by leaf.
                        tris.resize((size_t)triNext);
                        for (int64_t i = 0; i < nNodes; ++i) {
                                const int32_t rec = recOf[(size_t)i];
                                if (rec == INT32_MIN || rec >= 0) continue;
                                const int64_t start = ~(int64_t)rec;
                                const BVH2Node &nd = nodes[i];
                                for (int32_t j = 0; j < nd.count; ++j) {
                                        const int64_t k = (int64_t)nd.offset + j;
                                        const PrimId_C &pid = prims[k];
                                        TriRec &t = tris[(size_t)(start + j)];
                                        std::memset(&t, 0, sizeof t);
                                        int64_t mesh = -1, base = -1;
                                        if (pid.type == 0) {
                                                mesh = pid.id1;
                                                base = pid.id2;
                                        } else if ((pid.type == 1 || pid.type == 2 || pid.type == 3) && pid.id2 != -1) {
                                                mesh = pid.id2 >> 32;
                                                base = (pid.id2 & 0xFFFFFFFFLL) * 3;
                                        }
                                        int32_t flags = 0;
                                        if (mesh < 0 && base < 0) {
                                                flags = kTriSkip;
                                        } else {
                                                if (mesh < 0 || mesh >= nMeshes || base < 0 || base + 2 >= vtxCounts[mesh]) {
                                                        set_error("gz_gpu_upload_scene: primId %lld references mesh %lld "
                                                                  "index %lld outside the mesh arrays", (long long)k,
                                                                  (long long)mesh, (long long)base);
                                                        free_scene(s);
                                                        return nullptr;
                                                }
                                                const TriangleMesh_C &m = meshes[mesh];
                                                float v[3][3];
                                                for (int c = 0; c < 3; ++c) {
                                                        const int64_t vi = m.vertexIndices[base + c];
                                                        if (vi < 0 || vi * 4 + 2 >= ptsCounts[mesh]) {
                                                                set_error("gz_gpu_upload_scene: vertex index %lld outside mesh %lld "
                                                                          "(%lld floats)", (long long)vi, (long long)mesh,
                                                                          (long long)ptsCounts[mesh]);
                                                                free_scene(s);
                                                                return nullptr;
                                                        }
                                                        v[c][0] = m.points[vi * 4];
                                                        v[c][1] = m.points[vi * 4 + 1];
                                                        v[c][2] = m.points[vi * 4 + 2];
                                                }
                                                t.p0 = make_float4(v[0][0], v[0][1], v[0][2], 0.0f);
                                                t.p1 = make_float4(v[1][0], v[1][1], v[1][2], 0.0f);
                                                t.p2 = make_float4(v[2][0], v[2][1], v[2][2], 0.0f);
                                        }
                                        t.p0.w = as_float(j == 0 ? nd.count : 0);
                                        t.p1.w = as_float((int32_t)k);
                                        t.p2.w = as_float(flags);
                                }
                        }
                }
        } catch (const std::bad_alloc &) {
                if (gz_last_error()[0] == 0) set_error("gz_gpu_upload_scene: out of host memory");
                free_scene(s);
                return nullptr;
        }

        s->nodeRecords = (int64_t)pairs.size();
        bool ok = hip_ok(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate") &&
                  upload(&s->dNodes, pairs.data(), (int64_t)pairs.size(), s, "upload nodes") &&
                  upload(&s->dTris, tris.data(), (int64_t)tris.size(), s, "upload triangles") &&
                  scene_upload_common(s, prims, nPrims, meshes, nMeshes, ptsCounts, faceCounts, vtxCounts, materials,
                                      nMaterials);
        if (!ok) {
                free_scene(s);
                return nullptr;
        }
        return s;
}

extern "C" void gz_gpu_free_scene(void *handle) {
        if (!handle) return;
        free_scene(static_cast<Scene *>(handle));
}

extern "C" int gz_gpu_scene_info(void *handle, GzSceneInfo *info) {
        clear_error();
        if (!handle || !info) {
                set_error("gz_gpu_scene_info: null argument");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        info->nodeCount = s->nodeCount;
        info->interiorCount = s->interiorCount;
        info->primCount = s->primCount;
        info->triRecordCount = s->triCount;
        info->meshCount = s->meshCount;
        info->materialCount = s->materialCount;
        info->deviceBytes = s->deviceBytes;
        info->maxDepth = s->maxDepth;
        info->rootIsLeaf = (!s->empty && s->rootDesc < 0) ? 1 : 0;
        return GZ_OK;
}

extern "C" int gz_gpu_check_traversal(void *handle, void *stream) {
        clear_error();
        if (!handle) {
                set_error("gz_gpu_check_traversal: null handle");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return check_stream(s, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_set_trace_config(void *handle, const GzTraceConfig *config) {
        clear_error();
        if (!handle || !config) {
                set_error("gz_gpu_set_trace_config: null argument");
                return GZ_E_INVALID;
        }
        if (config->persistent < 0 || config->persistent > 3 ||
            (config->ldsStack != 0 && config->ldsStack != 8 && config->ldsStack != 16) || config->refillLanes < 1 ||
            config->refillLanes > 64 || config->gridScale < 0 || config->gridScale > 16 || config->chunk < 1 ||
            config->chunk > 4096 || config->binMinRays < 0 || config->recordTimes < 0 || config->recordTimes > 1) {
                set_error("gz_gpu_set_trace_config: persistent 0..3, ldsStack 0/8/16, refillLanes 1..64, "
                          "gridScale 0..16, chunk 1..4096, binMinRays >= 0, recordTimes 0/1");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        s->traceConfig = *config;
        return GZ_OK;
}

extern "C" int gz_gpu_get_trace_config(void *handle, GzTraceConfig *config) {
        clear_error();
        if (!handle || !config) {
                set_error("gz_gpu_get_trace_config: null argument");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        *config = s->traceConfig;
        return GZ_OK;
}
