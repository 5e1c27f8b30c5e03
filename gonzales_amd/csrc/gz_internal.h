// Internal declarations shared by the library's translation units (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>

#include "gonzales_amd.h"

namespace gz {

// Child-pair record of one interior BVH2 node (64 bytes, one half of a 128-byte cache line).
// Holds both children's boxes, which the reference reads from two separate 32-byte nodes per
// interior visit (kernel.mojo:314-327), and both children's descriptors:
//   desc >= 0 : the child is interior, desc is its record index;
//   desc <  0 : the child is a leaf, ~desc is the index of its first triangle record.
struct alignas(16) NodePair {
        float4 a;  // left.minX, left.minY, left.minZ, left.maxX
        float4 b;  // left.maxY, left.maxZ, right.minX, right.minY
        float4 c;  // right.minZ, right.maxX, right.maxY, right.maxZ
        int4 d;    // left desc, right desc, 0, 0
};
static_assert(sizeof(NodePair) == 64, "NodePair");

// One primitive of a leaf, pre-gathered (48 bytes): the three vertices the reference fetches
// through PrimId -> TriangleMesh_C -> vertexIndices -> points (kernel.mojo:266-300).
//   p0.w : (first record of a leaf only) number of records in the leaf, as int bits
//   p1.w : index into primIds of this primitive (the reference's hitIndex), as int bits
//   p2.w : flags as int bits; bit 0 = skip (the reference `continue`s: id2 == -1 or an
//          unknown type, kernel.mojo:273-277)
struct alignas(16) TriRec {
        float4 p0;
        float4 p1;
        float4 p2;
};
static_assert(sizeof(TriRec) == 48, "TriRec");

enum : int { kTriSkip = 1 };

// Device-resident mesh view for the shading kernels: TriangleMesh_C's pointers plus the element
// counts passed at upload, so a kernel can reject an index the reference would read past.
struct DevMesh {
        const float *points;
        const int64_t *faceIndices;
        const int64_t *vertexIndices;
        int64_t pointFloats;
        int64_t vertexIndexCount;
};

// Light and material tables of the MIS integrator (gz_mis.hip), set by gz_gpu_set_mis_scene.
struct MisScene;
void release_mis(MisScene *m);
// Device path queues of the MIS trace, per stream (gz_mis.hip).
struct MisQueues;
void release_mis_queues(MisQueues *m);
struct Scene;
// gz_gpu_mis_trace_device over one scene with the queue storage taken from *queues (created on
// first use, grow-only, freed with release_mis_queues) instead of the handle's own, and traversal
// overflows raised in *err; for a caller that renders several scenes one after another on one
// stream and keeps one storage for all of them.
int mis_trace_shared(Scene *s, GzPath *d_paths, int64_t count, const GzMisParams *params, hipStream_t stream,
                     MisQueues **queues, int *err);

// Per-stream device state of a scene handle (allocated zeroed at the stream's first use):
//   dErr     : bit 0 = a traversal stack overflowed, bit 1 = a shading kernel met a bad index;
//              raised by the launches on the stream, read and cleared by
//              gz_gpu_check_traversal(stream), so concurrent streams never see each other's flag;
//   dCounter : the ray counters of the persistent traversal kernels (one, or one per XCD: eight),
//              reset by a kernel at the start of each launch; launches on one stream run in order,
//              so they can share them.
//   dBin     : the ray binning's storage (sorted batch, keys, counts), grow-only, binBytes bytes;
//   ev       : with recordTimes, events before the binning, before and after the traversal kernel
//              of the last launch (created at first use); timed: whether they hold one.
struct StreamSlot {
        int *dErr = nullptr;
        unsigned long long *dCounter = nullptr;
        uint32_t *dBin = nullptr;
        size_t binBytes = 0;
        hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
        bool timed = false;
};

struct Scene {
        std::mutex mu;
        std::map<hipStream_t, StreamSlot> slots;
        uint64_t misGeneration = 0;  // bumped by every gz_gpu_set_mis_scene (batches check it)
        MisScene *mis = nullptr;
        MisQueues *misQueues = nullptr;
        GzMisLaunch misLaunch{1, 256, 1, 256, 3, {0, 0, 0}};
        bool misLaunchSet = false;  // false: misLaunch follows the scene (default_mis_launch)
        int device = 0;
        hipStream_t stream = nullptr;  // used by the synchronous host-buffer entry points

        int64_t nodeCount = 0, interiorCount = 0, primCount = 0, triCount = 0;
        int64_t interfacePrims = 0;  // surface primitives with materialIndex < 0 (interfaces)
        int64_t meshCount = 0, materialCount = 0;
        int32_t maxDepth = 0;
        int32_t rootDesc = 0;
        bool empty = true;  // no BVH nodes: every query misses

        NodePair *dNodes = nullptr;
        int64_t nodeRecords = 0;  // child-pair records allocated at dNodes
        TriRec *dTris = nullptr;
        uint32_t *dBuildOrder = nullptr;  // device build: caller's primitive index per leaf position
        PrimId_C *dPrims = nullptr;
        // The three vertices of every primId in primId order (3 float4 per primitive, p0.w = 1 for a
        // primitive whose mesh or vertex indices are out of range), gathered at the first
        // gz_gpu_set_mis_scene: the shading kernels read a hit's vertices with one load, not
        // through the primId -> mesh -> vertex index -> point chain.
        float4 *dPrimTris = nullptr;
        DevMesh *dMeshes = nullptr;
        Material_C *dMaterials = nullptr;
        void **meshBuffers = nullptr;  // per-mesh device arrays to free
        int64_t meshBufferCount = 0;
        int64_t deviceBytes = 0;

        // Grow-only staging for the synchronous entry points.
        void *dStage = nullptr;
        size_t stageBytes = 0;

        // Traversal launch configuration.
        GzTraceConfig traceConfig{3, 16, 32, 1, 128, 1 << 18, 0, 0};
        bool boundsSafe = false;  // every node box coordinate is finite with magnitude <= 2^40
        int cuCount = 256;
};

void set_error(const char *fmt, ...);
void clear_error();
bool hip_ok(hipError_t e, const char *what);
// Allocation-failure injection (gz_error.cpp): alloc_fault_arm() at the start of an entry point
// reads GZ_FAIL_ALLOC_AFTER; alloc_fault(what) before each allocation returns true (error set) on
// the armed one.
void alloc_fault_arm();
bool alloc_fault(const char *what);

// Makes the scene's device current; returns false (error set) on failure.
bool enter_device(Scene *s);
// The slot of `stream` on this handle (the caller holds s->mu); nullptr (error set) on failure.
StreamSlot *stream_slot(Scene *s, hipStream_t stream);
// Reads and clears the error flag of `stream` (synchronizes the stream): GZ_OK, GZ_E_OVERFLOW or
// GZ_E_HIP.
int check_stream(Scene *s, hipStream_t stream);
// Ensures the staging buffer holds at least `bytes`.
bool ensure_stage(Scene *s, size_t bytes);

// Uploads primIds, materials and the per-mesh arrays, creates the handle's stream, error flag
// and counters (gz_scene.hip); shared by the host-BVH upload and the device build.
bool scene_upload_common(Scene *s, const PrimId_C *prims, int64_t nPrims, const TriangleMesh_C *meshes,
                         int64_t nMeshes, const int64_t *ptsCounts, const int64_t *faceCounts,
                         const int64_t *vtxCounts, const Material_C *materials, int64_t nMaterials);
// Frees a handle and everything it owns.
void scene_release(Scene *s);

// Kernel launchers (gz_traverse.hip).
int launch_trace(Scene *s, const Ray_C *d_rays, const float *d_tMax, int64_t n, GzHit *d_hits,
                 uint8_t *d_occ, GzTraceStats *d_stats, bool anyHit, hipStream_t stream);
int launch_expand(Scene *s, const GzHit *d_hits, const float *d_tMax, int64_t n, Intersection_C *d_out,
                  hipStream_t stream);

}  // namespace gz
