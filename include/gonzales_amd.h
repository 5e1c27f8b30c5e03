/*
 * gonzales_amd.h — C-ABI of the MI355X (gfx950) hot path of the Gonzales path tracer.
 *
 * Drop-in boundary. The structs are byte-identical to the reference's
 * Sources/mojoKernel/include/mojo_kernel.h:4-58, so the Swift caller
 * (Accelerators/BoundingHierarchy.swift, Core/Tile.swift) keeps its marshalling code and only
 * swaps the imported module and the `mojo_gpu_*` symbol prefix for `gz_gpu_*`.
 *
 * Conventions of this library:
 *   - every entry point is `extern "C"`, takes plain pointers and sizes, no HIP or torch types;
 *     a stream is passed as `void *` (a hipStream_t, NULL = the legacy default stream);
 *   - `*_device` entry points take device pointers and are asynchronous on the given stream;
 *     every other entry point takes host pointers and is synchronous (results ready on return),
 *     as the reference's are (kernel.mojo:608-673, 788-846);
 *   - the `gz_gpu_*` functions that mirror `mojo_gpu_*` keep their `void`/pointer returns; on
 *     failure they leave outputs untouched (kernel.mojo:672-673) and record a message that
 *     gz_last_error() returns; new entry points return GZ_OK (0) or a negative GZ_E_* code;
 *   - a scene handle may be used from several host threads at once: each call serialises on an
 *     internal per-handle lock (the reference shares one DeviceContext across tile tasks,
 *     TileRenderer.swift:77-107).
 */
#ifndef GONZALES_AMD_H
#define GONZALES_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------- */
/* Structs shared with the reference (mojo_kernel.h:4-58).                                   */
/* ---------------------------------------------------------------------------------------- */

/* mojo_kernel.h:4-9. type: 0 triangle (id1 mesh, id2 = 3*triangle), 1 geometric primitive,
 * 2 transformed primitive, 3 area light; for 1..3 id2 = mesh<<32 | triangle, or -1 for a
 * non-triangle shape (BoundingHierarchy.swift:21-58, 479-484). 32 bytes. */
struct PrimId_C {
        int64_t id1;
        int64_t id2;
        int64_t materialIndex;
        int8_t type;
};

/* mojo_kernel.h:11-15. type: 0 none, 1 diffuse, 2 area light. 28 bytes. */
struct Material_C {
        int8_t type;
        float albedo[3];
        float emission[3];
};

/* mojo_kernel.h:17-21. points: 4 floats per vertex (x, y, z, 1); indices are int64. */
struct TriangleMesh_C {
        const float *points;
        const int64_t *faceIndices;
        const int64_t *vertexIndices;
};

/* mojo_kernel.h:23-26. 24 bytes. */
struct Ray_C {
        float orgX, orgY, orgZ;
        float dirX, dirY, dirZ;
};

/* mojo_kernel.h:28-34. 48 bytes; a miss carries PrimId(-1, -1, 0, 0), tHit = tMax. */
struct Intersection_C {
        struct PrimId_C primId;
        float tHit;
        float u;
        float v;
        int8_t hit;
};

/* mojo_kernel.h:36-44. 88 bytes. */
struct PathState_C {
        struct Ray_C ray;
        float throughput[3];
        float estimate[3];
        float albedo[3];
        uint64_t pcgState;
        uint64_t pcgInc;
        int8_t active;
};

/* mojo_kernel.h:46-51. Depth-first BVH2: left child = index + 1, right child = offset;
 * count == 0 marks an interior node, count > 0 a leaf over primIds[offset, offset+count). */
struct BVH2Node {
        float boundsMinX, boundsMinY, boundsMinZ;
        float boundsMaxX, boundsMaxY, boundsMaxZ;
        int32_t offset;
        int32_t count;
};

/* mojo_kernel.h:53-58 (the C definition, 32 bytes; the Mojo side declares two more fields,
 * kernel.mojo:208-215, which no C caller fills). */
struct SceneDescriptor2_C {
        const struct BVH2Node *bvh2Nodes;
        const struct PrimId_C *primIds;
        const struct TriangleMesh_C *meshes;
        int64_t meshCount;
};

/* ---------------------------------------------------------------------------------------- */
/* Types introduced by this library.                                                         */
/* ---------------------------------------------------------------------------------------- */

/* Compact closest-hit record written by the device-resident traversal (16 bytes).
 * prim is the index into the scene's primIds array (the reference's `hitIndex`,
 * kernel.mojo:306) or -1 on a miss; on a miss t = tMax and u = v = 0 (kernel.mojo:360-364). */
struct GzHit {
        int32_t prim;
        float t;
        float u;
        float v;
};

/* Per-ray traversal work counters (used for the algorithmic-bytes figure):
 * interior = interior nodes whose two child boxes were tested, tris = triangle tests run. */
struct GzTraceStats {
        uint32_t interior;
        uint32_t tris;
};

/* Scene summary filled by gz_gpu_scene_info. */
struct GzSceneInfo {
        int64_t nodeCount;      /* BVH2 nodes passed at upload */
        int64_t interiorCount;  /* child-pair records on the device */
        int64_t primCount;      /* primIds */
        int64_t triRecordCount; /* 48-byte triangle records on the device (leaf order) */
        int64_t meshCount;
        int64_t materialCount;
        int64_t deviceBytes;    /* device memory held by the handle */
        int32_t maxDepth;       /* deepest root-to-leaf path (edges) */
        int32_t rootIsLeaf;
};

enum {
        GZ_OK = 0,
        GZ_E_INVALID = -1,    /* bad argument or malformed scene */
        GZ_E_NO_DEVICE = -2,  /* no usable gfx950 device */
        GZ_E_HIP = -3,        /* a HIP runtime call failed */
        GZ_E_OVERFLOW = -4,   /* a traversal needed more than GZ_STACK_CAPACITY stack entries */
        GZ_E_ALLOC = -5       /* device or host allocation failed */
};

/* Traversal stack capacity per ray, the reference's InlineArray[Int, 64] (kernel.mojo:250).
 * The reference does not check it; this library stops the ray and reports GZ_E_OVERFLOW. */
#define GZ_STACK_CAPACITY 64

/* ---------------------------------------------------------------------------------------- */
/* Entry points that replace the reference's GPU half of mojo_kernel.h.                       */
/* ---------------------------------------------------------------------------------------- */

/* Replaces mojo_gpu_available (mojo_kernel.h:65; kernel.mojo:407-410). True when a gfx950
 * device is visible and the kernels in this library load on it. */
bool gz_gpu_available(void);

/* Replaces mojo_gpu_upload_scene (mojo_kernel.h:70-75; kernel.mojo:430-587). Same arguments:
 * per-mesh point counts are in float elements (4 per vertex), index counts in int64 elements
 * (BoundingHierarchy.swift:160-172). Copies and re-lays out the scene into device memory
 * (64-byte child-pair node records, 48-byte triangle records in leaf order); keeps no host
 * pointer. Returns NULL on failure (kernel.mojo:583-585) with gz_last_error() set. */
void *gz_gpu_upload_scene(const struct BVH2Node *bvh2Nodes, int64_t bvh2NodesCount,
                          const struct PrimId_C *primIds, int64_t primIdsCount,
                          const struct TriangleMesh_C *meshes, int64_t meshCount,
                          const int64_t *meshPointsCounts, const int64_t *meshFaceIndicesCounts,
                          const int64_t *meshVertexIndicesCounts, const struct Material_C *materials,
                          int64_t materialCount);

/* Replaces mojo_gpu_traverse_batch (mojo_kernel.h:77-78; kernel.mojo:608-673): closest hit of
 * `count` host rays with per-ray tMax, results into host `results`. NULL handle or count == 0
 * is a no-op (kernel.mojo:616-622). */
void gz_gpu_traverse_batch(void *handle, const struct Ray_C *rays, const float *tMaxValues,
                           int64_t count, struct Intersection_C *results);

/* Replaces mojo_gpu_shade_batch (mojo_kernel.h:80-81; kernel.mojo:788-846): the diffuse
 * wavefront shade of the reference's Tile.render loop (kernel.mojo:676-773), in place. */
void gz_gpu_shade_batch(void *handle, struct PathState_C *paths, int64_t count,
                        const struct Intersection_C *intersections);

/* Replaces mojo_gpu_free_scene (mojo_kernel.h:91; kernel.mojo:848-854). NULL is a no-op. */
void gz_gpu_free_scene(void *handle);

/* ---------------------------------------------------------------------------------------- */
/* Entry points that replace the reference's CPU half of mojo_kernel.h (host C++, no device). */
/* Re-entrant, no global state: callers run them concurrently, one per tile task, as           */
/* TileRenderer.swift:77-107 does. The 64-entry traversal stack is checked (the reference's is   */
/* not, kernel.mojo:250): a ray that would overflow it stops with the closest hit found so far   */
/* and gz_last_error() on the calling thread says so; otherwise gz_last_error() is "".          */
/* ---------------------------------------------------------------------------------------- */

/* Replaces mojo_traverse_bvh2 (mojo_kernel.h:60-62; kernel.mojo:369-373): closest hit of one ray
 * over the BVH2 / PrimId / TriangleMesh_C arrays the descriptor points at (traverse_bvh2_core,
 * kernel.mojo:223-364). A descriptor with no nodes (bvh2Nodes NULL) gives a miss. */
void gz_traverse_bvh2(const struct SceneDescriptor2_C *scene, const struct Ray_C *ray, float tMax,
                      struct Intersection_C *result);

/* Replaces mojo_cpu_traverse_batch (mojo_kernel.h:84-85; kernel.mojo:378-390): the same for
 * `count` rays with per-ray tMax, one after another on the calling thread. */
void gz_cpu_traverse_batch(const struct SceneDescriptor2_C *scene, const struct Ray_C *rays,
                           const float *tMaxValues, int64_t count, struct Intersection_C *results);

/* Replaces mojo_cpu_shade_batch (mojo_kernel.h:87-89; kernel.mojo:393-403): the diffuse shade
 * step of Tile.render (shade_core, kernel.mojo:676-773) for `count` paths, in place. */
void gz_cpu_shade_batch(struct PathState_C *paths, int64_t count, const struct Intersection_C *intersections,
                        const struct TriangleMesh_C *meshes, const struct Material_C *materials);

/* ---------------------------------------------------------------------------------------- */
/* New entry points (status-returning).                                                        */
/* ---------------------------------------------------------------------------------------- */

/* Message of the last failure on the calling thread ("" if none). */
const char *gz_last_error(void);

/* Fills *info for a handle returned by gz_gpu_upload_scene. */
int gz_gpu_scene_info(void *handle, struct GzSceneInfo *info);

/* Occlusion of `count` host rays: occluded[i] = 1 iff the closest-hit traversal of ray i with
 * tMaxValues[i] reports a hit, which is what BoundingHierarchy.intersect(scene:ray:tHit:) ->
 * Bool answers (BoundingHierarchy.swift:348-384) and Visibility.occluded consumes
 * (Core/Visibility.swift:5-24). Runs the any-hit kernel, which stops at the first accepted
 * triangle. Synchronous, host buffers. */
int gz_gpu_occluded_batch(void *handle, const struct Ray_C *rays, const float *tMaxValues,
                          int64_t count, uint8_t *occluded);

/* No counterpart in the reference. For callers that cannot allocate device memory (serving
 * loops and harnesses that forbid hipMalloc): gz_gpu_traverse_batch's closest hits, the same
 * records, with every device buffer the call needs (rays, tMax, hits, records, the error flag and
 * the ray counter) in a fixed 46 MiB array that the library holds on each GPU it runs on until
 * the process ends, instead of the handle's grow-only staging. At most GZ_NOALLOC_MAX_RAYS rays
 * per call: more is GZ_E_INVALID and nothing runs. Calls on one GPU share its array and run one
 * at a time (a lock per GPU); the array keeps the last call's contents. Synchronous, host
 * buffers; returns GZ_OK, GZ_E_OVERFLOW (no results written), GZ_E_INVALID or GZ_E_HIP. */
#define GZ_NOALLOC_MAX_RAYS 524288
int gz_gpu_traverse_batch_noalloc(void *handle, const struct Ray_C *rays, const float *tMaxValues,
                                  int64_t count, struct Intersection_C *results);

/* Device-resident closest hit: d_rays, d_tMax, d_hits are device pointers; asynchronous on
 * `stream`. The batch form behind mojo_gpu_traverse_batch, without the host round trip. */
int gz_gpu_traverse_closest_device(void *handle, const struct Ray_C *d_rays, const float *d_tMax,
                                   int64_t count, struct GzHit *d_hits, void *stream);

/* Device-resident any-hit (occlusion): d_occluded[i] in {0, 1}. Asynchronous on `stream`. */
int gz_gpu_traverse_any_device(void *handle, const struct Ray_C *d_rays, const float *d_tMax,
                               int64_t count, uint8_t *d_occluded, void *stream);

/* Device-resident closest hit that also counts per-ray work (d_stats may be NULL for hits only).
 * Same results as gz_gpu_traverse_closest_device; slower, used for the byte accounting. */
int gz_gpu_traverse_closest_stats_device(void *handle, const struct Ray_C *d_rays,
                                         const float *d_tMax, int64_t count, struct GzHit *d_hits,
                                         struct GzTraceStats *d_stats, void *stream);

/* Device-resident any-hit with per-ray work counters (counted up to the early exit). */
int gz_gpu_traverse_any_stats_device(void *handle, const struct Ray_C *d_rays, const float *d_tMax,
                                     int64_t count, uint8_t *d_occluded,
                                     struct GzTraceStats *d_stats, void *stream);

/* Expands compact hits into the reference's 48-byte Intersection_C records (primIds gathered on
 * the device; misses become PrimId(-1, -1, 0, 0), tHit = tMax). d_tMax supplies tMax for misses.
 * Asynchronous on `stream`. */
int gz_gpu_expand_hits_device(void *handle, const struct GzHit *d_hits, const float *d_tMax,
                              int64_t count, struct Intersection_C *d_results, void *stream);

/* Launch configuration of the traversal kernels (no reference counterpart: the reference launches
 * one thread per ray in blocks of 256, kernel.mojo:649-662). Results do not depend on it.
 *   persistent  : 0 = one lane per ray; 1 = persistent lanes refilled from an atomic ray
 *                 counter; 2 = persistent lanes that trace a fixed stride of rays; 3 = as 1 with
 *                 the batch cut into eight ranges, one per XCD, each with its own counter (a
 *                 wave draws from its XCD's range first, then from the others);
 *   ldsStack    : stack entries kept in LDS per lane (0, 8 or 16; the rest in private memory);
 *   refillLanes : idle lanes (1..64) that trigger a refill of a wave's lanes (modes 1, 3);
 *   gridScale   : persistent grid = gridScale x the blocks resident per CU x CUs (0 = 1);
 *   chunk       : rays a wave takes from a counter per atomic (modes 1, 3; 1..4096); also the
 *                 queue entries a wave of the MIS full wavefront's path-ray kernel takes per
 *                 atomic (default 128);
 *   binMinRays  : batches of at least this many rays (and at least 1) are traced in the order of
 *                 their origin's cell in a 16^3 grid over the scene box (a counting sort on the
 *                 device, into storage the handle keeps per stream: 34 bytes per ray plus 16 KiB
 *                 per block of up to 65536 rays), so that the waves in flight read overlapping
 *                 parts of the tree; each ray's result still goes to its own index. 0 = always in
 *                 the given order;
 *   recordTimes : 1 = record HIP events around the binning and the traversal kernel of each
 *                 launch, for gz_gpu_trace_times (0 = none). */
struct GzTraceConfig {
        int32_t persistent;
        int32_t ldsStack;
        int32_t refillLanes;
        int32_t gridScale;
        int32_t chunk;
        int32_t binMinRays;
        int32_t recordTimes;
        int32_t reserved;
};

int gz_gpu_set_trace_config(void *handle, const struct GzTraceConfig *config);
int gz_gpu_get_trace_config(void *handle, struct GzTraceConfig *config);

/* With recordTimes set: waits for `stream` and returns in *binMs and *traceMs the milliseconds of
 * the binning kernels (0 if the batch was traced as given) and of the traversal kernel of the last
 * closest-hit or any-hit launch on `stream`. GZ_E_INVALID if none was recorded. No counterpart in
 * the reference (for per-kernel rates). */
int gz_gpu_trace_times(void *handle, void *stream, float *binMs, float *traceMs);

/* Waits for `stream` and returns GZ_E_OVERFLOW if any traversal launched on `stream` since the
 * last call stopped on a full stack (GZ_E_INVALID if a shading kernel met an index out of range),
 * and clears the flag; else GZ_OK. Each stream a handle is used with has its own flag (and its
 * own ray counter for the persistent traversal), so calls on concurrent streams never read or
 * clear each other's. No counterpart in the reference (its stack is unchecked,
 * kernel.mojo:250). */
int gz_gpu_check_traversal(void *handle, void *stream);

/* Device-resident diffuse shade: the kernel behind gz_gpu_shade_batch without the host round
 * trip (kernel.mojo:676-786). d_paths is updated in place. Asynchronous on `stream`. */
int gz_gpu_shade_device(void *handle, struct PathState_C *d_paths, int64_t count,
                        const struct Intersection_C *d_isects, void *stream);

/* ---------------------------------------------------------------------------------------- */
/* MIS path tracing: the bounce of VolumePathIntegrator (Integrator/VolumePathIntegrator.swift) */
/* ---------------------------------------------------------------------------------------- */

/* Material parameters the MIS integrator shades with (the reference's Material/ types with
 * constant textures), read by the kernel as each type's getBsdf does:
 *   GZ_MAT_DIFFUSE           reflectance (Diffuse.swift)
 *   GZ_MAT_CONDUCTOR         conductorEta, conductorK, roughness (alpha = max(r, 1e-3), never
 *                            remapped; Conductor.swift:3-17); reflectance is white
 *   GZ_MAT_DIELECTRIC        eta, roughness (alpha = sqrt(r) when remapped; Dielectric.swift:44-54)
 *   GZ_MAT_COATED_CONDUCTOR  eta and roughness of the coating, conductorRoughness, conductorEta,
 *                            conductorK, reflectance, thickness, albedo, g, maxDepth, nSamples
 *                            (alphas remapped when asked, then max(., 1e-3); CoatedConductor.swift)
 * flags: GZ_MAT_REMAP_ROUGHNESS. */
struct GzMaterial {
        int32_t type;
        int32_t flags;
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
};
enum { GZ_MAT_REMAP_ROUGHNESS = 1 };

/* One triangle area light (AreaLight.swift:3-49): constant radiance on triangle `triangle` of
 * mesh `mesh`; its index is the id1 of its type-3 PrimId. */
struct GzAreaLight {
        float brightness[3];
        int32_t mesh;
        int64_t triangle;
};

/* Path state of the wavefront MIS integrator (PathState, VolumePathIntegrator.swift:19-31, plus
 * the Z-Sobol sampler state of the path and of the light chooser). 112 bytes. */
struct GzPath {
        struct Ray_C ray;
        float throughput[3];
        float estimate[3];
        float albedo[3];
        float firstNormal[3];
        float filterWeight;
        int32_t pixelX, pixelY, sampleIndex;
        int32_t dim;      /* next dimension of the path's Z-Sobol sampler */
        int32_t lightDim; /* next dimension of the light chooser's sampler (Morton index 0) */
        int32_t bounce;
        int32_t flags;    /* GZ_PATH_ACTIVE | GZ_PATH_SPECULAR */
        int32_t pad[2];
};

struct GzMisParams {
        int32_t maxDepth;       /* Integrator "maxdepth" (RenderConfiguration.swift:143-146) */
        int32_t flags;          /* GZ_MIS_FIX_LIGHT_PDF | GZ_MIS_FIX_SKY_DIRECTION */
        int64_t zsSeed;         /* Z-Sobol seed and constants, as gz_zsobol_params returns them */
        int32_t zsLog2Spp;
        int32_t zsBase4Digits;
        uint64_t rrSeed;        /* Russian-roulette stream */
};

enum { GZ_PATH_ACTIVE = 1, GZ_PATH_SPECULAR = 2 };
enum { GZ_MAT_DIFFUSE = 1, GZ_MAT_CONDUCTOR = 2, GZ_MAT_DIELECTRIC = 3, GZ_MAT_COATED_CONDUCTOR = 4 };
/* Off (the default) reproduces the reference: the light pdf of a BSDF-sampled direction toward a
 * triangle light is 0 (Sampleable.swift:57-58 passes tHit = 0), so BSDF samples of area lights
 * get MIS weight 0. On, the pdf is computed from the chosen triangle.
 * GZ_MIS_FIX_SKY_DIRECTION: off reproduces InfiniteLight.equalAreaSquareToSphere
 * (InfiniteLight.swift:112-127), whose x and y lack the factor sqrt(2 - r^2), so its directions
 * are shorter than 1 away from the poles and equator; on, the factor is applied (unit directions,
 * uniform on the sphere). */
enum { GZ_MIS_FIX_LIGHT_PDF = 1, GZ_MIS_FIX_SKY_DIRECTION = 2 };

/* Sets the shading scene of a handle: materials indexed by PrimId.materialIndex (area-light
 * surfaces shade with materials[0], as AreaLight.computeSurfaceInteraction leaves the index at 0),
 * the triangle area lights indexed by their PrimId.id1, and an optional constant infinite light
 * (skyRadiance: 3 floats, NULL = none). Lights are chosen by power (PowerLightSampler.swift):
 * the area lights in index order, then the sky. */
int gz_gpu_set_mis_scene(void *handle, const struct GzMaterial *materials, int32_t materialCount,
                         const struct GzAreaLight *lights, int32_t lightCount, const float *skyRadiance);

/* One bounce (oneBounce + shadeBounce, VolumePathIntegrator.swift:373-445) of every active path:
 * closest hit, emission, one light sample with MIS (shadow ray + BSDF ray), continuation and
 * Russian roulette. Paths that stop clear GZ_PATH_ACTIVE. Asynchronous on `stream`. */
int gz_gpu_mis_bounce_device(void *handle, struct GzPath *d_paths, int64_t count,
                             const struct GzMisParams *params, void *stream);

/* Runs bounces until every path has stopped (VolumePathIntegrator.bounces, :448-463).
 * Asynchronous on `stream`. With a queued launch configuration (the default) bounce 0 runs over
 * all paths and every later bounce over a device queue of the paths still active; the handle
 * keeps the queue storage per stream (grow-only: the first call with a larger count allocates,
 * so make one such call before capturing these launches in a graph). */
int gz_gpu_mis_trace_device(void *handle, struct GzPath *d_paths, int64_t count,
                            const struct GzMisParams *params, void *stream);

/* Launch configuration of gz_gpu_mis_trace_device, _trace_stats_device and _adjoint_device
 * (results never depend on it; no counterpart in the reference):
 *   queued    : 1 = persistent waves over device path queues, compacted after every bounce by a
 *               wave ballot and prefix popcount (one atomic per wave); 2 = the same queues with
 *               each bounce split in a closest-hit kernel (lanes refilled from the queue, hits
 *               sorted by material class into two queues) and one shade kernel per class, the
 *               diffuse one small; 3 = the full wavefront: as 2, with the shade kernels tracing
 *               nothing (they queue each path's shadow ray and BSDF ray), an any-hit and a
 *               closest-hit kernel trace those two ray queues, and a finish kernel per class adds
 *               the two strategies' terms; 0 = every bounce launches one lane per path of all
 *               `count` (stopped paths idle in their waves). The adjoint replays queue 1 for 1-3;
 *   block     : threads per workgroup, 64, 128 or 256;
 *   gridScale : persistent grid = gridScale x the workgroups resident per CU x CUs (1..16);
 *   chunk     : queue entries a wave takes per atomic, a multiple of 64 in 64..4096;
 *   drainLevel: 0, or 1..64 with a queued mode: the paths left at that bounce level run all their
 *               remaining bounces in one launch (a lane per path), in place of one launch
 *               sequence per level for the few paths that Russian roulette leaves (the
 *               adjoint's replay drains from the same level);
 *   reserved  : zero.
 * Until gz_gpu_set_mis_launch sets one, a handle uses {1, 256, 1, 256, 3} if every material is
 * diffuse and it has at most 65536 primitives, else {3, 256, 1, 64, 3} (gz_gpu_get_mis_launch
 * returns the one in use). */
struct GzMisLaunch {
        int32_t queued;
        int32_t block;
        int32_t gridScale;
        int32_t chunk;
        int32_t drainLevel;
        int32_t reserved[3];
};

int gz_gpu_set_mis_launch(void *handle, const struct GzMisLaunch *launch);
int gz_gpu_get_mis_launch(void *handle, struct GzMisLaunch *launch);

/* Frees the path-queue storage the MIS entry points keep per handle and stream (about 170 bytes
 * per path of the largest launch, kept so that later launches of that size allocate nothing);
 * the next MIS launch allocates it again. Waits for the work queued on those streams first. No
 * counterpart in the reference, which renders one scene per process: for callers that render
 * many large scenes in turn. */
int gz_gpu_release_mis_queues(void *handle);


/* Traversal work of one path, summed over its bounces and all rays of each bounce (closest hit,
 * shadow ray, BSDF ray): interior-node visits, triangle tests, bounces run. */
struct GzPathStats {
        uint32_t interior;
        uint32_t tris;
        uint32_t bounces;
};

/* gz_gpu_mis_trace_device that also adds each path's work to d_stats[i] (zero it first). For
 * measurement: the counting costs a little time. No counterpart in the reference. */
int gz_gpu_mis_trace_stats_device(void *handle, struct GzPath *d_paths, int64_t count,
                                  const struct GzMisParams *params, struct GzPathStats *d_stats, void *stream);

/* Path-replay adjoint input of one path: its forward radiance (the estimate the forward trace
 * left) and dLoss/d(radiance) of that path sample. */
struct GzPathAdjoint {
        float radiance[3];
        float adjoint[3];
};

/* Path-replay backward pass (no counterpart in the reference, which has no derivatives):
 * replays every bounce of the `count` paths from the same starting records the forward trace
 * used (same samples, streams and decisions) and adds dLoss/dtheta to d_grad, a
 * [materialCount][4] double array (zero it first): entries 0..2 the reflectance of
 * GZ_MAT_DIFFUSE materials, entry 3 the (isotropic) roughness of GZ_MAT_CONDUCTOR materials
 * with roughness > 1e-3. Directions, MIS weights and Russian-roulette decisions are held fixed
 * (detached sampling). The path records are overwritten as by gz_gpu_mis_trace_device.
 * Asynchronous on `stream`. */
int gz_gpu_mis_adjoint_device(void *handle, struct GzPath *d_paths, const struct GzPathAdjoint *d_adjoint,
                              int64_t count, const struct GzMisParams *params, double *d_grad, void *stream);

/* Replay record (no counterpart in the reference): the forward pass keeps, for each of the first
 * `levels` bounce levels and each path, the closest hit of the path's ray, the closest hit of its
 * BSDF-sampled ray and its shadow ray's visibility (33 B per path and level, in HBM), so that the
 * backward pass replays the paths without tracing those rays again. gz_mis_record_bytes: the size
 * of such a record (-1 for bad arguments).
 * gz_gpu_mis_trace_record_device: gz_gpu_mis_trace_device that fills d_record. Only the full
 * wavefront launch (GzMisLaunch.queued 3) records, and only its levels below drainLevel; every
 * other entry is marked as not traced. Images and path records are those of
 * gz_gpu_mis_trace_device.
 * gz_gpu_mis_adjoint_record_device: gz_gpu_mis_adjoint_device over the paths of such a trace (the
 * same starting records, count, parameters and record); it traces the rays the record lacks.
 * Gradients equal gz_gpu_mis_adjoint_device's up to the order of the float64 additions. */
int64_t gz_mis_record_bytes(int64_t count, int32_t levels);
int gz_gpu_mis_trace_record_device(void *handle, struct GzPath *d_paths, int64_t count,
                                   const struct GzMisParams *params, void *d_record, int32_t levels, void *stream);
int gz_gpu_mis_adjoint_record_device(void *handle, struct GzPath *d_paths, const struct GzPathAdjoint *d_adjoint,
                                     int64_t count, const struct GzMisParams *params, double *d_grad,
                                     const void *d_record, int32_t levels, void *stream);

/* ---------------------------------------------------------------------------------------- */
/* Camera samples, primary rays and film (Core/Tile.swift:27-59, :170-219; Core/Film.swift).  */
/* ---------------------------------------------------------------------------------------- */

/* A pinhole PerspectiveCamera with its film and pixel filter, as the host builds it
 * (Camera/PerspectiveCamera.swift:3-23, CameraTransform.swift:3-27): rasterToCamera and
 * cameraToWorld are row-major 4x4 matrices. Paths are numbered pixel-major over the crop
 * [x0, x1) x [y0, y1): path i is sample i % spp of pixel i / spp (row-major in the crop). */
struct GzCamera {
        float rasterToCamera[16];
        float cameraToWorld[16];
        int32_t resX, resY;
        int32_t x0, y0, x1, y1;
        int32_t spp;
        int32_t filterType;      /* GZ_FILTER_GAUSSIAN or GZ_FILTER_BOX */
        float filterRadius[2];   /* the filter's support */
        float filterSigma;
        float lensRadius;        /* must be 0: thin-lens depth of field is not supported */
        float focalDistance;
        int32_t pad;
        int64_t zsSeed;          /* the Z-Sobol constants of the film (gz_zsobol_params) */
        int32_t zsLog2Spp;
        int32_t zsBase4Digits;
};
enum { GZ_FILTER_GAUSSIAN = 0, GZ_FILTER_BOX = 1 };

/* Fresh GzPath records for paths first .. first+count-1 of the crop: startPixelSample (dim 0),
 * getCameraSample (Sampler.swift:69-83), generateRay, filterWeight = filter value / filter pdf
 * (Tile.swift:27-59); dim = 4, lightDim = 0, GZ_PATH_ACTIVE. Asynchronous on `stream`.
 * Replaces the reference's primary-ray loop (Tile.swift:27-59), which has no C-ABI form. */
int gz_gpu_camera_paths_device(const struct GzCamera *camera, int64_t first, int64_t count,
                               struct GzPath *d_paths, void *stream);

/* Film.add(sample:) for `count` finished paths in the layout gz_gpu_camera_paths_device writes
 * (whole pixels: count is a multiple of spp), each pixel's samples added in sample order to
 * the resX*resY images: d_light RGB (estimate x iso/100, clamped to maxComponentValue when it is
 * > 0), d_weight (filterWeight), d_albedo RGB, d_normal RGB (|normalized first normal|)
 * (Core/Film.swift:107-136, Image/Image.swift:21-30). Asynchronous on `stream`. */
int gz_gpu_film_add_device(const struct GzCamera *camera, const struct GzPath *d_paths, int64_t count, float iso,
                           float maxComponentValue, float *d_light, float *d_weight, float *d_albedo,
                           float *d_normal, void *stream);

/* Pixel.normalized for `pixels` pixels (Core/Pixel.swift:12-16): black stays black, otherwise
 * light / weight into d_out (RGB); pixels with light but zero weight (the reference throws) are
 * set to NaN and counted into *d_zeroWeight. Asynchronous on `stream`. */
int gz_gpu_film_resolve_device(const float *d_light, const float *d_weight, int64_t pixels, float *d_out,
                               int32_t *d_zeroWeight, void *stream);

/* ---------------------------------------------------------------------------------------- */
/* Multi-scene batch.                                                                        */
/* ---------------------------------------------------------------------------------------- */

/* Many scenes in one launch sequence (no counterpart in the reference, which renders one scene
 * per process): each scene is a handle with gz_gpu_set_mis_scene done, a camera (whole frame or
 * crop), MIS parameters and a film (device images of resX*resY pixels as gz_gpu_film_add_device
 * takes them, zeroed by the caller). */
struct GzBatchScene {
        void *handle;
        struct GzCamera camera;
        struct GzMisParams params;
        float iso;
        float maxComponentValue;
        float *d_light;
        float *d_weight;
        float *d_albedo;
        float *d_normal;
};

struct GzBatchInfo {
        int64_t totalPaths;      /* path samples of all scenes */
        int32_t rounds;          /* launch sequences one render runs */
        int32_t items;           /* (scene, whole pixels) pieces over all rounds */
        int32_t roundPaths;      /* path slots of a round: the largest round's paths (<= the roundPaths asked) */
        int32_t maxItemsPerRound;
        int32_t levels;          /* bounce levels of a round: 1 + the largest maxDepth */
        int32_t allBsdfs;        /* 1 if some scene of the rounds has a non-diffuse material */
        int32_t soloScenes;      /* scenes that render alone (more than soloPaths paths) */
        int32_t sideScenes;      /* of those, the ones on the batch's side streams (lanes 1-3) */
};

/* Plans the batch: the paths of all scenes (in scene order, whole pixels) are cut into rounds of
 * at most roundPaths path slots, and device tables are built for them; device state (about 280 B
 * per slot) is allocated for the largest round only. A scene with more than soloPaths paths
 * (soloPaths > 0; GZ_BATCH_SOLO_PATHS is the measured break-even on MI355X) fills the chip alone:
 * it stays out of the rounds and each render runs it with the single-scene launch sequence
 * (gz_gpu_camera_paths_device, gz_gpu_mis_trace_device under the scene's own GzMisLaunch,
 * gz_gpu_film_add_device) in pieces of at most roundPaths paths, in four lanes: lane 0 on the
 * render's stream after the rounds, lanes 1-3 on side streams of the batch's own that start with
 * the render (they wait for the work queued on the render's stream before the call, and that
 * stream waits for them at the end), so that one scene's bounce-level tails overlap the other
 * lanes' work. Each scene goes to the lane with the fewest paths so far (lane 0 counting the
 * rounds' paths); a scene whose film shares a buffer with an earlier solo scene's goes to that
 * scene's lane. Each side lane in use has path slots of its own (allocated here, 112 B per slot)
 * and each lane one queue storage for its scenes (allocated by the first render, about 170 B per
 * slot); soloPaths = 0 puts every scene in the rounds.
 * The batch snapshots each scene's MIS tables and parameters: a gz_gpu_set_mis_scene on one of
 * its scenes afterwards makes gz_batch_render_device refuse to run (GZ_E_INVALID; create the
 * batch again). Free the batch before its scenes. All scenes must be on one device. NULL on
 * failure, with gz_last_error() set. */
#define GZ_BATCH_SOLO_PATHS 4194304
void *gz_batch_create(const struct GzBatchScene *scenes, int32_t sceneCount, int64_t roundPaths,
                      int64_t soloPaths);
int gz_batch_info(void *batch, struct GzBatchInfo *info);
/* Launch shape of the batch's bounce levels (queued 1 or 3, as in GzMisLaunch). queued 3 bounces
 * the items of the scenes whose own default launch is the wavefront with the wavefront kernels and
 * the others with the bounce kernel, level by level. The default is {3, 256, 1, 64, 3} when some
 * scene's default is the wavefront, else {1, 256, 1, 256, 3}. */
int gz_batch_set_launch(void *batch, const struct GzMisLaunch *launch);
/* Renders every path of every scene once and adds it to the scenes' films, asynchronously on
 * `stream`. Each round is the same launch sequence (camera rays, device path queues per bounce
 * level, film) driven by device tables and a device round cursor. useGraph = 1 captures that
 * sequence once as a HIP graph (on the first call after creation or gz_batch_set_launch) and
 * replays it once per round; 0 launches it directly. Both give identical films. The scenes that
 * render alone follow the rounds (lane 0) or run beside them on the side streams (lanes 1-3),
 * launched directly. Calls on one batch must not overlap (they share its path slots and queues). */
int gz_batch_render_device(void *batch, int32_t useGraph, void *stream);
/* gz_gpu_check_traversal for a batch: waits for `stream`, returns GZ_E_OVERFLOW if a traversal
 * of any scene's paths stopped on a full stack since the last call (the batch's own flag; then
 * cleared), else GZ_OK. */
int gz_batch_check(void *batch, void *stream);
void gz_batch_free(void *batch);

/* ---------------------------------------------------------------------------------------- */
/* Z-Sobol sampler (replaces the Swift ZSobolSampler, Sampler/ZSobolSampler.swift).            */
/* ---------------------------------------------------------------------------------------- */

/* Sampler constants of ZSobolSampler.init (ZSobolSampler.swift:114-122). */
struct GzZSobolParams {
        int64_t seed;                /* the reference's .pbrt path hard-codes 1234 (Random.swift:66) */
        int32_t log2SamplesPerPixel; /* trailing zero bits of spp (non-powers of 2 truncate) */
        int32_t nBase4Digits;        /* log2(nextPow2(max(w, h))) + (log2spp + 1) / 2 */
};

/* Host: fills *out as ZSobolSampler(samplesPerPixel:fullResolution:seed:) does. */
int gz_zsobol_params(int64_t samplesPerPixel, int64_t resX, int64_t resY, int64_t seed,
                     struct GzZSobolParams *out);

/* For each i < count: startPixelSample(pixel (d_px[i], d_py[i]), index d_sampleIndex[i], dim
 * startDim) (ZSobolSampler.swift:130-136), then the calls of `pattern` (host array of nCalls
 * entries, each 1, 2 or 3 = get1D / get2D / get3D, ZSobolSampler.swift:203-280), values in call
 * order. layout 0: d_out[i * stride + k]; layout 1: d_out[k * count + i]; stride = sum(pattern).
 * nCalls <= 64. Asynchronous on `stream`. */
int gz_zsobol_samples_device(const struct GzZSobolParams *params, const int32_t *d_px, const int32_t *d_py,
                             const int64_t *d_sampleIndex, int64_t count, int32_t startDim,
                             const int8_t *pattern, int32_t nCalls, float *d_out, int32_t layout,
                             void *stream);

/* Same for every sample of a pixel tile in the order of Tile.render (Core/Tile.swift:28-29):
 * sample i is pixel (x0 + (i / spp) % width, y0 + (i / spp) / width), index i % spp. */
int gz_zsobol_tile_device(const struct GzZSobolParams *params, int32_t x0, int32_t y0, int32_t width,
                          int32_t height, int32_t spp, int32_t startDim, const int8_t *pattern,
                          int32_t nCalls, float *d_out, int32_t layout, void *stream);

/* Host: the 1024 x 52 Sobol generator matrices (column-per-bit uint32) the sampler uses, as the
 * reference's SobolGenerator emits them (SobolGenerator/main.swift:11-76). */
int gz_sobol_matrices(uint32_t *out);

/* ---------------------------------------------------------------------------------------- */
/* Host-side helper (no device work): the SAH BVH2 builder the Swift host runs before upload.  */
/* ---------------------------------------------------------------------------------------- */

/* Restates BoundingHierarchyBuilder (Accelerators/BoundingHierarchyBuilder.swift:66-339):
 * binned SAH, 12 buckets on the centroid bounds' longest axis, leaves of up to 4 primitives,
 * Hoare partition by bucket, depth-first flatten (left = i + 1, right = offset).
 * primBounds: n x 6 floats (minX, minY, minZ, maxX, maxY, maxZ) per primitive.
 * nodesOut: capacity 2n - 1 (n >= 1). primOrderOut: n entries; the primitive stored at BVH
 * position k is primOrderOut[k] (BoundingHierarchyBuilder.swift:102-108).
 * Returns the node count (0 for n == 0), or a negative GZ_E_* code. */
int64_t gz_bvh2_build(const float *primBounds, int64_t n, struct BVH2Node *nodesOut,
                      int64_t *primOrderOut);

/* ---------------------------------------------------------------------------------------- */
/* Mesh ingest (Shape/PlyMesh.swift, Shape/Triangle.swift:560-647).                           */
/* ---------------------------------------------------------------------------------------- */

struct GzPlyInfo {
        int64_t vertexCount;
        int64_t faceCount;
        int64_t indexCount;     /* vertex indices of all faces (3 per triangle) */
        int32_t hasNormals;
        int32_t hasUvs;
        int32_t hasFaceIndices;
        int32_t bigEndian;
};

/* Reads a binary PLY file held in memory as PlyMesh does (PlyMesh.swift:24-344): the formats,
 * vertex property sets and face lists it accepts, and nothing else (ASCII PLY is rejected there
 * and here). Call with points == NULL to fill `info`, then with arrays of those sizes:
 * points/normals 3 floats per vertex, uvs 2 per vertex, indices indexCount int64 (faces that
 * are not triangles are kept as they are, as the reference does), faceIndices one per face.
 * Returns GZ_OK or GZ_E_INVALID (malformed, unsupported or truncated) with gz_last_error() set. */
int gz_ply_parse(const uint8_t *data, int64_t size, struct GzPlyInfo *info, float *points, float *normals,
                 float *uvs, int64_t *indices, int64_t *faceIndices);

/* ---------------------------------------------------------------------------------------- */
/* BVH build on the device (the next step of the reference's roadmap,                         */
/* docs/A0_development_roadmap.md:143-146; its builder is the host SAH of                     */
/* Accelerators/BoundingHierarchyBuilder.swift).                                              */
/* ---------------------------------------------------------------------------------------- */

/* maxLeaf: subtrees of at most this many primitives become one leaf (1..64, default 4).
 * sah: 0 = linear BVH (Morton order, Karras hierarchy); 1 = top-down binned SAH with the host
 * builder's decisions (12 buckets on the longest centroid axis, BoundingHierarchyBuilder.swift:
 * 199-339): the nodes hold the same primitive sets as gz_bvh2_build's, except that a node the host
 * keeps as a leaf of several primitives is one leaf only if it has at most maxLeaf.
 * reserved: zero. */
struct GzBuildOptions {
        int32_t maxLeaf;
        int32_t sah;
        int32_t reserved[6];
};

/* gz_gpu_upload_scene without the BVH: uploads the primIds (any order), meshes and materials and
 * builds a linear BVH on the device (63-bit Morton codes of the triangle-box centroids, radix
 * sort, Karras hierarchy, bottom-up boxes, subtrees of <= maxLeaf primitives collapsed into
 * leaves). The handle keeps the primIds in leaf order, so hit primitive indices are leaf-order
 * positions (gz_gpu_export_bvh2 gives each position's index in the caller's array). Every other
 * entry point takes the handle as one from gz_gpu_upload_scene. NULL on failure with
 * gz_last_error() set; options may be NULL. */
void *gz_gpu_build_scene(const struct PrimId_C *primIds, int64_t primIdsCount,
                         const struct TriangleMesh_C *meshes, int64_t meshCount,
                         const int64_t *meshPointsCounts, const int64_t *meshFaceIndicesCounts,
                         const int64_t *meshVertexIndicesCounts, const struct Material_C *materials,
                         int64_t materialCount, const struct GzBuildOptions *options);

/* The handle's tree in the reference's BVH2 layout (BVH2Node.swift: depth-first from node 0,
 * left child at index + 1, right child at `offset`, a leaf over [offset, offset + count) of the
 * primIds in leaf order), for host consumers such as BoundingHierarchy's CPU traversal.
 * order (primIdsCount entries, may be NULL): the caller's index of the primitive at leaf-order
 * position j. Returns the node count, or a negative GZ_E_* code (GZ_E_INVALID when more than
 * nodeCapacity nodes; 2 x primIdsCount always suffices). */
int64_t gz_gpu_export_bvh2(void *handle, struct BVH2Node *nodes, int64_t nodeCapacity, int32_t *order);

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
static_assert(sizeof(struct PrimId_C) == 32, "PrimId_C layout");
static_assert(sizeof(struct Material_C) == 28, "Material_C layout");
static_assert(sizeof(struct TriangleMesh_C) == 24, "TriangleMesh_C layout");
static_assert(sizeof(struct Ray_C) == 24, "Ray_C layout");
static_assert(sizeof(struct Intersection_C) == 48, "Intersection_C layout");
static_assert(offsetof(struct Intersection_C, tHit) == 32, "Intersection_C.tHit");
static_assert(offsetof(struct Intersection_C, hit) == 44, "Intersection_C.hit");
static_assert(sizeof(struct PathState_C) == 88, "PathState_C layout");
static_assert(offsetof(struct PathState_C, pcgState) == 64, "PathState_C.pcgState");
static_assert(offsetof(struct PathState_C, active) == 80, "PathState_C.active");
static_assert(sizeof(struct BVH2Node) == 32, "BVH2Node layout");
static_assert(sizeof(struct SceneDescriptor2_C) == 32, "SceneDescriptor2_C layout");
static_assert(sizeof(struct GzHit) == 16, "GzHit layout");
static_assert(sizeof(struct GzMaterial) == 96, "GzMaterial layout");
static_assert(sizeof(struct GzAreaLight) == 24, "GzAreaLight layout");
static_assert(sizeof(struct GzPath) == 112, "GzPath layout");
static_assert(sizeof(struct GzMisParams) == 32, "GzMisParams layout");
static_assert(sizeof(struct GzPathStats) == 12, "GzPathStats layout");
static_assert(sizeof(struct GzCamera) == 200, "GzCamera layout");
static_assert(sizeof(struct GzPathAdjoint) == 24, "GzPathAdjoint layout");
static_assert(sizeof(struct GzMisLaunch) == 32, "GzMisLaunch layout");
static_assert(sizeof(struct GzBatchScene) == 280, "GzBatchScene layout");
static_assert(sizeof(struct GzBatchInfo) == 40, "GzBatchInfo layout");
static_assert(sizeof(struct GzBuildOptions) == 32, "GzBuildOptions layout");
static_assert(sizeof(struct GzPlyInfo) == 40, "GzPlyInfo layout");
#endif

#endif /* GONZALES_AMD_H */
