// Many scenes rendered in one launch sequence on one device (the MIS integrator of gz_mis.hip
// over several scene handles, each with its own camera, film, parameters and BVH).
//
// The paths of all scenes are cut into rounds of at most `roundPaths` path slots. A round holds
// items: whole pixels of one scene, in scene order, so a round has at most one item more than
// there are scenes. A round is one fixed sequence of launches that reads everything it needs
// from device tables and a device round cursor:
//   camera  one lane per slot: the item's camera sample and primary ray (film_device.h);
//   clear   the per-level, per-item queue counters (a kernel, no memset: graph-safe);
//   per level 0 .. maxDepth:
//     plan    one lane: each item's chunk count and their running sum;
//     bounce  persistent waves take one chunk per atomic; a chunk belongs to one item, so the
//             wave's scene is uniform and its MisArgs come from the scene table by scalar loads;
//             survivors go to the item's queue of the next level by ballot + prefix popcount;
//   film    one lane per pixel: the pixel's samples into its scene's film, in sample order;
//   next    the round cursor moves on.
// Captured once as a HIP graph, the sequence is replayed once per round with no host work
// between rounds; the eager mode launches the same sequence directly.
#include <algorithm>
#include <cstring>
#include <vector>

#include "film_device.h"
#include "mis_device.h"

#pragma clang fp contract(off)

namespace gz {

struct BatchItem {
        int32_t scene;
        int32_t count;      // path slots (whole pixels of the scene's spp)
        int64_t first;      // first path of the item in the scene's crop numbering
        int32_t offset;     // first slot in the round's path buffer
        int32_t pixOffset;  // first pixel of the item among the round's pixels
};

struct BatchRound {
        int32_t firstItem;
        int32_t nItems;
        int32_t total;   // path slots
        int32_t pixels;
};

struct BatchFilm {
        float *light, *weight, *albedo, *normal;
        float iso, maxComponent;
        int32_t resX, resY, spp, pad;
};

// The round's item containing position `pos` of a running sum `key` (largest i with key(i) <= pos).
template <typename Key>
__device__ __forceinline__ int find_item(int n, int pos, Key &&key) {
        int lo = 0, hi = n;  // answer in [lo, hi)
        while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (key(mid) <= pos)
                        lo = mid;
                else
                        hi = mid;
        }
        return lo;
}

__global__ void k_batch_begin(int *cursor) {
        if (threadIdx.x == 0 && blockIdx.x == 0) cursor[0] = 0;
}

__global__ void k_batch_next(int *cursor) {
        if (threadIdx.x == 0 && blockIdx.x == 0) cursor[0] = cursor[0] + 1;
}

constexpr int kBatchBlock = 256;

__global__ __launch_bounds__(kBatchBlock) void k_batch_camera(const int *__restrict__ cursor,
                                                              const BatchRound *__restrict__ rounds,
                                                              const BatchItem *__restrict__ items,
                                                              const GzCamera *__restrict__ cams,
                                                              const uint32_t *__restrict__ table,
                                                              GzPath *__restrict__ paths) {
        const BatchRound R = rounds[cursor[0]];
        const int t = (int)(blockIdx.x * kBatchBlock + threadIdx.x);
        if (t >= R.total) return;
        const BatchItem *its = items + R.firstItem;
        const int k = find_item(R.nItems, t, [&](int i) { return its[i].offset; });
        const BatchItem it = its[k];
        paths[t] = camera_path(cams[it.scene], table, it.first + (t - it.offset));
}

// Per level: n[level][item] queue entries, head[level] chunks handed out, base[level][item]
// the chunk running sum (base[level][nItems] = the level's chunks).
struct BatchCounters {
        unsigned *n;
        unsigned *head;
        unsigned *base;
        int maxItems;
        int levels;
};

__global__ void k_batch_clear(BatchCounters c) {
        const int nn = (c.levels + 1) * c.maxItems;
        for (int k = threadIdx.x; k < nn; k += blockDim.x) c.n[k] = 0u;
        for (int k = threadIdx.x; k < c.levels; k += blockDim.x) c.head[k] = 0u;
}

__global__ void k_batch_plan(const int *__restrict__ cursor, const BatchRound *__restrict__ rounds,
                             const BatchItem *__restrict__ items, BatchCounters c, int level, unsigned chunk) {
        if (threadIdx.x != 0 || blockIdx.x != 0) return;
        const BatchRound R = rounds[cursor[0]];
        unsigned *base = c.base + (size_t)level * (c.maxItems + 1);
        unsigned sum = 0;
        for (int i = 0; i < R.nItems; ++i) {
                base[i] = sum;
                const unsigned n = level == 0 ? (unsigned)items[R.firstItem + i].count : c.n[level * c.maxItems + i];
                sum += (n + chunk - 1) / chunk;
        }
        base[R.nItems] = sum;
}

// One bounce level of a round: waves take chunks of the level's work items through one global
// counter; survivors go to the item's next-level queue (appends held over four runs of 64 and
// written with one atomic). kDrain: a lane runs each path's remaining bounces instead, and no
// queue is written (the last levels of a round, in one launch).
template <bool kAll, bool kDrain>
__global__ __launch_bounds__(kMisBlock, kAll ? 2 : 4) void k_batch_bounce(const int *__restrict__ cursor,
                                                            const BatchRound *__restrict__ rounds,
                                                            const BatchItem *__restrict__ items,
                                                            const MisArgs *__restrict__ args, GzPath *__restrict__ paths,
                                                            const int *__restrict__ qin, int *__restrict__ qout,
                                                            BatchCounters c, int level, unsigned chunk) {
        const BatchRound R = rounds[cursor[0]];
        const BatchItem *its = items + R.firstItem;
        const unsigned *base = c.base + (size_t)level * (c.maxItems + 1);
        const unsigned total = base[R.nItems];
        unsigned *nOut = c.n + (size_t)(level + 1) * c.maxItems;
        const unsigned *nIn = c.n + (size_t)level * c.maxItems;
        const int lane = (int)(threadIdx.x & 63);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        for (;;) {
                unsigned g = 0;
                if (lane == 0) g = atomicAdd(&c.head[level], 1u);
                g = (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl(g, 0));
                if (g >= total) break;
                const int k = __builtin_amdgcn_readfirstlane(
                        find_item(R.nItems, (int)g, [&](int i) { return (int)base[i]; }));
                const BatchItem it = its[k];
                const MisArgs &a = args[__builtin_amdgcn_readfirstlane(it.scene)];
                const unsigned n = level == 0 ? (unsigned)it.count : nIn[k];
                const unsigned e0 = (g - base[k]) * chunk;
                const unsigned e1 = n - e0 < chunk ? n : e0 + chunk;
                if (kDrain) {
                        for (unsigned j = e0 + (unsigned)lane; j < e1; j += 64) {
                                const int idx = level == 0 ? it.offset + (int)j : qin[it.offset + (int)j];
                                for (int b = 0; b <= a.maxDepth; ++b)
                                        if (!bounce_one<false, kAll, false>(a, paths, idx, nullptr, AdjArgs{nullptr, nullptr},
                                                                           nullptr))
                                                break;
                        }
                        continue;
                }
                int *const qo[1] = {qout + it.offset};
                unsigned *const on[1] = {&nOut[k]};
                Keep4 keep;
                unsigned bits[1] = {0u};
                int runs = 0;
                for (unsigned j0 = e0; j0 < e1; j0 += 64) {
                        const unsigned j = j0 + (unsigned)lane;
                        int idx = -1;
                        if (j < e1) idx = level == 0 ? it.offset + (int)j : qin[it.offset + (int)j];
                        const bool go = idx >= 0 && bounce_one<false, kAll, false>(a, paths, idx, nullptr,
                                                                                   AdjArgs{nullptr, nullptr}, nullptr);
                        keep.set(runs, idx);
                        bits[0] |= (go ? 1u : 0u) << runs;
                        if (++runs == 4) {
                                flush_appends<1>(keep, bits, runs, qo, on, lane, below);
                                runs = 0;
                                bits[0] = 0u;
                        }
                }
                if (runs > 0) flush_appends<1>(keep, bits, runs, qo, on, lane, below);
        }
}

__global__ __launch_bounds__(kBatchBlock) void k_batch_film(const int *__restrict__ cursor,
                                                            const BatchRound *__restrict__ rounds,
                                                            const BatchItem *__restrict__ items,
                                                            const BatchFilm *__restrict__ films,
                                                            const GzPath *__restrict__ paths) {
        const BatchRound R = rounds[cursor[0]];
        const int p = (int)(blockIdx.x * kBatchBlock + threadIdx.x);
        if (p >= R.pixels) return;
        const BatchItem *its = items + R.firstItem;
        const int k = find_item(R.nItems, p, [&](int i) { return its[i].pixOffset; });
        const BatchItem it = its[k];
        const BatchFilm f = films[it.scene];
        film_add_pixel(f.resX, f.resY, f.spp, paths + it.offset + (int64_t)(p - it.pixOffset) * f.spp, f.iso,
                       f.maxComponent, f.light, f.weight, f.albedo, f.normal);
}

struct Batch {
        int device = 0;
        int nScenes = 0;
        std::vector<Scene *> scenes;
        bool all = false;  // some scene has a non-diffuse material
        int levels = 0;    // maxDepth + 1 over the scenes
        int64_t totalPaths = 0;
        int roundPaths = 0;
        int maxRoundPixels = 0;
        int maxItems = 0;
        std::vector<BatchRound> hRounds;
        std::vector<BatchItem> hItems;
        GzMisLaunch launch{1, 256, 1, 256, 3, {0, 0, 0}};
        int cuCount = 256;

        int *dCursor = nullptr;
        BatchRound *dRounds = nullptr;
        BatchItem *dItems = nullptr;
        MisArgs *dArgs = nullptr;
        GzCamera *dCams = nullptr;
        BatchFilm *dFilms = nullptr;
        GzPath *dPaths = nullptr;
        int *dQ[2] = {nullptr, nullptr};
        BatchCounters ctr{nullptr, nullptr, nullptr, 0, 0};
        const uint32_t *sobol = nullptr;

        hipStream_t captureStream = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
};

static void release_graph(Batch *b) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
}

static void free_batch(Batch *b) {
        if (!b) return;
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != b->device) (void)hipSetDevice(b->device);
        release_graph(b);
        if (b->captureStream) (void)hipStreamDestroy(b->captureStream);
        void *ptrs[] = {b->dCursor, b->dRounds, b->dItems, b->dArgs, b->dCams, b->dFilms, b->dPaths,
                        b->dQ[0],   b->dQ[1],  b->ctr.n,  b->ctr.head, b->ctr.base};
        for (void *p : ptrs)
                if (p) (void)hipFree(p);
        delete b;
}

template <class T>
static bool upload(T **dst, const T *src, size_t count, const char *what) {
        if (!hip_ok(hipMalloc(dst, std::max<size_t>(count, 1) * sizeof(T)), what)) return false;
        return count == 0 || hip_ok(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice), what);
}

// Enqueues one round's launch sequence on `st`.
static bool enqueue_round(Batch *b, hipStream_t st) {
        const unsigned chunk = (unsigned)b->launch.chunk;
        const int block = b->launch.block;
        auto kern = b->all ? k_batch_bounce<true, false> : k_batch_bounce<false, false>;
        auto kdrain = b->all ? k_batch_bounce<true, true> : k_batch_bounce<false, true>;
        const int drainLevel = b->launch.drainLevel;
        int perCU = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, block, 0) != hipSuccess || perCU < 1) perCU = 1;
        const int64_t want = (int64_t)b->cuCount * perCU * b->launch.gridScale;
        const int64_t need = ((int64_t)b->roundPaths + chunk - 1) / chunk * 64 / block + 1;
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min(want, need));
        const unsigned camBlocks = (unsigned)((b->roundPaths + kBatchBlock - 1) / kBatchBlock);
        const unsigned filmBlocks = (unsigned)((b->maxRoundPixels + kBatchBlock - 1) / kBatchBlock);
        hipLaunchKernelGGL(k_batch_camera, dim3(camBlocks), dim3(kBatchBlock), 0, st, b->dCursor, b->dRounds, b->dItems,
                           b->dCams, b->sobol, b->dPaths);
        hipLaunchKernelGGL(k_batch_clear, dim3(1), dim3(256), 0, st, b->ctr);
        for (int level = 0; level < b->levels; ++level) {
                hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, st, b->dCursor, b->dRounds, b->dItems, b->ctr,
                                   level, chunk);
                const bool drain = drainLevel > 0 && level >= drainLevel;
                hipLaunchKernelGGL(drain ? kdrain : kern, dim3(grid), dim3(block), 0, st, b->dCursor, b->dRounds,
                                   b->dItems, b->dArgs, b->dPaths, b->dQ[level & 1], b->dQ[(level + 1) & 1], b->ctr,
                                   level, chunk);
                if (drain) break;
        }
        hipLaunchKernelGGL(k_batch_film, dim3(filmBlocks), dim3(kBatchBlock), 0, st, b->dCursor, b->dRounds, b->dItems,
                           b->dFilms, b->dPaths);
        hipLaunchKernelGGL(k_batch_next, dim3(1), dim3(64), 0, st, b->dCursor);
        return hip_ok(hipGetLastError(), "batch round launch");
}

static bool capture(Batch *b) {
        release_graph(b);
        if (!b->captureStream &&
            !hip_ok(hipStreamCreateWithFlags(&b->captureStream, hipStreamNonBlocking), "batch capture stream"))
                return false;
        if (!hip_ok(hipStreamBeginCapture(b->captureStream, hipStreamCaptureModeThreadLocal), "begin capture"))
                return false;
        const bool ok = enqueue_round(b, b->captureStream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->captureStream, &g);
        if (!ok || !hip_ok(e, "end capture")) {
                if (g) (void)hipGraphDestroy(g);
                return false;
        }
        b->graph = g;
        if (!hip_ok(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate")) {
                release_graph(b);
                return false;
        }
        return true;
}

}  // namespace gz

using namespace gz;

extern "C" void *gz_batch_create(const GzBatchScene *scenes, int32_t sceneCount, int64_t roundPaths) {
        clear_error();
        if (!scenes || sceneCount <= 0 || roundPaths <= 0 || roundPaths > 0x7fffffffLL) {
                set_error("gz_batch_create: need scenes and 0 < roundPaths < 2^31");
                return nullptr;
        }
        Batch *b = new Batch;
        b->nScenes = sceneCount;
        b->roundPaths = (int)roundPaths;
        std::vector<MisArgs> args((size_t)sceneCount);
        std::vector<GzCamera> cams((size_t)sceneCount);
        std::vector<BatchFilm> films((size_t)sceneCount);
        int maxDepth = 0;
        for (int32_t s = 0; s < sceneCount; ++s) {
                const GzBatchScene &bs = scenes[s];
                Scene *sc = static_cast<Scene *>(bs.handle);
                if (!sc) {
                        set_error("gz_batch_create: scene %d has no handle", s);
                        free_batch(b);
                        return nullptr;
                }
                if (!camera_ok(&bs.camera, "gz_batch_create")) {
                        free_batch(b);
                        return nullptr;
                }
                if (!bs.d_light || !bs.d_weight || !bs.d_albedo || !bs.d_normal) {
                        set_error("gz_batch_create: scene %d lacks a film image", s);
                        free_batch(b);
                        return nullptr;
                }
                if (bs.camera.spp > b->roundPaths) {
                        set_error("gz_batch_create: roundPaths %d below scene %d's spp %d", b->roundPaths, s,
                                  bs.camera.spp);
                        free_batch(b);
                        return nullptr;
                }
                std::lock_guard<std::mutex> lock(sc->mu);
                if (s == 0) {
                        b->device = sc->device;
                        b->cuCount = sc->cuCount;
                } else if (sc->device != b->device) {
                        set_error("gz_batch_create: scene %d is on device %d, scene 0 on %d", s, sc->device, b->device);
                        free_batch(b);
                        return nullptr;
                }
                if (!enter_device(sc) || !mis_args(sc, &bs.params, args[(size_t)s])) {
                        free_batch(b);
                        return nullptr;
                }
                if (!mis_all_diffuse(sc)) b->all = true;
                maxDepth = std::max(maxDepth, (int)bs.params.maxDepth);
                cams[(size_t)s] = bs.camera;
                films[(size_t)s] = BatchFilm{bs.d_light,       bs.d_weight,         bs.d_albedo,  bs.d_normal,
                                             bs.iso,           bs.maxComponentValue, bs.camera.resX, bs.camera.resY,
                                             bs.camera.spp,    0};
        }
        b->levels = maxDepth + 1;
        // Rounds of whole pixels, scenes in order.
        int used = 0, pixels = 0, firstItem = 0;
        auto close_round = [&]() {
                if (used == 0) return;
                b->hRounds.push_back(BatchRound{firstItem, (int)b->hItems.size() - firstItem, used, pixels});
                b->maxRoundPixels = std::max(b->maxRoundPixels, pixels);
                b->maxItems = std::max(b->maxItems, (int)b->hItems.size() - firstItem);
                firstItem = (int)b->hItems.size();
                used = 0;
                pixels = 0;
        };
        for (int32_t s = 0; s < sceneCount; ++s) {
                const GzCamera &c = cams[(size_t)s];
                const int64_t cropPixels = crop_paths(&c) / c.spp;
                int64_t done = 0;
                b->totalPaths += crop_paths(&c);
                while (done < cropPixels) {
                        const int64_t room = (b->roundPaths - used) / c.spp;
                        if (room == 0) {
                                close_round();
                                continue;
                        }
                        const int64_t take = std::min(room, cropPixels - done);
                        b->hItems.push_back(BatchItem{s, (int)(take * c.spp), done * c.spp, used, pixels});
                        used += (int)(take * c.spp);
                        pixels += (int)take;
                        done += take;
                }
        }
        close_round();
        b->ctr.maxItems = std::max(b->maxItems, 1);
        b->ctr.levels = b->levels;
        const size_t slots = (size_t)b->roundPaths;
        int cur = -1;
        if (!hip_ok(hipGetDevice(&cur), "get device") || (cur != b->device && !hip_ok(hipSetDevice(b->device), "set device")) ||
            !(b->sobol = device_sobol_table()) || !upload(&b->dRounds, b->hRounds.data(), b->hRounds.size(), "batch rounds") ||
            !upload(&b->dItems, b->hItems.data(), b->hItems.size(), "batch items") ||
            !upload(&b->dArgs, args.data(), args.size(), "batch scene arguments") ||
            !upload(&b->dCams, cams.data(), cams.size(), "batch cameras") ||
            !upload(&b->dFilms, films.data(), films.size(), "batch films") ||
            !hip_ok(hipMalloc(&b->dCursor, sizeof(int)), "batch cursor") ||
            !hip_ok(hipMalloc(&b->dPaths, slots * sizeof(GzPath)), "batch paths") ||
            !hip_ok(hipMalloc(&b->dQ[0], slots * sizeof(int)), "batch queue") ||
            !hip_ok(hipMalloc(&b->dQ[1], slots * sizeof(int)), "batch queue") ||
            !hip_ok(hipMalloc(&b->ctr.n, (size_t)(b->levels + 1) * b->ctr.maxItems * sizeof(unsigned)), "batch counters") ||
            !hip_ok(hipMalloc(&b->ctr.head, (size_t)b->levels * sizeof(unsigned)), "batch counters") ||
            !hip_ok(hipMalloc(&b->ctr.base, (size_t)b->levels * (b->ctr.maxItems + 1) * sizeof(unsigned)),
                    "batch counters")) {
                free_batch(b);
                return nullptr;
        }
        return b;
}

extern "C" int gz_batch_info(void *batch, GzBatchInfo *info) {
        clear_error();
        if (!batch || !info) {
                set_error("gz_batch_info: bad arguments");
                return GZ_E_INVALID;
        }
        const Batch *b = static_cast<Batch *>(batch);
        info->totalPaths = b->totalPaths;
        info->rounds = (int32_t)b->hRounds.size();
        info->items = (int32_t)b->hItems.size();
        info->roundPaths = b->roundPaths;
        info->maxItemsPerRound = b->maxItems;
        info->levels = b->levels;
        info->allBsdfs = b->all ? 1 : 0;
        return GZ_OK;
}

extern "C" int gz_batch_set_launch(void *batch, const GzMisLaunch *launch) {
        clear_error();
        if (!batch || !launch) {
                set_error("gz_batch_set_launch: bad arguments");
                return GZ_E_INVALID;
        }
        const GzMisLaunch &c = *launch;
        if (c.queued != 1 || (c.block != 64 && c.block != 128 && c.block != 256) || c.gridScale < 1 ||
            c.gridScale > 16 || c.chunk < 64 || c.chunk > 4096 || c.chunk % 64 != 0 || c.drainLevel < 0 ||
            c.drainLevel > 64 || c.reserved[0] != 0 || c.reserved[1] != 0 || c.reserved[2] != 0) {
                set_error("gz_batch_set_launch: queued %d block %d gridScale %d chunk %d drainLevel %d out of range "
                          "(queued must be 1)",
                          c.queued, c.block, c.gridScale, c.chunk, c.drainLevel);
                return GZ_E_INVALID;
        }
        Batch *b = static_cast<Batch *>(batch);
        b->launch = c;
        release_graph(b);  // captured with the old shape
        return GZ_OK;
}

extern "C" int gz_batch_render_device(void *batch, int32_t useGraph, void *stream) {
        clear_error();
        if (!batch) {
                set_error("gz_batch_render_device: null batch");
                return GZ_E_INVALID;
        }
        Batch *b = static_cast<Batch *>(batch);
        int cur = -1;
        if (!hip_ok(hipGetDevice(&cur), "get device") || (cur != b->device && !hip_ok(hipSetDevice(b->device), "set device")))
                return GZ_E_HIP;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (useGraph && !b->exec && !capture(b)) return GZ_E_HIP;
        hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, b->dCursor);
        if (!hip_ok(hipGetLastError(), "batch begin")) return GZ_E_HIP;
        for (size_t r = 0; r < b->hRounds.size(); ++r) {
                if (useGraph) {
                        if (!hip_ok(hipGraphLaunch(b->exec, st), "graph launch")) return GZ_E_HIP;
                } else if (!enqueue_round(b, st)) {
                        return GZ_E_HIP;
                }
        }
        return GZ_OK;
}

extern "C" void gz_batch_free(void *batch) { free_batch(static_cast<Batch *>(batch)); }
