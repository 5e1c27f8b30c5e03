// BVH2 closest-hit and any-hit traversal for gfx950, and the entry points that drive it.
//
// Semantics follow the reference's traverse_bvh2_core (Sources/mojoKernel/kernel.mojo:223-364)
// step for step, so hit indices, t, u and v are bit-identical to the reference's CPU path:
//   - rdir = 1/dir, orgRdir = org*rdir, nearIsMin = rdir >= 0 (kernel.mojo:233-243);
//   - the root box is never tested; traversal starts at node 0;
//   - at an interior node both children's boxes are tested against the current closest t
//     (kernel.mojo:314-327, slab test kernel.mojo:171-206 with its select-form min/max and the
//     x1.0000003 slack); both hit -> go to the nearer (left wins ties) and push the other;
//   - leaves test their primitives in order with Moller-Trumbore (kernel.mojo:122-156); a hit
//     with t == current t is accepted, so the later primitive wins ties.
// Products, sums and differences are rounded one by one (`fp contract(off)`), divisions are
// IEEE-correct (hipcc's default), and min/max are written as the reference's selects so NaN
// propagates the same way.
//
// MI355X layout: one lane per ray; the interior visit reads one 64-byte child-pair record
// instead of two 32-byte nodes, a leaf reads 48-byte pre-gathered triangle records instead of
// PrimId -> mesh -> index -> point chains (five dependent loads per vertex in the reference).
#include <algorithm>

#include "traverse_device.h"

#pragma clang fp contract(off)

namespace gz {

// One ray of a binned batch in its traced order: the Ray_C, its tMax and its index in the caller's
// arrays (32 bytes, two 16-byte loads).
struct alignas(16) SortedRay {
        float4 a;  // orgX, orgY, orgZ, dirX
        float4 b;  // dirY, dirZ, tMax, index (uint32 bits)
};

// Starts lane L on entry j of a binned batch; returns the ray's index in the caller's arrays.
__device__ __forceinline__ int64_t lane_start_sorted(Lane &L, const SortedRay *__restrict__ sorted, int64_t j,
                                                     int rootDesc, bool sceneSafe) {
        const SortedRay r = sorted[j];
        const Ray_C ray{r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y};
        lane_init(L, ray, r.b.z, rootDesc, sceneSafe);
        return (int64_t)__float_as_uint(r.b.w);
}

// One lane per ray, one ray per lane.
template <bool kAny, bool kStats, int kLds>
__global__ __launch_bounds__(kBlock) void k_trace(const NodePair *__restrict__ nodes, const TriRec *__restrict__ tris,
                                                  int rootDesc, bool sceneSafe, const Ray_C *__restrict__ rays,
                                                  const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
                                                  uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats,
                                                  int *__restrict__ err, const SortedRay *__restrict__ order) {
        const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (j >= n) return;
        Lane L;
        int64_t i = j;
        if (order)
                i = lane_start_sorted(L, order, j, rootDesc, sceneSafe);
        else
                lane_start(L, rays, tMaxIn, i, rootDesc, sceneSafe);
        Stack<kLds> S;
        int st;
        do {
                st = step<kAny, kStats, kLds>(L, S, nodes, tris);
        } while (st == kStepGo);
        lane_finish<kAny, kStats>(L, i, st == kStepOverflow, hits, occ, stats, err);
}

// Persistent lanes with ray refill. A wave keeps its lanes busy: once `refill` lanes are idle,
// each idle lane takes the next ray index from the wave's reserve, a range of `chunk`
// consecutive rays grabbed from the global counter with one atomic (so the counter sees one
// atomic per `chunk` rays, not one per refill). The long tail of deep rays no longer idles the
// other lanes of a wave. A wave leaves once the counter is exhausted and its lanes are idle.
template <bool kAny, bool kStats, int kLds>
__global__ __launch_bounds__(kBlock) void k_trace_persistent(
        const NodePair *__restrict__ nodes, const TriRec *__restrict__ tris, int rootDesc, bool sceneSafe,
        const Ray_C *__restrict__ rays, const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
        uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats, int *__restrict__ err,
        unsigned long long *__restrict__ counter, int refill, int chunk, const SortedRay *__restrict__ order) {
        Stack<kLds> S;
        Lane L;
        int64_t ray = -1;
        // Wave-uniform state: the reserve [resBase, resBase + resCount), whether the global counter
        // has run past n (no further atomics), and whether no ray is left to hand out at all.
        unsigned long long resBase = 0, resCount = 0;
        bool counterDone = false, exhausted = false;
        const int laneId = (int)(threadIdx.x & 63);
        const uint64_t below = laneId == 0 ? 0ull : (~0ull >> (64 - laneId));
        const unsigned long long un = (unsigned long long)n;
        for (;;) {
                if (!exhausted) {
                        const uint64_t idle = __ballot(ray < 0);
                        const int nIdle = __popcll(idle);
                        if (nIdle >= refill || nIdle == 64) {
                                const unsigned long long rank = (unsigned long long)__popcll(idle & below);
                                const unsigned long long need = (unsigned long long)nIdle;
                                unsigned long long fresh = 0, take = 0;
                                if (resCount < need && !counterDone) {
                                        // Top up with a new chunk, at least enough for every idle lane.
                                        take = need - resCount > (unsigned long long)chunk ? need - resCount
                                                                                        : (unsigned long long)chunk;
                                        const int leader = __ffsll((unsigned long long)idle) - 1;
                                        if (laneId == leader) fresh = atomicAdd(counter, take);
                                        fresh = __shfl(fresh, leader);
                                        if (fresh + take >= un) counterDone = true;
                                }
                                if (ray < 0) {
                                        // The first resCount idle lanes drain the reserve, the others
                                        // start on the fresh chunk.
                                        unsigned long long mine = ~0ull;
                                        if (rank < resCount)
                                                mine = resBase + rank;
                                        else if (rank - resCount < take)
                                                mine = fresh + (rank - resCount);
                                        if (mine < un) {
                                                if (order) {
                                                        ray = lane_start_sorted(L, order, (int64_t)mine, rootDesc,
                                                                                sceneSafe);
                                                } else {
                                                        ray = (int64_t)mine;
                                                        lane_start(L, rays, tMaxIn, ray, rootDesc, sceneSafe);
                                                }
                                        }
                                }
                                if (take > 0) {
                                        const unsigned long long used = need - resCount;  // <= take
                                        resBase = fresh + used;
                                        resCount = take - used;
                                } else {
                                        const unsigned long long used = need < resCount ? need : resCount;
                                        resBase += used;
                                        resCount -= used;
                                }
                                if (counterDone && (resCount == 0 || resBase >= un)) exhausted = true;
                        }
                }
                if (ray >= 0) {
                        const int st = step<kAny, kStats, kLds>(L, S, nodes, tris);
                        if (st != kStepGo) {
                                lane_finish<kAny, kStats>(L, ray, st == kStepOverflow, hits, occ, stats, err);
                                ray = -1;
                        }
                }
                if (exhausted && __ballot(ray >= 0) == 0ull) break;
        }
}

// Persistent lanes with ray refill, one ray range per XCD. The batch (in its traced order) is cut
// into eight consecutive ranges with a counter each; a wave takes chunks from the range of its XCD
// (workgroups go to the eight XCDs in turn, so blockIdx.x mod 8 names it) and, once that range is
// handed out, from the next ones in turn. Each XCD's L2 then serves waves that trace neighbouring
// rays of the order (with the binning: rays that start in neighbouring cells). Refill as in
// k_trace_persistent; a wave leaves once all eight ranges are handed out and its lanes are idle.
template <bool kAny, bool kStats, int kLds>
__global__ __launch_bounds__(kBlock) void k_trace_xcd(
        const NodePair *__restrict__ nodes, const TriRec *__restrict__ tris, int rootDesc, bool sceneSafe,
        const Ray_C *__restrict__ rays, const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
        uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats, int *__restrict__ err,
        unsigned long long *__restrict__ counters, int refill, int chunk, const SortedRay *__restrict__ order) {
        Stack<kLds> S;
        Lane L;
        int64_t ray = -1;
        const unsigned long long un = (unsigned long long)n;
        // wave-uniform: the range drawn from, ranges used up so far, the reserve [resBase, resEnd)
        int cur = (int)(blockIdx.x & 7u), tried = 0;
        unsigned long long resBase = 0, resEnd = 0;
        bool exhausted = false;
        const int laneId = (int)(threadIdx.x & 63);
        const uint64_t below = laneId == 0 ? 0ull : (~0ull >> (64 - laneId));
        for (;;) {
                if (!exhausted) {
                        const uint64_t idle = __ballot(ray < 0);
                        const int nIdle = __popcll(idle);
                        if (nIdle >= refill || nIdle == 64) {
                                const unsigned long long rank = (unsigned long long)__popcll(idle & below);
                                const unsigned long long need = (unsigned long long)nIdle;
                                const unsigned long long resCount = resEnd - resBase;
                                unsigned long long fresh = 0, freshEnd = 0;
                                if (resCount < need && tried < 8) {
                                        const unsigned long long take =
                                                need - resCount > (unsigned long long)chunk ? need - resCount
                                                                                           : (unsigned long long)chunk;
                                        const unsigned long long rs = un * (unsigned long long)cur / 8ull;
                                        const unsigned long long re = un * (unsigned long long)(cur + 1) / 8ull;
                                        const int leader = __ffsll((unsigned long long)idle) - 1;
                                        unsigned long long old = 0;
                                        if (laneId == leader) old = atomicAdd(&counters[cur], take);
                                        old = __shfl(old, leader);
                                        if (old < re - rs) {
                                                fresh = rs + old;
                                                freshEnd = fresh + take < re ? fresh + take : re;
                                        }
                                        if (old + take >= re - rs) {  // this range is handed out
                                                cur = (cur + 1) & 7;
                                                ++tried;
                                        }
                                }
                                if (ray < 0) {
                                        unsigned long long mine = ~0ull;
                                        if (rank < resCount)
                                                mine = resBase + rank;
                                        else if (rank - resCount < freshEnd - fresh)
                                                mine = fresh + (rank - resCount);
                                        if (mine < un) {
                                                if (order) {
                                                        ray = lane_start_sorted(L, order, (int64_t)mine, rootDesc,
                                                                                sceneSafe);
                                                } else {
                                                        ray = (int64_t)mine;
                                                        lane_start(L, rays, tMaxIn, ray, rootDesc, sceneSafe);
                                                }
                                        }
                                }
                                if (resCount >= need) {
                                        resBase += need;
                                } else {
                                        const unsigned long long used = need - resCount;
                                        resBase = fresh + used < freshEnd ? fresh + used : freshEnd;
                                        resEnd = freshEnd;
                                }
                                if (tried >= 8 && resBase == resEnd) exhausted = true;
                        }
                }
                if (ray >= 0) {
                        const int st = step<kAny, kStats, kLds>(L, S, nodes, tris);
                        if (st != kStepGo) {
                                lane_finish<kAny, kStats>(L, ray, st == kStepOverflow, hits, occ, stats, err);
                                ray = -1;
                        }
                }
                if (exhausted && __ballot(ray >= 0) == 0ull) break;
        }
}

// Strided persistent lanes: global lane g traces rays g, g + G, g + 2G, ... (G = lanes in the
// grid) back to back, starting the next ray the step after the previous one ends. Summing many
// rays per lane evens out the per-ray work, with no atomics; consecutive lanes still read
// consecutive rays. Each lane leaves after its last ray, so every wave exits.
template <bool kAny, bool kStats, int kLds>
__global__ __launch_bounds__(kBlock) void k_trace_strided(
        const NodePair *__restrict__ nodes, const TriRec *__restrict__ tris, int rootDesc, bool sceneSafe,
        const Ray_C *__restrict__ rays, const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
        uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats, int *__restrict__ err,
        const SortedRay *__restrict__ order) {
        Stack<kLds> S;
        const int64_t G = (int64_t)gridDim.x * kBlock;
        int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (j >= n) return;
        Lane L;
        int64_t ray = j;
        if (order)
                ray = lane_start_sorted(L, order, j, rootDesc, sceneSafe);
        else
                lane_start(L, rays, tMaxIn, ray, rootDesc, sceneSafe);
        for (;;) {
                const int st = step<kAny, kStats, kLds>(L, S, nodes, tris);
                if (st != kStepGo) {
                        lane_finish<kAny, kStats>(L, ray, st == kStepOverflow, hits, occ, stats, err);
                        j += G;
                        if (j >= n) break;
                        if (order) {
                                ray = lane_start_sorted(L, order, j, rootDesc, sceneSafe);
                        } else {
                                ray = j;
                                lane_start(L, rays, tMaxIn, ray, rootDesc, sceneSafe);
                        }
                }
        }
}

// ---- Ray binning: a counting sort of the batch by origin cell ------------------------------------
//
// The key of a ray is the Morton code of its origin's cell in a 2^kBinAxisBits-per-axis grid over
// the scene box (the union of the root's two child boxes; origins outside it are clamped to the
// nearest cell). Rays of one cell run next to each other, in no particular order within the cell,
// so the waves in flight at any time start in a compact part of the scene and walk overlapping
// parts of the tree: more of it is found in L2. The sort never uses global atomics:
//   k_bin_count  : per block of `per` rays, an LDS histogram of the keys, written as
//                  counts[key][block]; the keys kept as 16-bit values;
//   k_scan_*     : exclusive prefix sum of counts in [key][block] order (three kernels);
//   k_bin_place  : per block, the LDS copy of its offsets hands out ranks with LDS atomics;
//                  the ray, its tMax and its index go to entry rank of the sorted batch, which the
//                  traversal then reads in order.
constexpr int kBinAxisBits = 4;
constexpr uint32_t kBins = 1u << (3 * kBinAxisBits);
constexpr int kBinBlock = 1024;  // threads per block of the count and place kernels

// Rays per block of the count and place kernels: enough blocks for one per CU (256), each as large
// as that allows up to 65536 rays. The blocks the hardware sends to one XCD take consecutive logical
// blocks (bin_block), so each bin's runs from neighbouring blocks meet in one L2 and reach HBM as
// whole lines: on 16M rays, 8192-ray blocks in hardware order took 0.80 ms, in XCD order 0.76 ms,
// 65536-ray blocks of 1024 threads 0.60 ms.
static uint32_t bin_per(int64_t n) {
        uint32_t per = 4096;
        while (per < 65536u && (int64_t)per * 256 < n) per <<= 1;
        return per;
}

// The logical block of a count or place block (a bijection on [0, n)): hardware block b runs on
// XCD b mod 8, and XCD x takes the x-th run of consecutive logical blocks.
__device__ __forceinline__ uint32_t bin_block(uint32_t b, uint32_t n) {
        const uint32_t x = b & 7u, q = n >> 3, r = n & 7u;
        return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ uint32_t spread3_4(uint32_t v) {  // bits b3..b0 -> b9, b6, b3, b0
        return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6);
}

__device__ __forceinline__ uint32_t cell_of(float x, float lo, float scale, uint32_t cells) {
        const float q = (x - lo) * scale;
        // NaN and values below 0 go to cell 0, values past the top to the last cell
        if (!(q > 0.0f)) return 0u;
        if (q >= (float)cells) return cells - 1u;
        return (uint32_t)q;
}

struct BinBox {
        float lox, loy, loz, sx, sy, sz;
};

__device__ __forceinline__ BinBox bin_box(const NodePair *__restrict__ nodes, int rootDesc) {
        const NodePair root = nodes[rootDesc];
        const float lox = fminf(root.a.x, root.b.z), loy = fminf(root.a.y, root.b.w), loz = fminf(root.a.z, root.c.x);
        const float hix = fmaxf(root.a.w, root.c.y), hiy = fmaxf(root.b.x, root.c.z), hiz = fmaxf(root.b.y, root.c.w);
        constexpr float cells = (float)(1 << kBinAxisBits);
        return BinBox{lox, loy, loz, cells / (hix - lox), cells / (hiy - loy), cells / (hiz - loz)};
}

__device__ __forceinline__ uint32_t bin_key(const BinBox &bb, const Ray_C *__restrict__ rays, int64_t i) {
        const float *o = reinterpret_cast<const float *>(rays + i);
        constexpr uint32_t cells = 1u << kBinAxisBits;
        const uint32_t x = cell_of(o[0], bb.lox, bb.sx, cells);
        const uint32_t y = cell_of(o[1], bb.loy, bb.sy, cells);
        const uint32_t z = cell_of(o[2], bb.loz, bb.sz, cells);
        return spread3_4(x) | (spread3_4(y) << 1) | (spread3_4(z) << 2);
}

__global__ __launch_bounds__(kBinBlock) void k_bin_count(const NodePair *__restrict__ nodes, int rootDesc,
                                                         const Ray_C *__restrict__ rays, int64_t n,
                                                         uint16_t *__restrict__ keys, uint32_t *__restrict__ counts,
                                                         uint32_t nBlocks, uint32_t per) {
        __shared__ uint32_t hist[kBins];
        for (uint32_t k = threadIdx.x; k < kBins; k += kBinBlock) hist[k] = 0u;
        __syncthreads();
        const BinBox bb = bin_box(nodes, rootDesc);
        const uint32_t lb = bin_block(blockIdx.x, nBlocks);
        const int64_t base = (int64_t)lb * per;
        for (uint32_t j = threadIdx.x; j < per; j += kBinBlock) {
                const int64_t i = base + j;
                if (i >= n) break;
                const uint32_t key = bin_key(bb, rays, i);
                keys[i] = (uint16_t)key;
                atomicAdd(&hist[key], 1u);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < kBins; k += kBinBlock) counts[(size_t)k * nBlocks + lb] = hist[k];
}

// Exclusive scan of m values: k_scan_reduce sums each 4096-value tile, k_scan_top scans the tile
// sums (one block), k_scan_down scans each tile and adds its tile's offset.
constexpr int kScanTile = 4096;

__device__ __forceinline__ uint32_t block_exclusive_1024(uint32_t v, uint32_t *part, uint32_t &total) {
        const uint32_t t = threadIdx.x;
        part[t] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024u; off <<= 1) {
                const uint32_t w = t >= off ? part[t - off] : 0u;
                __syncthreads();
                part[t] += w;
                __syncthreads();
        }
        total = part[1023];
        const uint32_t r = part[t] - v;
        __syncthreads();
        return r;
}

__global__ __launch_bounds__(1024) void k_scan_reduce(const uint32_t *__restrict__ v, int64_t m,
                                                      uint32_t *__restrict__ tileSums) {
        __shared__ uint32_t part[1024];
        const int64_t b = (int64_t)blockIdx.x * kScanTile + 4 * (int64_t)threadIdx.x;
        uint32_t s = 0u;
        for (int k = 0; k < 4; ++k)
                if (b + k < m) s += v[b + k];
        uint32_t total;
        (void)block_exclusive_1024(s, part, total);
        if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t *__restrict__ tileSums, uint32_t tiles) {
        __shared__ uint32_t part[1024];
        const uint32_t per = (tiles + 1023u) / 1024u;
        const uint32_t b = threadIdx.x * per;
        uint32_t s = 0u;
        for (uint32_t k = 0; k < per; ++k)
                if (b + k < tiles) s += tileSums[b + k];
        uint32_t total;
        uint32_t run = block_exclusive_1024(s, part, total);
        for (uint32_t k = 0; k < per; ++k)
                if (b + k < tiles) {
                        const uint32_t c = tileSums[b + k];
                        tileSums[b + k] = run;
                        run += c;
                }
}

__global__ __launch_bounds__(1024) void k_scan_down(uint32_t *__restrict__ v, int64_t m,
                                                    const uint32_t *__restrict__ tileOffsets) {
        __shared__ uint32_t part[1024];
        const int64_t b = (int64_t)blockIdx.x * kScanTile + 4 * (int64_t)threadIdx.x;
        uint32_t x[4], s = 0u;
        for (int k = 0; k < 4; ++k) {
                x[k] = b + k < m ? v[b + k] : 0u;
                s += x[k];
        }
        uint32_t total;
        uint32_t run = block_exclusive_1024(s, part, total) + tileOffsets[blockIdx.x];
        for (int k = 0; k < 4; ++k)
                if (b + k < m) {
                        v[b + k] = run;
                        run += x[k];
                }
}

__global__ __launch_bounds__(kBinBlock) void k_bin_place(const uint16_t *__restrict__ keys,
                                                         const Ray_C *__restrict__ rays,
                                                         const float *__restrict__ tMaxIn, int64_t n,
                                                         const uint32_t *__restrict__ offsets, uint32_t nBlocks,
                                                         uint32_t per, SortedRay *__restrict__ sorted) {
        __shared__ uint32_t next[kBins];
        const uint32_t lb = bin_block(blockIdx.x, nBlocks);
        for (uint32_t k = threadIdx.x; k < kBins; k += kBinBlock) next[k] = offsets[(size_t)k * nBlocks + lb];
        __syncthreads();
        const int64_t base = (int64_t)lb * per;
        for (uint32_t j = threadIdx.x; j < per; j += kBinBlock) {
                const int64_t i = base + j;
                if (i >= n) break;
                const Ray_C r = rays[i];
                SortedRay o;
                o.a = make_float4(r.orgX, r.orgY, r.orgZ, r.dirX);
                o.b = make_float4(r.dirY, r.dirZ, tMaxIn[i], __uint_as_float((uint32_t)i));
                sorted[atomicAdd(&next[keys[i]], 1u)] = o;
        }
}

__global__ void k_reset_counter(unsigned long long *counter) { *counter = 0ull; }
__global__ void k_reset_counters(unsigned long long *counters) {
        if (threadIdx.x < 8) counters[threadIdx.x] = 0ull;
}

// Empty scene (no BVH2 nodes): every ray misses (the guard of BoundingHierarchy.swift:353,392).
__global__ __launch_bounds__(kBlock) void k_miss(const float *__restrict__ tMaxIn, int64_t n, GzHit *__restrict__ hits,
                                                 uint8_t *__restrict__ occ, GzTraceStats *__restrict__ stats) {
        const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        if (hits) hits[i] = GzHit{-1, tMaxIn[i], 0.0f, 0.0f};
        if (occ) occ[i] = 0;
        if (stats) stats[i] = GzTraceStats{0u, 0u};
}

// kernel.mojo:360-364: Intersection_C from a compact hit.
__global__ __launch_bounds__(kBlock) void k_expand(const GzHit *__restrict__ hits, const float *__restrict__ tMaxIn,
                                                   const PrimId_C *__restrict__ prims, int64_t primCount,
                                                   int64_t n, Intersection_C *__restrict__ out) {
        const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        if (i >= n) return;
        const GzHit h = hits[i];
        Intersection_C r;
        memset(&r, 0, sizeof r);
        if (h.prim >= 0 && h.prim < primCount) {
                r.primId = prims[h.prim];
                r.tHit = h.t;
                r.u = h.u;
                r.v = h.v;
                r.hit = 1;
        } else {
                r.primId.id1 = -1;
                r.primId.id2 = -1;
                r.tHit = tMaxIn[i];
        }
        out[i] = r;
}

static inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <bool kAny, bool kStats, int kLds>
static void launch_variant(Scene *s, const StreamSlot &slot, const Ray_C *d_rays, const float *d_tMax, int64_t n,
                           GzHit *d_hits, uint8_t *d_occ, GzTraceStats *d_stats, hipStream_t stream,
                           const SortedRay *order) {
        const GzTraceConfig &c = s->traceConfig;
        if (c.persistent == 2) {
                auto kern = k_trace_strided<kAny, kStats, kLds>;
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
                const unsigned g = (unsigned)std::min<int64_t>(want, grid_for(n));
                hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, stream, s->dNodes, s->dTris, s->rootDesc,
                                   s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats, slot.dErr, order);
        } else if (c.persistent == 3) {
                auto kern = k_trace_xcd<kAny, kStats, kLds>;
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
                const unsigned g = (unsigned)std::min<int64_t>(want, grid_for(n));
                hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, stream, slot.dCounter);
                hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, stream, s->dNodes, s->dTris, s->rootDesc,
                                   s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats, slot.dErr, slot.dCounter,
                                   c.refillLanes > 0 && c.refillLanes <= 64 ? c.refillLanes : 16,
                                   c.chunk >= 1 && c.chunk <= 4096 ? c.chunk : 64, order);
        } else if (c.persistent == 1) {
                auto kern = k_trace_persistent<kAny, kStats, kLds>;
                int perCU = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, kBlock, 0) != hipSuccess || perCU < 1)
                        perCU = 1;
                const int64_t want = (int64_t)s->cuCount * perCU * (c.gridScale > 0 ? c.gridScale : 1);
                const unsigned g = (unsigned)std::min<int64_t>(want, grid_for(n));
                unsigned long long *ctr = slot.dCounter;
                hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, ctr);
                hipLaunchKernelGGL(kern, dim3(g), dim3(kBlock), 0, stream, s->dNodes, s->dTris, s->rootDesc,
                                   s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats, slot.dErr, ctr,
                                   c.refillLanes > 0 && c.refillLanes <= 64 ? c.refillLanes : 16,
                                   c.chunk >= 1 && c.chunk <= 4096 ? c.chunk : 64, order);
        } else {
                hipLaunchKernelGGL((k_trace<kAny, kStats, kLds>), dim3(grid_for(n)), dim3(kBlock), 0, stream, s->dNodes,
                                   s->dTris, s->rootDesc, s->boundsSafe, d_rays, d_tMax, n, d_hits, d_occ, d_stats,
                                   slot.dErr, order);
        }
}

template <bool kAny, bool kStats>
static void launch_lds(Scene *s, const StreamSlot &slot, const Ray_C *d_rays, const float *d_tMax, int64_t n,
                       GzHit *d_hits, uint8_t *d_occ, GzTraceStats *d_stats, hipStream_t stream,
                       const SortedRay *order) {
        if (s->traceConfig.ldsStack >= 16)
                launch_variant<kAny, kStats, 16>(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream, order);
        else if (s->traceConfig.ldsStack >= 8)
                launch_variant<kAny, kStats, 8>(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream, order);
        else
                launch_variant<kAny, kStats, 0>(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, stream, order);
}

// The launches of one traversal with the error flag and ray counter of `slot`, tracing the rays of
// the sorted batch `order` (null: the caller's arrays in their order).
static int launch_trace_slot(Scene *s, const StreamSlot *slot, const Ray_C *d_rays, const float *d_tMax, int64_t n,
                             GzHit *d_hits, uint8_t *d_occ, GzTraceStats *d_stats, bool anyHit, hipStream_t stream,
                             const SortedRay *order = nullptr) {
        if (s->empty) {
                hipLaunchKernelGGL(k_miss, dim3(grid_for(n)), dim3(kBlock), 0, stream, d_tMax, n,
                                   anyHit ? nullptr : d_hits, anyHit ? d_occ : nullptr, d_stats);
        } else if (anyHit) {
                if (d_stats)
                        launch_lds<true, true>(s, *slot, d_rays, d_tMax, n, nullptr, d_occ, d_stats, stream, order);
                else
                        launch_lds<true, false>(s, *slot, d_rays, d_tMax, n, nullptr, d_occ, nullptr, stream, order);
        } else {
                if (d_stats)
                        launch_lds<false, true>(s, *slot, d_rays, d_tMax, n, d_hits, nullptr, d_stats, stream, order);
                else
                        launch_lds<false, false>(s, *slot, d_rays, d_tMax, n, d_hits, nullptr, nullptr, stream, order);
        }
        return hip_ok(hipGetLastError(), "traversal launch") ? GZ_OK : GZ_E_HIP;
}

static bool trace_size_ok(int64_t n) {
        if (n <= (int64_t)kBlock * 0x7fffffffLL) return true;
        set_error("traversal batch of %lld rays is too large", (long long)n);
        return false;
}

// Bytes of the binning storage for n rays: the sorted batch (32 B per ray), keys (2 B per ray,
// padded), counts[key][block], the scan's tile sums.
static size_t bin_bytes(int64_t n, uint32_t per, uint32_t &nBlocks, int64_t &countEntries, uint32_t &tiles) {
        nBlocks = (uint32_t)((n + per - 1) / per);
        countEntries = (int64_t)kBins * nBlocks;
        tiles = (uint32_t)((countEntries + kScanTile - 1) / kScanTile);
        return sizeof(SortedRay) * (size_t)n + ((2 * (size_t)n + 255) & ~(size_t)255) + 4 * (size_t)countEntries +
               4 * (size_t)tiles;
}

// The batch sorted by origin cell in the slot's binning storage, or null when the batch is
// traced as given (binning off, below binMinRays, no interior node, more than 2^32 - 1 rays) or its
// storage cannot be had (error set: `failed`).
static const SortedRay *bin_rays(Scene *s, StreamSlot *slot, const Ray_C *d_rays, const float *d_tMax, int64_t n,
                                 hipStream_t stream, bool &failed) {
        failed = false;
        const int32_t minRays = s->traceConfig.binMinRays;
        if (minRays <= 0 || n < minRays || s->empty || s->rootDesc < 0 || n > 0xffffffffLL) return nullptr;
        uint32_t nBlocks, tiles;
        int64_t entries;
        const uint32_t per = bin_per(n);
        const size_t bytes = bin_bytes(n, per, nBlocks, entries, tiles);
        if (tiles > 1024u * 1024u) return nullptr;  // beyond the one-block scan of tile sums
        if (slot->binBytes < bytes) {
                if (slot->dBin) {
                        s->deviceBytes -= (int64_t)slot->binBytes;
                        (void)hipFree(slot->dBin);
                        slot->dBin = nullptr;
                        slot->binBytes = 0;
                }
                void *p = nullptr;
                if (alloc_fault("ray binning") || !hip_ok(hipMalloc(&p, bytes), "hipMalloc(ray binning)")) {
                        failed = true;
                        return nullptr;
                }
                slot->dBin = static_cast<uint32_t *>(p);
                slot->binBytes = bytes;
                s->deviceBytes += (int64_t)bytes;
        }
        SortedRay *sorted = reinterpret_cast<SortedRay *>(slot->dBin);
        uint16_t *keys = reinterpret_cast<uint16_t *>(sorted + n);
        uint32_t *counts = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(keys) + ((2 * (size_t)n + 255) & ~(size_t)255));
        uint32_t *tileSums = counts + entries;
        hipLaunchKernelGGL(k_bin_count, dim3(nBlocks), dim3(kBinBlock), 0, stream, s->dNodes, s->rootDesc, d_rays, n,
                           keys, counts, nBlocks, per);
        hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(1024), 0, stream, counts, entries, tileSums);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, tileSums, tiles);
        hipLaunchKernelGGL(k_scan_down, dim3(tiles), dim3(1024), 0, stream, counts, entries, tileSums);
        hipLaunchKernelGGL(k_bin_place, dim3(nBlocks), dim3(kBinBlock), 0, stream, keys, d_rays, d_tMax, n, counts,
                           nBlocks, per, sorted);
        return sorted;
}

int launch_trace(Scene *s, const Ray_C *d_rays, const float *d_tMax, int64_t n, GzHit *d_hits, uint8_t *d_occ,
                 GzTraceStats *d_stats, bool anyHit, hipStream_t stream) {
        if (n <= 0) return GZ_OK;
        if (!trace_size_ok(n)) return GZ_E_INVALID;
        StreamSlot *slot = stream_slot(s, stream);
        if (!slot) return GZ_E_ALLOC;
        const bool timing = s->traceConfig.recordTimes == 1;
        if (timing) {
                for (hipEvent_t &e : slot->ev)
                        if (!e && !hip_ok(hipEventCreate(&e), "hipEventCreate")) return GZ_E_HIP;
                if (!hip_ok(hipEventRecord(slot->ev[0], stream), "hipEventRecord")) return GZ_E_HIP;
        }
        bool failed = false;
        const SortedRay *order = bin_rays(s, slot, d_rays, d_tMax, n, stream, failed);
        if (failed) return GZ_E_ALLOC;
        if (timing && !hip_ok(hipEventRecord(slot->ev[1], stream), "hipEventRecord")) return GZ_E_HIP;
        const int rc = launch_trace_slot(s, slot, d_rays, d_tMax, n, d_hits, d_occ, d_stats, anyHit, stream, order);
        if (rc == GZ_OK && timing) {
                if (!hip_ok(hipEventRecord(slot->ev[2], stream), "hipEventRecord")) return GZ_E_HIP;
                slot->timed = true;
        }
        return rc;
}

int launch_expand(Scene *s, const GzHit *d_hits, const float *d_tMax, int64_t n, Intersection_C *d_out,
                  hipStream_t stream) {
        if (n <= 0) return GZ_OK;
        hipLaunchKernelGGL(k_expand, dim3(grid_for(n)), dim3(kBlock), 0, stream, d_hits, d_tMax, s->dPrims, s->primCount, n, d_out);
        return hip_ok(hipGetLastError(), "expand launch") ? GZ_OK : GZ_E_HIP;
}

}  // namespace gz

using namespace gz;

static Scene *scene_of(void *h, const char *fn) {
        if (!h) {
                set_error("%s: null scene handle", fn);
                return nullptr;
        }
        return static_cast<Scene *>(h);
}

static int check_overflow_sync(Scene *s) { return check_stream(s, s->stream); }

extern "C" int gz_gpu_traverse_closest_device(void *handle, const Ray_C *d_rays, const float *d_tMax, int64_t count,
                                              GzHit *d_hits, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_closest_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_hits))) {
                set_error("gz_gpu_traverse_closest_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, d_hits, nullptr, nullptr, false, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_traverse_closest_stats_device(void *handle, const Ray_C *d_rays, const float *d_tMax,
                                                    int64_t count, GzHit *d_hits, GzTraceStats *d_stats,
                                                    void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_closest_stats_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_hits))) {
                set_error("gz_gpu_traverse_closest_stats_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, d_hits, nullptr, d_stats, false, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_traverse_any_device(void *handle, const Ray_C *d_rays, const float *d_tMax, int64_t count,
                                          uint8_t *d_occluded, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_any_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_occluded))) {
                set_error("gz_gpu_traverse_any_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, nullptr, d_occluded, nullptr, true, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_traverse_any_stats_device(void *handle, const Ray_C *d_rays, const float *d_tMax, int64_t count,
                                                uint8_t *d_occluded, GzTraceStats *d_stats, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_any_stats_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_rays || !d_tMax || !d_occluded))) {
                set_error("gz_gpu_traverse_any_stats_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_trace(s, d_rays, d_tMax, count, nullptr, d_occluded, d_stats, true, static_cast<hipStream_t>(stream));
}

extern "C" int gz_gpu_trace_times(void *handle, void *stream, float *binMs, float *traceMs) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_trace_times");
        if (!s) return GZ_E_INVALID;
        if (!binMs || !traceMs) {
                set_error("gz_gpu_trace_times: null output");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        auto it = s->slots.find(static_cast<hipStream_t>(stream));
        if (it == s->slots.end() || !it->second.timed) {
                set_error("gz_gpu_trace_times: no timed traversal on this stream (set recordTimes)");
                return GZ_E_INVALID;
        }
        const StreamSlot &slot = it->second;
        if (!hip_ok(hipEventSynchronize(slot.ev[2]), "hipEventSynchronize") ||
            !hip_ok(hipEventElapsedTime(binMs, slot.ev[0], slot.ev[1]), "hipEventElapsedTime") ||
            !hip_ok(hipEventElapsedTime(traceMs, slot.ev[1], slot.ev[2]), "hipEventElapsedTime"))
                return GZ_E_HIP;
        return GZ_OK;
}

extern "C" int gz_gpu_expand_hits_device(void *handle, const GzHit *d_hits, const float *d_tMax, int64_t count,
                                         Intersection_C *d_results, void *stream) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_expand_hits_device");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!d_hits || !d_tMax || !d_results))) {
                set_error("gz_gpu_expand_hits_device: bad arguments");
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        return launch_expand(s, d_hits, d_tMax, count, d_results, static_cast<hipStream_t>(stream));
}

// Host-buffer batch: stage rays + tMax, trace, expand to Intersection_C, copy back, wait.
extern "C" void gz_gpu_traverse_batch(void *handle, const Ray_C *rays, const float *tMaxValues, int64_t count,
                                      Intersection_C *results) {
        clear_error();
        if (!handle || count <= 0) return;  // kernel.mojo:616-622
        if (!rays || !tMaxValues || !results) {
                set_error("gz_gpu_traverse_batch: null buffer");
                return;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return;
        const size_t n = (size_t)count;
        const size_t offT = (n * sizeof(Ray_C) + 255) & ~(size_t)255;
        const size_t offH = (offT + n * sizeof(float) + 255) & ~(size_t)255;
        const size_t offH2 = (offH + n * sizeof(GzHit) + 255) & ~(size_t)255;
        const size_t total = offH2 + n * sizeof(Intersection_C);
        if (!ensure_stage(s, total)) return;
        char *base = static_cast<char *>(s->dStage);
        Ray_C *dR = reinterpret_cast<Ray_C *>(base);
        float *dT = reinterpret_cast<float *>(base + offT);
        GzHit *dH = reinterpret_cast<GzHit *>(base + offH);
        Intersection_C *dI = reinterpret_cast<Intersection_C *>(base + offH2);
        hipStream_t st = s->stream;
        if (!hip_ok(hipMemcpyAsync(dR, rays, n * sizeof(Ray_C), hipMemcpyHostToDevice, st), "copy rays") ||
            !hip_ok(hipMemcpyAsync(dT, tMaxValues, n * sizeof(float), hipMemcpyHostToDevice, st), "copy tMax"))
                return;
        if (launch_trace(s, dR, dT, count, dH, nullptr, nullptr, false, st) != GZ_OK) return;
        if (launch_expand(s, dH, dT, count, dI, st) != GZ_OK) return;
        if (check_overflow_sync(s) != GZ_OK) return;
        if (!hip_ok(hipMemcpyAsync(results, dI, n * sizeof(Intersection_C), hipMemcpyDeviceToHost, st), "copy results"))
                return;
        (void)hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
}

extern "C" int gz_gpu_occluded_batch(void *handle, const Ray_C *rays, const float *tMaxValues, int64_t count,
                                     uint8_t *occluded) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_occluded_batch");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!rays || !tMaxValues || !occluded))) {
                set_error("gz_gpu_occluded_batch: bad arguments");
                return GZ_E_INVALID;
        }
        if (count == 0) return GZ_OK;
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        const size_t n = (size_t)count;
        const size_t offT = (n * sizeof(Ray_C) + 255) & ~(size_t)255;
        const size_t offO = (offT + n * sizeof(float) + 255) & ~(size_t)255;
        if (!ensure_stage(s, offO + n)) return GZ_E_ALLOC;
        char *base = static_cast<char *>(s->dStage);
        Ray_C *dR = reinterpret_cast<Ray_C *>(base);
        float *dT = reinterpret_cast<float *>(base + offT);
        uint8_t *dO = reinterpret_cast<uint8_t *>(base + offO);
        hipStream_t st = s->stream;
        if (!hip_ok(hipMemcpyAsync(dR, rays, n * sizeof(Ray_C), hipMemcpyHostToDevice, st), "copy rays") ||
            !hip_ok(hipMemcpyAsync(dT, tMaxValues, n * sizeof(float), hipMemcpyHostToDevice, st), "copy tMax"))
                return GZ_E_HIP;
        int rc = launch_trace(s, dR, dT, count, nullptr, dO, nullptr, true, st);
        if (rc != GZ_OK) return rc;
        rc = check_overflow_sync(s);
        if (rc != GZ_OK) return rc;
        if (!hip_ok(hipMemcpyAsync(occluded, dO, n, hipMemcpyDeviceToHost, st), "copy occluded") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        return GZ_OK;
}

// ---- Allocation-free host-buffer batch ----------------------------------------------------------
//
// gz_gpu_traverse_batch_noalloc stages its rays, tMax, hits and records, its error flag and the
// persistent kernel's ray counter in `scratch`, a file-scope device array of fixed size that every
// GPU the library runs on holds (its own copy, at its own address) until the process ends. In host
// code `scratch` is a stand-in for the symbol, not a device address: it is never passed to a
// kernel or a copy. hipGetSymbolAddress gives each device's address once, with that device
// current, and only that address is used. Layout (each part 256-byte aligned):
//   [0, 256) : the error flag (int at 0) and the ray counter (u64 at 8)
//   rays, tMax, hits, Intersection_C records for up to GZ_NOALLOC_MAX_RAYS rays
namespace gz {

static constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct NoAllocLayout {
        size_t rays, tMax, hits, records, end;
};

static constexpr NoAllocLayout noalloc_layout(size_t n) {
        const size_t r = 256;
        const size_t t = align256(r + n * sizeof(Ray_C));
        const size_t h = align256(t + n * sizeof(float));
        const size_t i = align256(h + n * sizeof(GzHit));
        return NoAllocLayout{r, t, h, i, i + n * sizeof(Intersection_C)};
}

constexpr size_t kNoAllocScratchBytes = align256(noalloc_layout((size_t)GZ_NOALLOC_MAX_RAYS).end);
static_assert(kNoAllocScratchBytes <= (size_t)48 << 20, "scratch stays within 48 MiB");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex gScratchMu[kMaxDevices];   // one call at a time per device owns its scratch
static void *gScratchPtr[kMaxDevices];       // device address of `scratch` per device, or null

__global__ void k_clear_flag(int *flag) { *flag = 0; }

// The device address of `scratch` on the current device `dev` (the caller holds gScratchMu[dev]).
static void *scratch_address(int dev) {
        if (!gScratchPtr[dev]) {
                void *p = nullptr;
                if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || !p)
                        return nullptr;
                gScratchPtr[dev] = p;
        }
        return gScratchPtr[dev];
}

}  // namespace gz

extern "C" int gz_gpu_traverse_batch_noalloc(void *handle, const Ray_C *rays, const float *tMaxValues, int64_t count,
                                             Intersection_C *results) {
        clear_error();
        Scene *s = scene_of(handle, "gz_gpu_traverse_batch_noalloc");
        if (!s) return GZ_E_INVALID;
        if (count < 0 || (count > 0 && (!rays || !tMaxValues || !results))) {
                set_error("gz_gpu_traverse_batch_noalloc: bad arguments");
                return GZ_E_INVALID;
        }
        if (count > GZ_NOALLOC_MAX_RAYS) {
                set_error("gz_gpu_traverse_batch_noalloc: %lld rays exceed the fixed staging of %d rays",
                          (long long)count, GZ_NOALLOC_MAX_RAYS);
                return GZ_E_INVALID;
        }
        if (count == 0) return GZ_OK;
        if (s->device < 0 || s->device >= kMaxDevices) {
                set_error("gz_gpu_traverse_batch_noalloc: device %d is beyond the %d this entry serves", s->device,
                          kMaxDevices);
                return GZ_E_INVALID;
        }
        std::lock_guard<std::mutex> dlock(gScratchMu[s->device]);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        char *base = static_cast<char *>(scratch_address(s->device));
        if (!base) return GZ_E_HIP;
        const size_t n = (size_t)count;
        const NoAllocLayout L = noalloc_layout(n);
        StreamSlot slot{reinterpret_cast<int *>(base), reinterpret_cast<unsigned long long *>(base + 8)};
        Ray_C *dR = reinterpret_cast<Ray_C *>(base + L.rays);
        float *dT = reinterpret_cast<float *>(base + L.tMax);
        GzHit *dH = reinterpret_cast<GzHit *>(base + L.hits);
        Intersection_C *dI = reinterpret_cast<Intersection_C *>(base + L.records);
        hipStream_t st = s->stream;
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, st, slot.dErr);
        if (!hip_ok(hipGetLastError(), "clear error flag") ||
            !hip_ok(hipMemcpyAsync(dR, rays, n * sizeof(Ray_C), hipMemcpyHostToDevice, st), "copy rays") ||
            !hip_ok(hipMemcpyAsync(dT, tMaxValues, n * sizeof(float), hipMemcpyHostToDevice, st), "copy tMax"))
                return GZ_E_HIP;
        int rc = launch_trace_slot(s, &slot, dR, dT, count, dH, nullptr, nullptr, false, st);
        if (rc != GZ_OK) return rc;
        rc = launch_expand(s, dH, dT, count, dI, st);
        if (rc != GZ_OK) return rc;
        int flag = 0;
        if (!hip_ok(hipMemcpyAsync(&flag, slot.dErr, sizeof(int), hipMemcpyDeviceToHost, st), "read error flag") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        if (flag & 1) {
                set_error("traversal stack overflow: a ray needed more than %d entries", GZ_STACK_CAPACITY);
                return GZ_E_OVERFLOW;
        }
        if (!hip_ok(hipMemcpyAsync(results, dI, n * sizeof(Intersection_C), hipMemcpyDeviceToHost, st), "copy results") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
                return GZ_E_HIP;
        return GZ_OK;
}
