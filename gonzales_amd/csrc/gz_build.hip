// BVH construction on the device: a linear BVH (Morton order, Karras 2012 hierarchy) written
// straight into the traversal kernels' layout. The reference builds its SAH BVH2 on the host
// (Accelerators/BoundingHierarchyBuilder.swift:199-339) and uploads it; here only the primIds
// and the meshes go to the device and the tree is made there:
//
//   gather   one lane per primitive: PrimId -> mesh -> vertexIndices -> points (the decode of
//            kernel.mojo:266-300), the triangle's box (Triangle.worldBound) and the box centroid;
//            the centroids' bounds by wave reduction + one atomic per wave;
//   morton   63-bit codes (21 bits per axis) of the centroids in their bounds;
//   sort     LSD radix sort of (code, primitive) pairs, 8 passes of 8 bits: per-block digit counts,
//            one exclusive scan, a stable scatter ranked by wave ballots (no block waits on
//            another); the sorted order is the leaf order;
//   hierarchy one lane per internal node i of the N - 1: its key range and split by the longest
//            common prefix of the codes (ties broken by position), children and parents;
//   validate every internal node has one parent (the root none), every leaf one, and each child's
//            range lies strictly inside its parent's: a tree, so every climb and descent ends;
//            a scene that fails is rejected;
//   refit    one lane per leaf climbs to the root; at each node the second child to arrive
//            (an atomic counter) writes the union of the two child boxes (exact min/max);
//   emit     every internal node whose range holds more than maxLeaf primitives becomes a
//            64-byte child-pair record; a child range of at most maxLeaf primitives becomes a
//            leaf: its primitives are consecutive in the sorted order, so the leaf is a run of
//            48-byte triangle records. primIds are stored in the same order, so a hit's
//            primitive index is a leaf-order position, as after a host build.
//
// Every loop is bounded (climbs and walks by the node count, with every index checked before it
// is used), no lane waits on another, a node record is written only by its own lane and a leaf
// count only by the leaf's parent.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gz_internal.h"

namespace gz {

constexpr int kBuildBlock = 256;

// flags[0] bits
enum : int { kFlagUnsafe = 1, kFlagSkip = 2, kFlagBadIndex = 4, kFlagBadTree = 8 };

struct Box {
        float lo[3], hi[3];
};

// Order-preserving map of a float to an unsigned key (for atomicMin/atomicMax on floats).
__device__ __forceinline__ unsigned fkey(float f) {
        const unsigned b = __float_as_uint(f);
        return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fval(unsigned k) {
        return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
        v &= 0x1fffffull;
        v = (v | (v << 32)) & 0x1f00000000ffffull;
        v = (v | (v << 16)) & 0x1f0000ff0000ffull;
        v = (v | (v << 8)) & 0x100f00f00f00f00full;
        v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
        v = (v | (v << 2)) & 0x1249249249249249ull;
        return v;
}

// Primitive k: its vertices (w = 0), its box and its centroid; unknown or non-triangle
// primitives (the reference skips them in a leaf) get an empty box and are flagged. Indices
// outside the mesh arrays set err bit 4 (the handle is then rejected).
__global__ __launch_bounds__(kBuildBlock) void k_gather(const PrimId_C *__restrict__ prims, int n,
                                                        const DevMesh *__restrict__ meshes, int64_t nMeshes,
                                                        float4 *__restrict__ verts, Box *__restrict__ boxes,
                                                        float4 *__restrict__ cent, unsigned *__restrict__ cb,
                                                        int *__restrict__ flags) {
        const int k = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        bool valid = false;
        float c[3] = {0.0f, 0.0f, 0.0f};
        if (k < n) {
                const PrimId_C pr = prims[k];
                int64_t mesh = -1, base = -1;
                if (pr.type == 0) {
                        mesh = pr.id1;
                        base = pr.id2;
                } else if ((pr.type == 1 || pr.type == 2 || pr.type == 3) && pr.id2 != -1) {
                        mesh = pr.id2 >> 32;
                        base = (pr.id2 & 0xFFFFFFFFLL) * 3;
                }
                Box b;
                float v[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                int skip = 0;
                if (mesh < 0 && base < 0) {
                        skip = kTriSkip;
                } else if (mesh < 0 || mesh >= nMeshes || base < 0 || base + 2 >= meshes[mesh].vertexIndexCount) {
                        atomicOr(&flags[0], kFlagBadIndex);
                        skip = kTriSkip;
                } else {
                        const DevMesh m = meshes[mesh];
                        for (int q = 0; q < 3; ++q) {
                                const int64_t vi = m.vertexIndices[base + q];
                                if (vi < 0 || vi * 4 + 2 >= m.pointFloats) {
                                        atomicOr(&flags[0], kFlagBadIndex);
                                        skip = kTriSkip;
                                        break;
                                }
                                v[q][0] = m.points[vi * 4];
                                v[q][1] = m.points[vi * 4 + 1];
                                v[q][2] = m.points[vi * 4 + 2];
                        }
                }
                if (skip) {
                        for (int a = 0; a < 3; ++a) {
                                b.lo[a] = INFINITY;
                                b.hi[a] = -INFINITY;
                        }
                } else {
                        for (int a = 0; a < 3; ++a) {
                                b.lo[a] = fminf(v[0][a], fminf(v[1][a], v[2][a]));
                                b.hi[a] = fmaxf(v[0][a], fmaxf(v[1][a], v[2][a]));
                                c[a] = 0.5f * (b.lo[a] + b.hi[a]);
                                // the traversal's fast slab form needs finite boxes within 2^40
                                if (!(b.lo[a] >= -0x1p40f && b.hi[a] <= 0x1p40f)) atomicOr(&flags[0], kFlagUnsafe);
                        }
                        valid = true;
                }
                if (skip) atomicOr(&flags[0], kFlagSkip);
                verts[3 * (size_t)k] = make_float4(v[0][0], v[0][1], v[0][2], 0.0f);
                verts[3 * (size_t)k + 1] = make_float4(v[1][0], v[1][1], v[1][2], 0.0f);
                verts[3 * (size_t)k + 2] = make_float4(v[2][0], v[2][1], v[2][2], __int_as_float(skip));
                boxes[k] = b;
                cent[k] = make_float4(c[0], c[1], c[2], valid ? 1.0f : 0.0f);
        }
        // centroid bounds: wave min/max, one atomic per wave and axis
        for (int a = 0; a < 3; ++a) {
                unsigned lo = valid ? fkey(c[a]) : 0xffffffffu, hi = valid ? fkey(c[a]) : 0u;
                for (int off = 32; off >= 1; off >>= 1) {
                        lo = min(lo, (unsigned)__shfl_xor((int)lo, off));
                        hi = max(hi, (unsigned)__shfl_xor((int)hi, off));
                }
                if ((threadIdx.x & 63) == 0) {
                        atomicMin(&cb[a], lo);
                        atomicMax(&cb[3 + a], hi);
                }
        }
}

__global__ void k_build_init(unsigned *cb, int *flags) {
        if (threadIdx.x < 3) {
                cb[threadIdx.x] = 0xffffffffu;
                cb[3 + threadIdx.x] = 0u;
        }
        if (threadIdx.x == 0) {
                flags[0] = 0;
                flags[1] = 0;  // emitted interior nodes
                flags[2] = 0;  // deepest interior node (edges from the root)
                flags[3] = 0;  // leaves
        }
}

__global__ __launch_bounds__(kBuildBlock) void k_morton(const float4 *__restrict__ cent, int n,
                                                        const unsigned *__restrict__ cb, uint64_t *__restrict__ keys,
                                                        uint32_t *__restrict__ vals) {
        const int k = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        if (k >= n) return;
        const float4 c = cent[k];
        uint64_t code = 0x7fffffffffffffffull;  // skipped primitives go last
        if (c.w != 0.0f) {
                const float p[3] = {c.x, c.y, c.z};
                uint64_t q[3];
                for (int a = 0; a < 3; ++a) {
                        const float lo = fval(cb[a]), hi = fval(cb[3 + a]);
                        const float ext = hi - lo;
                        float t = ext > 0.0f ? (p[a] - lo) / ext : 0.0f;
                        t = fminf(fmaxf(t, 0.0f), 1.0f);
                        q[a] = (uint64_t)fminf(t * 2097152.0f, 2097151.0f);
                }
                code = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
        }
        keys[k] = code;
        vals[k] = (uint32_t)k;
}

// ---- LSD radix sort of (key, value) pairs, 8-bit digits --------------------------------------
constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / 64;

__device__ __forceinline__ unsigned digit_of(uint64_t k, int shift) { return (unsigned)(k >> shift) & 255u; }

// Per block: how many of its keys have each digit; counts is digit-major, [256][nBlocks].
__global__ __launch_bounds__(kSortBlock) void k_digit_counts(const uint64_t *__restrict__ keys, int n, int shift,
                                                             unsigned *__restrict__ counts, int nBlocks) {
        __shared__ unsigned c[256];
        for (int k = threadIdx.x; k < 256; k += kSortBlock) c[k] = 0u;
        __syncthreads();
        const int i = (int)(blockIdx.x * kSortBlock + threadIdx.x);
        if (i < n) atomicAdd(&c[digit_of(keys[i], shift)], 1u);
        __syncthreads();
        for (int k = threadIdx.x; k < 256; k += kSortBlock) counts[(size_t)k * nBlocks + blockIdx.x] = c[k];
}

// Exclusive scan of `total` counts in place, by one block: each thread scans a contiguous
// segment, the segment sums are scanned in LDS.
__global__ __launch_bounds__(1024) void k_digit_scan(unsigned *__restrict__ counts, int total) {
        __shared__ unsigned sums[1024];
        const int t = (int)threadIdx.x;
        const int per = (total + 1023) / 1024;
        const int lo = min(t * per, total), hi = min(lo + per, total);
        unsigned s = 0;
        for (int k = lo; k < hi; ++k) s += counts[k];
        sums[t] = s;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
                const unsigned v = t >= off ? sums[t - off] : 0u;
                __syncthreads();
                sums[t] += v;
                __syncthreads();
        }
        unsigned run = sums[t] - s;
        for (int k = lo; k < hi; ++k) {
                const unsigned c = counts[k];
                counts[k] = run;
                run += c;
        }
}

// Stable scatter: a key's slot is its digit's offset for the block + the keys of that digit in
// earlier waves of the block + the lanes before it in its wave with the same digit (lanes with
// the same digit found by one ballot per digit bit).
__global__ __launch_bounds__(kSortBlock) void k_digit_scatter(const uint64_t *__restrict__ keysIn,
                                                              const uint32_t *__restrict__ valsIn, int n, int shift,
                                                              const unsigned *__restrict__ offsets, int nBlocks,
                                                              uint64_t *__restrict__ keysOut,
                                                              uint32_t *__restrict__ valsOut, int *__restrict__ flags) {
        __shared__ unsigned waveCount[kSortWaves][256];
        const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
        for (int k = threadIdx.x; k < kSortWaves * 256; k += kSortBlock) (&waveCount[0][0])[k] = 0u;
        __syncthreads();
        const int i = (int)(blockIdx.x * kSortBlock + threadIdx.x);
        const bool valid = i < n;
        const uint64_t key = valid ? keysIn[i] : 0ull;
        const unsigned d = digit_of(key, shift);
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
                const uint64_t m = __ballot(valid && ((d >> b) & 1u));
                peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const unsigned rank = (unsigned)__popcll(peers & below);
        if (valid && rank == 0u) waveCount[wave][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (!valid) return;
        unsigned before = 0;
        for (int w = 0; w < wave; ++w) before += waveCount[w][d];
        const unsigned dst = offsets[(size_t)d * nBlocks + blockIdx.x] + before + rank;
        if (dst >= (unsigned)n) {
                atomicOr(&flags[0], kFlagBadTree);
                return;
        }
        keysOut[dst] = key;
        valsOut[dst] = valsIn[i];
}

// Length of the common prefix of sorted keys i and j (positions break ties), -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t *__restrict__ keys, int n, int i, int j) {
        if (j < 0 || j >= n) return -1;
        const uint64_t a = keys[i], b = keys[j];
        if (a == b) return 64 + __clz((unsigned)(i ^ j));
        return __clzll((long long)(a ^ b));
}

// Internal node i of the N - 1 (Karras, "Maximizing parallelism in the construction of BVHs,
// octrees and k-d trees", HPG 2012, section 4): range [first, last] and split; child c >= 0 is
// internal node c, c < 0 is leaf ~c (a sorted position).
__global__ __launch_bounds__(kBuildBlock) void k_hierarchy(const uint64_t *__restrict__ keys, int n,
                                                           int4 *__restrict__ node, int *__restrict__ parent,
                                                           int *__restrict__ leafParent, int *__restrict__ parents,
                                                           int *__restrict__ leafParents) {
        const int i = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        if (i >= n - 1) return;
        const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
        const int dmin = delta(keys, n, i, i - d);
        int lmax = 2;
        while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
                if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(keys, n, i, j);
        int s = 0;
        for (int div = 2;; div <<= 1) {
                const int t = (l + div - 1) / div;
                if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
                if (t <= 1) break;
        }
        const int g = i + s * d + (d < 0 ? -1 : 0);
        const int first = min(i, j), last = max(i, j);
        const int left = first == g ? ~g : g;
        const int right = last == g + 1 ? ~(g + 1) : g + 1;
        node[i] = make_int4(first, last, left, right);
        // every index below is in range: first <= g < last lie in [0, n) and g, g + 1 in [0, n - 1]
        if (g < first || g >= last) return;  // no children recorded: the validation rejects the tree
        if (left >= 0) {
                parent[left] = i;
                atomicAdd(&parents[left], 1);
        } else {
                leafParent[g] = i;
                atomicAdd(&leafParents[g], 1);
        }
        if (right >= 0) {
                parent[right] = i;
                atomicAdd(&parents[right], 1);
        } else {
                leafParent[g + 1] = i;
                atomicAdd(&leafParents[g + 1], 1);
        }
}

// A tree check: the root spans [0, n - 1] and has no parent, every other internal node and every
// leaf has exactly one, and each child's range is the left or right part of its parent's split
// (strictly smaller, so no path can come back to a node). Violations set kFlagBadTree.
__global__ __launch_bounds__(kBuildBlock) void k_validate(const int4 *__restrict__ node, int n,
                                                          const int *__restrict__ parents,
                                                          const int *__restrict__ leafParents, int *__restrict__ flags) {
        const int i = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        const int m = n - 1;
        bool bad = false;
        if (i < n && leafParents[i] != 1) bad = true;
        if (i < m) {
                const int4 nd = node[i];
                if (parents[i] != (i == 0 ? 0 : 1)) bad = true;
                if (i == 0 && (nd.x != 0 || nd.y != m)) bad = true;
                if (nd.x < 0 || nd.y >= n || nd.x >= nd.y) {
                        bad = true;
                } else {
                        const int g = nd.z >= 0 ? nd.z : ~nd.z;
                        if (g < nd.x || g >= nd.y) {
                                bad = true;
                        } else {
                                if (nd.z >= 0 ? (nd.z >= m || node[nd.z].x != nd.x || node[nd.z].y != g) : nd.x != g)
                                        bad = true;
                                const int r = nd.w >= 0 ? nd.w : ~nd.w;
                                if (r != g + 1 ||
                                    (nd.w >= 0 ? (nd.w >= m || node[nd.w].x != g + 1 || node[nd.w].y != nd.y) : nd.y != g + 1))
                                        bad = true;
                        }
                }
        }
        if (bad) atomicOr(&flags[0], kFlagBadTree);
}

__device__ __forceinline__ Box box_union(const Box &a, const Box &b) {
        Box r;
        for (int q = 0; q < 3; ++q) {
                r.lo[q] = fminf(a.lo[q], b.lo[q]);
                r.hi[q] = fmaxf(a.hi[q], b.hi[q]);
        }
        return r;
}

// The box of child descriptor c (internal node c, or leaf position ~c); false if c is out of range.
__device__ __forceinline__ bool child_box(int c, const Box *nodeBox, const Box *__restrict__ primBox,
                                          const uint32_t *__restrict__ order, int n, Box &out) {
        if (c >= 0) {
                if (c >= n - 1) return false;
                out = nodeBox[c];
                return true;
        }
        if (~c >= n) return false;
        const uint32_t k = order[~c];
        if (k >= (uint32_t)n) return false;
        out = primBox[k];
        return true;
}

// Whether internal node i becomes (part of) a leaf: a range of at most maxLeaf primitives, or in
// an SAH build one the host builder keeps whole.
__device__ __forceinline__ bool in_leaf(const int4 &nd, int i, int maxLeaf, const int *__restrict__ leafMark) {
        return leafMark ? leafMark[i] != 0 : nd.y - nd.x + 1 <= maxLeaf;
}

// Bottom-up boxes: the lane of leaf position p climbs while it is the second child to arrive.
// With them the number of child-pair records in each subtree (rec), which numbers the records
// depth first in k_emit.
__global__ __launch_bounds__(kBuildBlock) void k_refit(const int4 *__restrict__ node, const int *__restrict__ parent,
                                                       const int *__restrict__ leafParent,
                                                       const uint32_t *__restrict__ order,
                                                       const Box *__restrict__ primBox, Box *nodeBox,
                                                       unsigned *arrivals, int *rec, int maxLeaf,
                                                       const int *__restrict__ leafMark, int n,
                                                       int *__restrict__ flags) {
        const int p = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        if (p >= n) return;
        const int m = n - 1;
        int cur = leafParent[p];
        for (int steps = 0; cur >= 0; ++steps) {
                if (cur >= m || steps >= m) {
                        atomicOr(&flags[0], kFlagBadTree);
                        return;
                }
                __threadfence();
                if (atomicAdd(&arrivals[cur], 1u) == 0u) return;
                __threadfence();
                const int4 nd = node[cur];
                Box a, b;
                if (!child_box(nd.z, nodeBox, primBox, order, n, a) || !child_box(nd.w, nodeBox, primBox, order, n, b)) {
                        atomicOr(&flags[0], kFlagBadTree);
                        return;
                }
                nodeBox[cur] = box_union(a, b);
                const int rl = nd.z >= 0 && nd.z < m ? rec[nd.z] : 0, rr = nd.w >= 0 && nd.w < m ? rec[nd.w] : 0;
                rec[cur] = in_leaf(nd, cur, maxLeaf, leafMark) ? 0 : 1 + rl + rr;
                cur = cur == 0 ? -1 : parent[cur];
        }
}

// Triangle records and primIds in leaf (sorted) order.
__global__ __launch_bounds__(kBuildBlock) void k_sorted(const uint32_t *__restrict__ order,
                                                        const float4 *__restrict__ verts,
                                                        const PrimId_C *__restrict__ primsIn, int n,
                                                        TriRec *__restrict__ tris, PrimId_C *__restrict__ primsOut,
                                                        int *__restrict__ flags) {
        const int j = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        if (j >= n) return;
        uint32_t k = order[j];
        if (k >= (uint32_t)n) {
                atomicOr(&flags[0], kFlagBadTree);
                k = 0;
        }
        TriRec t;
        t.p0 = verts[3 * (size_t)k];
        t.p1 = verts[3 * (size_t)k + 1];
        t.p2 = verts[3 * (size_t)k + 2];
        t.p0.w = __int_as_float(0);
        t.p1.w = __int_as_float(j);
        tris[j] = t;
        primsOut[j] = primsIn[k];
}


// Child-pair records of the internal nodes holding more than maxLeaf primitives, leaf counts in
// the first triangle record of each leaf, and the depth / node / leaf counts.
__global__ __launch_bounds__(kBuildBlock) void k_emit(const int4 *__restrict__ node, const int *__restrict__ parent,
                                                      const Box *__restrict__ nodeBox, const Box *__restrict__ primBox,
                                                      const uint32_t *__restrict__ order, int n, int maxLeaf,
                                                      const int *__restrict__ leafMark, const int *__restrict__ rec,
                                                      NodePair *__restrict__ pairs, TriRec *__restrict__ tris,
                                                      int *__restrict__ flags) {
        const int i = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        const int m = n - 1;
        if (i >= m) return;
        const int4 nd = node[i];
        // a leaf: a range of at most maxLeaf primitives, or (SAH build) one the host builder keeps whole
        if (in_leaf(nd, i, maxLeaf, leafMark)) return;
        // the record's depth-first number: one per edge from the root, plus the records of the
        // left sibling's subtree on every edge that enters a right child
        int pre = 0;
        for (int c = i, steps = 0; c != 0; ++steps) {
                const int p = c >= 0 && c < m ? parent[c] : -1;
                if (p < 0 || p >= m || steps >= m) {
                        atomicOr(&flags[0], kFlagBadTree);
                        return;
                }
                const int4 pn = node[p];
                pre += 1;
                if (pn.w == c && pn.z >= 0 && pn.z < m) pre += rec[pn.z];
                c = p;
        }
        if (pre >= m) {
                atomicOr(&flags[0], kFlagBadTree);
                return;
        }
        int desc[2];
        Box cb[2];
        const int g = nd.z >= 0 ? nd.z : ~nd.z;  // split: left [first, g], right [g + 1, last]
        if (nd.x < 0 || nd.y >= n || g < nd.x || g >= nd.y) {
                atomicOr(&flags[0], kFlagBadTree);
                return;
        }
        for (int side = 0; side < 2; ++side) {
                const int c = side == 0 ? nd.z : nd.w;
                const int cFirst = side == 0 ? nd.x : g + 1;
                const int cLast = side == 0 ? g : nd.y;
                if (!child_box(c, nodeBox, primBox, order, n, cb[side])) {
                        atomicOr(&flags[0], kFlagBadTree);
                        return;
                }
                const int size = cLast - cFirst + 1;
                if (leafMark ? (c < 0 || leafMark[c] != 0) : size <= maxLeaf) {
                        desc[side] = ~cFirst;
                        tris[cFirst].p0.w = __int_as_float(size);
                        atomicAdd(&flags[3], 1);
                } else {
                        // depth-first numbers: the left child follows its parent, the right child
                        // the left child's subtree
                        desc[side] = side == 0 ? pre + 1 : pre + 1 + (nd.z >= 0 && nd.z < m ? rec[nd.z] : 0);
                }
        }
        NodePair p;
        p.a = make_float4(cb[0].lo[0], cb[0].lo[1], cb[0].lo[2], cb[0].hi[0]);
        p.b = make_float4(cb[0].hi[1], cb[0].hi[2], cb[1].lo[0], cb[1].lo[1]);
        p.c = make_float4(cb[1].lo[2], cb[1].hi[0], cb[1].hi[1], cb[1].hi[2]);
        p.d = make_int4(desc[0], desc[1], 0, 0);
        pairs[pre] = p;
        atomicAdd(&flags[1], 1);
        int depth = 0;
        for (int c = i; c != 0; c = parent[c]) {
                if (c < 0 || c >= m || depth >= m) {
                        atomicOr(&flags[0], kFlagBadTree);
                        return;
                }
                ++depth;
        }
        atomicMax(&flags[2], depth + 1);
}

// ---- Binned SAH hierarchy (GzBuildOptions.sah = 1) ----------------------------------------------
//
// The host builder's decisions (bvh_build.cpp after BoundingHierarchyBuilder.swift:199-339) made
// on the device, top-down: a node's bounds and centroid bounds, its longest centroid axis, 12
// buckets, the SAH costs in the same float arithmetic, and the split "bucket > best goes right".
// Every decision depends only on the set of primitives in a node, so the tree's nodes hold the
// same sets as the host build's (the order inside a child differs: a stable partition here, a
// Hoare partition there). Nodes of more than kSahSmall primitives are split level by level, one
// workgroup per node; smaller ones are finished by one lane each. The result is written in the
// hierarchy layout k_hierarchy writes (a node over [first, last] split at g has children g and
// g + 1, a leaf position ~p), which any binary tree over the final primitive order fits, so the
// validation, refit and emit kernels serve both builds. Where the host makes a leaf, this build
// splits the range at its middle: emit turns every node of at most maxLeaf primitives into one
// leaf. A node the SAH cannot split (flat centroids, an empty box) is split at its middle too.
constexpr int kSahSmall = 32;
constexpr int kSahBuckets = 12;

struct SahTask {
        int lo, hi, me;  // primitives [lo, hi) of the order array; the node's index
        int leaf;        // 1: inside a range the host builder makes one leaf
};

struct B3 {
        float mn[3], mx[3];
};
__device__ __forceinline__ B3 b3_empty() { return B3{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}}; }
// Swift's min/max on Comparable: min(x, y) = y < x ? y : x; max(x, y) = y >= x ? y : x.
__device__ __forceinline__ float sw_min(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float sw_max(float x, float y) { return y >= x ? y : x; }
// Bounds3.swift:104-114 with Bounds3f(first:second:) re-sorting the corners.
__device__ __forceinline__ B3 b3_unite(const B3 &a, const B3 &b) {
        B3 r, s;
        for (int d = 0; d < 3; ++d) {
                r.mn[d] = sw_min(a.mn[d], b.mn[d]);
                r.mx[d] = sw_max(a.mx[d], b.mx[d]);
        }
        for (int d = 0; d < 3; ++d) {
                s.mn[d] = sw_min(r.mn[d], r.mx[d]);
                s.mx[d] = sw_max(r.mn[d], r.mx[d]);
        }
        return s;
}
__device__ __forceinline__ float b3_area(const B3 &b) {  // Bounds3.swift:98-101
        const float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
        float s = dx * dy + dx * dz;
        s = s + dy * dz;
        return 2.0f * s;
}
__device__ __forceinline__ int b3_max_extent(const B3 &b) {  // Bounds3.swift:75-84
        const float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
        if (dx > dy && dx > dz) return 0;
        if (dy > dz) return 1;
        return 2;
}
// The host builder's centre: 0.5 pMin + 0.5 pMax (Bounds3.swift:168-171).
__device__ __forceinline__ float sah_center(const Box &b, int d) {
        const float h0 = 0.5f * b.lo[d];
        const float h1 = 0.5f * b.hi[d];
        return h0 + h1;
}
// bucket_of: Bounds3.offset on one axis, 12 buckets, Int() truncation, clamped to 0..11.
__device__ __forceinline__ int sah_bucket(const B3 &cb, float c, int d) {
        float r = c - cb.mn[d];
        if (cb.mx[d] > cb.mn[d]) r /= cb.mx[d] - cb.mn[d];
        const float f = 12.0f * r;
        if (f >= 12.0f) return kSahBuckets - 1;
        if (!(f > 0.0f)) return 0;
        const int b = (int)f;
        return b > kSahBuckets - 1 ? kSahBuckets - 1 : b;
}
// The SAH costs of the 11 bucket splits and the decision (bvh_build.cpp's build, in its order of
// float operations): true to split at bucket `best`, false for a leaf.
__device__ __forceinline__ bool sah_decide(const int *cnt, const B3 *bb, int count, const B3 &bounds, int &best) {
        float costs[kSahBuckets - 1];
        for (int k = 0; k < kSahBuckets - 1; ++k) costs[k] = 0.0f;
        int below = 0;
        B3 bBelow = b3_empty();
        for (int k = 0; k < kSahBuckets - 1; ++k) {
                bBelow = b3_unite(bBelow, bb[k]);
                below += cnt[k];
                costs[k] = costs[k] + (float)below * b3_area(bBelow);
        }
        int above = 0;
        B3 bAbove = b3_empty();
        for (int k = kSahBuckets - 1; k >= 1; --k) {
                bAbove = b3_unite(bAbove, bb[k]);
                above += cnt[k];
                costs[k - 1] = costs[k - 1] + (float)above * b3_area(bAbove);
        }
        best = -1;
        float minCost = INFINITY;
        for (int k = 0; k < kSahBuckets - 1; ++k)
                if (costs[k] < minCost) {
                        minCost = costs[k];
                        best = k;
                }
        const float leafCost = (float)count;
        minCost = 0.5f + minCost / b3_area(bounds);
        return count > 4 || minCost < leafCost;
}

// Records node `me` over [lo, hi) split before `mid` and its children: a child of one primitive
// is a leaf position, a larger one an internal node (index mid - 1 on the left, mid on the
// right) queued as a task of the small or the next big list.
__device__ __forceinline__ void sah_record(int lo, int mid, int hi, int me, int leaf, int4 *__restrict__ node,
                                           int *__restrict__ leafMark, int *__restrict__ parent,
                                           int *__restrict__ leafParent, int *__restrict__ parents,
                                           int *__restrict__ leafParents, SahTask *__restrict__ big, unsigned *nBig,
                                           SahTask *__restrict__ small, unsigned *nSmall, int cap,
                                           int *__restrict__ flags) {
        const int g = mid - 1;
        const int left = lo == g ? ~g : g;
        const int right = hi - 1 == mid ? ~mid : mid;
        node[me] = make_int4(lo, hi - 1, left, right);
        leafMark[me] = leaf;
        const int cl[2] = {lo, mid}, ch[2] = {mid, hi}, ci[2] = {left, right};
        for (int side = 0; side < 2; ++side) {
                const int c = ci[side];
                if (c < 0) {
                        leafParent[~c] = me;
                        atomicAdd(&leafParents[~c], 1);
                        continue;
                }
                parent[c] = me;
                atomicAdd(&parents[c], 1);
                const SahTask t{cl[side], ch[side], c, leaf};
                if (ch[side] - cl[side] <= kSahSmall) {
                        if (!small) continue;  // the lane's own stack takes it
                        const unsigned k = atomicAdd(nSmall, 1u);
                        if (k < (unsigned)cap)
                                small[k] = t;
                        else
                                atomicOr(&flags[0], kFlagBadTree);
                } else {
                        const unsigned k = atomicAdd(nBig, 1u);
                        if (k < (unsigned)cap)
                                big[k] = t;
                        else
                                atomicOr(&flags[0], kFlagBadTree);
                }
        }
}

// One level of big nodes, one workgroup per node (persistent over the level's list).
__global__ __launch_bounds__(kBuildBlock) void k_sah_level(const Box *__restrict__ primBox, uint32_t *__restrict__ order,
                                                           uint32_t *__restrict__ tmp, const SahTask *__restrict__ tasks,
                                                           const unsigned *nTasks, unsigned *head,
                                                           SahTask *__restrict__ nextBig, unsigned *nNextBig,
                                                           SahTask *__restrict__ small, unsigned *nSmall, int cap,
                                                           int4 *__restrict__ node, int *__restrict__ leafMark,
                                                           int *__restrict__ parent, int *__restrict__ leafParent,
                                                           int *__restrict__ parents, int *__restrict__ leafParents,
                                                           int n, int forceMid, int *__restrict__ flags) {
        constexpr int kWaves = kBuildBlock / 64;
        __shared__ unsigned sTask;
        __shared__ float sRed[kWaves][12];
        __shared__ float sB[12];  // bounds (mn, mx) then centroid bounds (mn, mx)
        __shared__ int sCnt[kSahBuckets];
        __shared__ unsigned sBB[kSahBuckets][6];
        __shared__ int sPlan[4];  // dim, split bucket (-2: middle, -1: buckets next), mid, host leaf
        __shared__ int sScan[kWaves][2];
        const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const unsigned total = *nTasks;
        for (;;) {
                if (tid == 0) sTask = atomicAdd(head, 1u);
                __syncthreads();
                const unsigned t = sTask;
                __syncthreads();
                if (t >= total || t >= (unsigned)cap) break;
                const SahTask task = tasks[t];
                const int lo = task.lo, hi = task.hi;
                if (lo < 0 || hi > n || hi - lo < 2 || task.me < 0 || task.me >= n - 1) {
                        if (tid == 0) atomicOr(&flags[0], kFlagBadTree);
                        continue;
                }
                // 1. bounds and centroid bounds
                float r[12];
                for (int q = 0; q < 3; ++q) {
                        r[q] = INFINITY;
                        r[3 + q] = -INFINITY;
                        r[6 + q] = INFINITY;
                        r[9 + q] = -INFINITY;
                }
                for (int pidx = lo + tid; pidx < hi; pidx += kBuildBlock) {
                        const uint32_t k = order[pidx];
                        if (k >= (uint32_t)n) continue;
                        const Box b = primBox[k];
                        for (int q = 0; q < 3; ++q) {
                                r[q] = fminf(r[q], b.lo[q]);
                                r[3 + q] = fmaxf(r[3 + q], b.hi[q]);
                                const float c = sah_center(b, q);
                                r[6 + q] = fminf(r[6 + q], c);
                                r[9 + q] = fmaxf(r[9 + q], c);
                        }
                }
                for (int q = 0; q < 12; ++q) {
                        const bool isMin = (q % 6) < 3;
                        for (int off = 32; off >= 1; off >>= 1) {
                                const float o = __shfl_xor(r[q], off);
                                r[q] = isMin ? fminf(r[q], o) : fmaxf(r[q], o);
                        }
                }
                if (lane == 0)
                        for (int q = 0; q < 12; ++q) sRed[wave][q] = r[q];
                if (tid < kSahBuckets) {
                        sCnt[tid] = 0;
                        for (int q = 0; q < 3; ++q) {
                                sBB[tid][q] = fkey(INFINITY);
                                sBB[tid][3 + q] = fkey(-INFINITY);
                        }
                }
                __syncthreads();
                if (tid == 0) {
                        for (int q = 0; q < 12; ++q) {
                                const bool isMin = (q % 6) < 3;
                                float v = sRed[0][q];
                                for (int w = 1; w < kWaves; ++w) v = isMin ? fminf(v, sRed[w][q]) : fmaxf(v, sRed[w][q]);
                                sB[q] = v;
                        }
                        B3 bounds, cb;
                        for (int q = 0; q < 3; ++q) {
                                bounds.mn[q] = sB[q];
                                bounds.mx[q] = sB[3 + q];
                                cb.mn[q] = sB[6 + q];
                                cb.mx[q] = sB[9 + q];
                        }
                        const int dim = b3_max_extent(cb);
                        sPlan[0] = dim;
                        // the host makes a leaf of an empty-area node or one with flat centroids
                        const bool hostLeaf = task.leaf || b3_area(bounds) == 0.0f || cb.mx[dim] == cb.mn[dim] ||
                                              !(cb.mx[dim] > cb.mn[dim]);
                        sPlan[1] = (forceMid || hostLeaf) ? -2 : -1;  // -2: split at the middle, -1: buckets next
                        sPlan[3] = hostLeaf ? 1 : 0;
                }
                __syncthreads();
                const int dim = sPlan[0];
                B3 cb;
                for (int q = 0; q < 3; ++q) {
                        cb.mn[q] = sB[6 + q];
                        cb.mx[q] = sB[9 + q];
                }
                const int leafHere = sPlan[3];
                int best = -2;
                if (sPlan[1] == -1) {
                        // 2. buckets
                        for (int pidx = lo + tid; pidx < hi; pidx += kBuildBlock) {
                                const uint32_t k = order[pidx];
                                if (k >= (uint32_t)n) continue;
                                const Box b = primBox[k];
                                const int bk = sah_bucket(cb, sah_center(b, dim), dim);
                                atomicAdd(&sCnt[bk], 1);
                                for (int q = 0; q < 3; ++q) {
                                        atomicMin(&sBB[bk][q], fkey(b.lo[q]));
                                        atomicMax(&sBB[bk][3 + q], fkey(b.hi[q]));
                                }
                        }
                        __syncthreads();
                        if (tid == 0) {
                                int cnt[kSahBuckets];
                                B3 bb[kSahBuckets];
                                for (int k = 0; k < kSahBuckets; ++k) {
                                        cnt[k] = sCnt[k];
                                        for (int q = 0; q < 3; ++q) {
                                                bb[k].mn[q] = fval(sBB[k][q]);
                                                bb[k].mx[q] = fval(sBB[k][3 + q]);
                                        }
                                }
                                B3 bounds;
                                for (int q = 0; q < 3; ++q) {
                                        bounds.mn[q] = sB[q];
                                        bounds.mx[q] = sB[3 + q];
                                }
                                int bst = -1;
                                sah_decide(cnt, bb, hi - lo, bounds, bst);  // more than kSahSmall (> 4): a split
                                int left = 0;
                                for (int k = 0; k <= bst && k < kSahBuckets; ++k) left += cnt[k];
                                sPlan[1] = (bst < 0 || left == 0 || left == hi - lo) ? -2 : bst;
                                sPlan[2] = lo + left;
                        }
                        __syncthreads();
                        best = sPlan[1];
                }
                int mid = lo + (hi - lo) / 2;
                if (best >= 0) {
                        mid = sPlan[2];
                        // 3. stable partition into tmp, then back into order
                        int offL = lo, offR = mid;
                        for (int base = lo; base < hi; base += kBuildBlock) {
                                const int pidx = base + tid;
                                uint32_t k = 0;
                                bool valid = pidx < hi, goLeft = false;
                                if (valid) {
                                        k = order[pidx];
                                        goLeft = k < (uint32_t)n && sah_bucket(cb, sah_center(primBox[k], dim), dim) <= best;
                                }
                                const uint64_t mL = __ballot(valid && goLeft), mR = __ballot(valid && !goLeft);
                                if (lane == 0) {
                                        sScan[wave][0] = __popcll(mL);
                                        sScan[wave][1] = __popcll(mR);
                                }
                                __syncthreads();
                                int preL = 0, preR = 0, totL = 0, totR = 0;
                                for (int w = 0; w < kWaves; ++w) {
                                        if (w < wave) {
                                                preL += sScan[w][0];
                                                preR += sScan[w][1];
                                        }
                                        totL += sScan[w][0];
                                        totR += sScan[w][1];
                                }
                                if (valid) {
                                        const int dst = goLeft ? offL + preL + __popcll(mL & below)
                                                               : offR + preR + __popcll(mR & below);
                                        if (dst >= lo && dst < hi)
                                                tmp[dst] = k;
                                        else
                                                atomicOr(&flags[0], kFlagBadTree);
                                }
                                offL += totL;
                                offR += totR;
                                __syncthreads();
                        }
                        if (tid == 0 && (offL != mid || offR != hi)) atomicOr(&flags[0], kFlagBadTree);
                        __threadfence_block();
                        __syncthreads();
                        for (int pidx = lo + tid; pidx < hi; pidx += kBuildBlock) order[pidx] = tmp[pidx];
                        __syncthreads();
                }
                // 4. the node and its children
                if (tid == 0)
                        sah_record(lo, mid, hi, task.me, leafHere, node, leafMark, parent, leafParent, parents, leafParents,
                                   nextBig, nNextBig, small, nSmall, cap, flags);
                __syncthreads();
        }
}

// Nodes of at most kSahSmall primitives, one lane each, depth first on a private stack; the
// host builder's steps, with its Hoare partition and its two-primitive sort.
__global__ __launch_bounds__(kBuildBlock) void k_sah_small(const Box *__restrict__ primBox, uint32_t *__restrict__ order,
                                                           const SahTask *__restrict__ tasks, const unsigned *nTasks,
                                                           int cap, int4 *__restrict__ node, int *__restrict__ leafMark,
                                                           int *__restrict__ parent,
                                                           int *__restrict__ leafParent, int *__restrict__ parents,
                                                           int *__restrict__ leafParents, int n, int *__restrict__ flags) {
        const unsigned t = blockIdx.x * kBuildBlock + threadIdx.x;
        if (t >= *nTasks || t >= (unsigned)cap) return;
        const SahTask root = tasks[t];
        if (root.lo < 0 || root.hi > n || root.hi - root.lo < 2 || root.hi - root.lo > kSahSmall || root.me < 0 ||
            root.me >= n - 1) {
                atomicOr(&flags[0], kFlagBadTree);
                return;
        }
        int sLo[kSahSmall], sHi[kSahSmall], sMe[kSahSmall], sLeaf[kSahSmall];
        int sp = 0;
        sLo[0] = root.lo;
        sHi[0] = root.hi;
        sMe[0] = root.me;
        sLeaf[0] = root.leaf;
        sp = 1;
        for (int guard = 0; sp > 0; ++guard) {
                if (guard > 2 * kSahSmall) {
                        atomicOr(&flags[0], kFlagBadTree);
                        return;
                }
                --sp;
                const int lo = sLo[sp], hi = sHi[sp], me = sMe[sp];
                int leaf = sLeaf[sp];
                const int count = hi - lo;
                B3 bounds = b3_empty(), cb = b3_empty();
                for (int p = lo; p < hi; ++p) {
                        const Box b = primBox[order[p] < (uint32_t)n ? order[p] : 0];
                        for (int q = 0; q < 3; ++q) {
                                bounds.mn[q] = fminf(bounds.mn[q], b.lo[q]);
                                bounds.mx[q] = fmaxf(bounds.mx[q], b.hi[q]);
                                const float c = sah_center(b, q);
                                cb.mn[q] = fminf(cb.mn[q], c);
                                cb.mx[q] = fmaxf(cb.mx[q], c);
                        }
                }
                const int dim = b3_max_extent(cb);
                int mid = lo + count / 2;
                if (!leaf && (b3_area(bounds) == 0.0f || cb.mx[dim] == cb.mn[dim] || !(cb.mx[dim] > cb.mn[dim])))
                        leaf = 1;  // the host's leaf of an empty-area node or flat centroids
                if (!leaf) {
                        if (count <= 2) {
                                const uint32_t a = order[lo], b = order[lo + 1];
                                if (a < (uint32_t)n && b < (uint32_t)n &&
                                    sah_center(primBox[b], dim) < sah_center(primBox[a], dim)) {
                                        order[lo] = b;
                                        order[lo + 1] = a;
                                }
                                mid = lo + 1;
                        } else {
                                int cnt[kSahBuckets];
                                B3 bb[kSahBuckets];
                                for (int k = 0; k < kSahBuckets; ++k) {
                                        cnt[k] = 0;
                                        bb[k] = b3_empty();
                                }
                                for (int p = lo; p < hi; ++p) {
                                        const Box b = primBox[order[p] < (uint32_t)n ? order[p] : 0];
                                        const int bk = sah_bucket(cb, sah_center(b, dim), dim);
                                        cnt[bk] += 1;
                                        for (int q = 0; q < 3; ++q) {
                                                bb[bk].mn[q] = fminf(bb[bk].mn[q], b.lo[q]);
                                                bb[bk].mx[q] = fmaxf(bb[bk].mx[q], b.hi[q]);
                                        }
                                }
                                int best = -1;
                                if (!sah_decide(cnt, bb, count, bounds, best)) {
                                        leaf = 1;  // the host's leaf of at most four primitives
                                } else if (best >= 0) {
                                        // Hoare partition: bucket > best to the second half
                                        int a = lo, z = hi;
                                        for (int step = 0; step <= count; ++step) {
                                                while (a < z) {
                                                        const uint32_t k = order[a];
                                                        if (sah_bucket(cb, sah_center(primBox[k < (uint32_t)n ? k : 0], dim), dim) > best)
                                                                break;
                                                        ++a;
                                                }
                                                if (a == z) break;
                                                --z;
                                                while (a < z) {
                                                        const uint32_t k = order[z];
                                                        if (!(sah_bucket(cb, sah_center(primBox[k < (uint32_t)n ? k : 0], dim), dim) > best))
                                                                break;
                                                        --z;
                                                }
                                                if (a == z) break;
                                                const uint32_t tswap = order[a];
                                                order[a] = order[z];
                                                order[z] = tswap;
                                                ++a;
                                        }
                                        if (a > lo && a < hi) mid = a;
                                }
                        }
                }
                const int g = mid - 1;
                const int left = lo == g ? ~g : g;
                const int right = hi - 1 == mid ? ~mid : mid;
                node[me] = make_int4(lo, hi - 1, left, right);
                leafMark[me] = leaf;
                const int cl[2] = {lo, mid}, ch[2] = {mid, hi}, ci[2] = {left, right};
                for (int side = 0; side < 2; ++side) {
                        const int c = ci[side];
                        if (c < 0) {
                                leafParent[~c] = me;
                                atomicAdd(&leafParents[~c], 1);
                                continue;
                        }
                        parent[c] = me;
                        atomicAdd(&parents[c], 1);
                        if (sp >= kSahSmall) {
                                atomicOr(&flags[0], kFlagBadTree);
                                return;
                        }
                        sLo[sp] = cl[side];
                        sHi[sp] = ch[side];
                        sMe[sp] = c;
                        sLeaf[sp] = leaf;
                        ++sp;
                }
        }
}

__global__ void k_sah_init(SahTask *big, unsigned *counters, int n) {
        if (threadIdx.x != 0 || blockIdx.x != 0) return;
        big[0] = SahTask{0, n, 0, 0};  // (the root of a small scene goes to the small list)
        counters[0] = 1;  // big tasks of the level
        counters[1] = 0;  // head
        counters[2] = 0;  // next level's big tasks
        counters[3] = 0;  // small tasks
}

__global__ void k_sah_roll(unsigned *counters) {
        if (threadIdx.x != 0 || blockIdx.x != 0) return;
        counters[0] = counters[2];
        counters[1] = 0;
        counters[2] = 0;
}

__global__ void k_sah_identity(uint32_t *order, int n) {
        const int i = (int)(blockIdx.x * kBuildBlock + threadIdx.x);
        if (i < n) order[i] = (uint32_t)i;
}

__global__ void k_root_leaf(TriRec *tris, int n) {
        if (threadIdx.x == 0 && blockIdx.x == 0) tris[0].p0.w = __int_as_float(n);
}

// SAH build: the whole scene is one leaf when the host builder keeps the root whole.
__global__ void k_root_leaf_marked(const int *leafMark, TriRec *tris, int n) {
        if (threadIdx.x == 0 && blockIdx.x == 0 && (n == 1 || leafMark[0] != 0)) tris[0].p0.w = __int_as_float(n);
}

template <class T>
static bool dalloc(T **p, size_t count, const char *what) {
        return hip_ok(hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T)), what);
}

}  // namespace gz

using namespace gz;

extern "C" void *gz_gpu_build_scene(const PrimId_C *prims, int64_t nPrims, const TriangleMesh_C *meshes,
                                    int64_t nMeshes, const int64_t *ptsCounts, const int64_t *faceCounts,
                                    const int64_t *vtxCounts, const Material_C *materials, int64_t nMaterials,
                                    const GzBuildOptions *options) {
        clear_error();
        const int maxLeaf = options ? options->maxLeaf : 4;
        const int sah = options ? options->sah : 0;
        if (nPrims < 0 || nMeshes < 0 || nMaterials < 0 || maxLeaf < 1 || maxLeaf > 64 || (sah != 0 && sah != 1)) {
                set_error("gz_gpu_build_scene: negative count, maxLeaf %d outside 1..64 or sah %d not 0 or 1", maxLeaf,
                          sah);
                return nullptr;
        }
        if (options)
                for (int k = 0; k < 6; ++k)
                        if (options->reserved[k] != 0) {
                                set_error("gz_gpu_build_scene: reserved option fields must be zero");
                                return nullptr;
                        }
        if ((nPrims > 0 && !prims) || (nMeshes > 0 && (!meshes || !ptsCounts || !vtxCounts)) ||
            (nMaterials > 0 && !materials)) {
                set_error("gz_gpu_build_scene: null array with a non-zero count");
                return nullptr;
        }
        if (nPrims >= (1LL << 30)) {
                set_error("gz_gpu_build_scene: %lld primitives (at most 2^30 - 1)", (long long)nPrims);
                return nullptr;
        }
        if (!gz_gpu_available()) {
                set_error("gz_gpu_build_scene: no gfx950 device visible");
                return nullptr;
        }
        Scene *s = new (std::nothrow) Scene();
        if (!s) {
                set_error("gz_gpu_build_scene: out of host memory");
                return nullptr;
        }
        if (!hip_ok(hipGetDevice(&s->device), "hipGetDevice") ||
            !scene_upload_common(s, prims, nPrims, meshes, nMeshes, ptsCounts, faceCounts, vtxCounts, materials,
                                 nMaterials)) {
                scene_release(s);
                return nullptr;
        }
        const int n = (int)nPrims;
        s->primCount = nPrims;
        s->meshCount = nMeshes;
        s->materialCount = nMaterials;
        s->empty = n == 0;
        if (n == 0) return s;
        const hipStream_t st = s->stream;
        float4 *verts = nullptr, *cent = nullptr;
        Box *primBox = nullptr, *nodeBox = nullptr;
        unsigned *cb = nullptr, *arrivals = nullptr, *counts = nullptr;
        int *flags = nullptr, *parent = nullptr, *leafParent = nullptr, *parents = nullptr, *leafParents = nullptr;
        uint64_t *keys = nullptr, *keys2 = nullptr;
        uint32_t *vals = nullptr, *vals2 = nullptr;
        int4 *node = nullptr;
        PrimId_C *sortedPrims = nullptr;
        SahTask *bigA = nullptr, *bigB = nullptr, *smallTasks = nullptr;
        unsigned *sahCounters = nullptr;
        int *leafMark = nullptr, *rec = nullptr;
        const int m = std::max(n - 1, 1);
        const int sortBlocks = (n + kSortBlock - 1) / kSortBlock;
        bool ok = dalloc(&verts, 3 * (size_t)n, "build vertices") && dalloc(&cent, n, "build centroids") &&
                  dalloc(&primBox, n, "build boxes") && dalloc(&nodeBox, m, "build node boxes") &&
                  dalloc(&cb, 6, "build bounds") && dalloc(&arrivals, m, "build counters") &&
                  dalloc(&counts, 256 * (size_t)sortBlocks, "build digit counts") &&
                  dalloc(&flags, 4, "build flags") && dalloc(&parent, m, "build parents") &&
                  dalloc(&leafParent, n, "build parents") && dalloc(&parents, m, "build parent counts") &&
                  dalloc(&leafParents, n, "build parent counts") && dalloc(&keys, n, "build keys") &&
                  dalloc(&keys2, n, "build keys") && dalloc(&vals, n, "build order") && dalloc(&vals2, n, "build order") &&
                  dalloc(&node, m, "build nodes") && dalloc(&rec, m, "build subtree records") &&
                  dalloc(&sortedPrims, n, "build primIds") &&
                  dalloc(&s->dTris, n, "build triangle records") && dalloc(&s->dNodes, m, "build node records") &&
                  (!sah || (dalloc(&bigA, n, "build SAH tasks") && dalloc(&bigB, n, "build SAH tasks") &&
                            dalloc(&smallTasks, n, "build SAH tasks") && dalloc(&sahCounters, 4, "build SAH counters") &&
                            dalloc(&leafMark, m, "build SAH leaves")));
        int h[4] = {0, 0, 0, 0};
        if (ok) {
                const unsigned gb = (unsigned)((n + kBuildBlock - 1) / kBuildBlock);
                hipLaunchKernelGGL(k_build_init, dim3(1), dim3(64), 0, st, cb, flags);
                hipLaunchKernelGGL(k_gather, dim3(gb), dim3(kBuildBlock), 0, st, s->dPrims, n, s->dMeshes, nMeshes, verts,
                                   primBox, cent, cb, flags);
                if (sah)  // the SAH build reorders the identity in place
                        hipLaunchKernelGGL(k_sah_identity, dim3(gb), dim3(kBuildBlock), 0, st, vals, n);
                else
                        hipLaunchKernelGGL(k_morton, dim3(gb), dim3(kBuildBlock), 0, st, cent, n, cb, keys, vals);
                // 8 stable passes of 8 bits; an even number of passes leaves the result in keys, vals
                uint64_t *kin = keys, *kout = keys2;
                uint32_t *vin = vals, *vout = vals2;
                for (int shift = 0; !sah && shift < 64; shift += 8) {
                        hipLaunchKernelGGL(k_digit_counts, dim3((unsigned)sortBlocks), dim3(kSortBlock), 0, st, kin, n,
                                           shift, counts, sortBlocks);
                        hipLaunchKernelGGL(k_digit_scan, dim3(1), dim3(1024), 0, st, counts, 256 * sortBlocks);
                        hipLaunchKernelGGL(k_digit_scatter, dim3((unsigned)sortBlocks), dim3(kSortBlock), 0, st, kin, vin,
                                           n, shift, counts, sortBlocks, kout, vout, flags);
                        std::swap(kin, kout);
                        std::swap(vin, vout);
                }
                ok = hip_ok(hipGetLastError(), "build launch");
        }
        uint32_t *order = vals;  // sorted primitive indices (leaf order)
        if (ok && n > 1) {
                const unsigned gb = (unsigned)((n + kBuildBlock - 1) / kBuildBlock);
                ok = hip_ok(hipMemsetAsync(arrivals, 0, sizeof(unsigned) * (size_t)m, st), "build counters") &&
                     hip_ok(hipMemsetAsync(parents, 0, sizeof(int) * (size_t)m, st), "build counters") &&
                     hip_ok(hipMemsetAsync(leafParents, 0, sizeof(int) * (size_t)n, st), "build counters");
                if (ok && sah) {
                        // top-down binned SAH: big nodes level by level, then the small ones
                        hipLaunchKernelGGL(k_sah_init, dim3(1), dim3(64), 0, st, n > kSahSmall ? bigA : smallTasks,
                                           sahCounters, n);
                        unsigned hc[4] = {n > kSahSmall ? 1u : 0u, 0u, 0u, n > kSahSmall ? 0u : 1u};
                        if (n <= kSahSmall)
                                ok = hip_ok(hipMemcpyAsync(sahCounters, hc, sizeof hc, hipMemcpyHostToDevice, st),
                                            "build SAH counters");
                        int perCU = 0;
                        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_sah_level, kBuildBlock, 0) !=
                                    hipSuccess ||
                            perCU < 1)
                                perCU = 1;
                        // a level halves no node in the worst case but always splits it: the level count is
                        // bounded by n; past 4096 levels every node is split at its middle
                        for (int level = 0; ok && hc[0] > 0 && level < n; ++level) {
                                const unsigned grid = (unsigned)std::min<int64_t>((int64_t)hc[0],
                                                                                  (int64_t)s->cuCount * perCU);
                                hipLaunchKernelGGL(k_sah_level, dim3(grid), dim3(kBuildBlock), 0, st, primBox, order, vals2,
                                                   bigA, sahCounters, sahCounters + 1, bigB, sahCounters + 2, smallTasks,
                                                   sahCounters + 3, n, node, leafMark, parent, leafParent, parents,
                                                   leafParents, n, level >= 4096 ? 1 : 0, flags);
                                hipLaunchKernelGGL(k_sah_roll, dim3(1), dim3(64), 0, st, sahCounters);
                                ok = hip_ok(hipGetLastError(), "build SAH level") &&
                                     hip_ok(hipMemcpyAsync(hc, sahCounters, sizeof hc, hipMemcpyDeviceToHost, st),
                                            "build SAH counters") &&
                                     hip_ok(hipStreamSynchronize(st), "build SAH level");
                                std::swap(bigA, bigB);
                        }
                        if (ok && hc[0] > 0) {
                                set_error("gz_gpu_build_scene: the SAH build did not finish");
                                ok = false;
                        }
                        if (ok && hc[3] > 0)
                                hipLaunchKernelGGL(k_sah_small, dim3((hc[3] + kBuildBlock - 1) / kBuildBlock),
                                                   dim3(kBuildBlock), 0, st, primBox, order, smallTasks, sahCounters + 3, n,
                                                   node, leafMark, parent, leafParent, parents, leafParents, n, flags);
                } else if (ok) {
                        const unsigned gi = (unsigned)((n - 1 + kBuildBlock - 1) / kBuildBlock);
                        hipLaunchKernelGGL(k_hierarchy, dim3(gi), dim3(kBuildBlock), 0, st, keys, n, node, parent,
                                           leafParent, parents, leafParents);
                }
        }
        if (ok) {
                const unsigned gb = (unsigned)((n + kBuildBlock - 1) / kBuildBlock);
                hipLaunchKernelGGL(k_sorted, dim3(gb), dim3(kBuildBlock), 0, st, order, verts, s->dPrims, n, s->dTris,
                                   sortedPrims, flags);
                if (n > 1) {
                        const unsigned gi = (unsigned)((n - 1 + kBuildBlock - 1) / kBuildBlock);
                        hipLaunchKernelGGL(k_validate, dim3(gb), dim3(kBuildBlock), 0, st, node, n, parents, leafParents,
                                           flags);
                        hipLaunchKernelGGL(k_refit, dim3(gb), dim3(kBuildBlock), 0, st, node, parent, leafParent, order,
                                           primBox, nodeBox, arrivals, rec, maxLeaf, sah ? leafMark : nullptr, n, flags);
                        if (sah || n > maxLeaf)
                                hipLaunchKernelGGL(k_emit, dim3(gi), dim3(kBuildBlock), 0, st, node, parent, nodeBox,
                                                   primBox, order, n, maxLeaf, sah ? leafMark : nullptr, rec, s->dNodes,
                                                   s->dTris, flags);
                }
                if (sah)
                        hipLaunchKernelGGL(k_root_leaf_marked, dim3(1), dim3(64), 0, st, leafMark, s->dTris, n);
                else if (n <= maxLeaf)
                        hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(64), 0, st, s->dTris, n);
                ok = ok && hip_ok(hipGetLastError(), "build launch") &&
                     hip_ok(hipMemcpyAsync(h, flags, sizeof h, hipMemcpyDeviceToHost, st), "build flags") &&
                     hip_ok(hipStreamSynchronize(st), "build");
        }
        if (ok && (h[0] & kFlagBadTree)) {
                set_error("gz_gpu_build_scene: the device build produced an inconsistent tree (rejected)");
                ok = false;
        }
        if (ok && (h[0] & kFlagBadIndex)) {
                set_error("gz_gpu_build_scene: a primId references a mesh, index or vertex outside the mesh arrays");
                ok = false;
        }
        if (ok) {
                // primIds in leaf order replace the caller's order on the device
                (void)hipFree(s->dPrims);
                s->dPrims = sortedPrims;
                sortedPrims = nullptr;
                s->dBuildOrder = order;
                vals = nullptr;
                // a root leaf: no pair record was emitted
                const bool rootLeaf = sah ? h[1] == 0 : n <= maxLeaf;
                s->nodeRecords = rootLeaf ? 0 : h[1];  // records numbered depth first, densely
                s->triCount = n;
                s->boundsSafe = !(h[0] & (kFlagUnsafe | kFlagSkip));  // finite boxes within 2^40, none empty
                if (rootLeaf) {
                        s->rootDesc = ~0;
                        s->interiorCount = 0;
                        s->maxDepth = 0;
                        s->nodeCount = 1;
                } else {
                        s->rootDesc = 0;
                        s->interiorCount = h[1];
                        s->maxDepth = h[2];
                        s->nodeCount = (int64_t)h[1] + h[3];
                }
                s->deviceBytes += (int64_t)(sizeof(TriRec) * (size_t)n + sizeof(NodePair) * (size_t)m);
        }
        void *tmp[] = {verts,       cent,    primBox, nodeBox, cb,   arrivals, counts,      flags,
                       parent,      leafParent, parents, leafParents, keys, keys2, vals, vals2,
                       node,        sortedPrims, bigA, bigB, smallTasks, sahCounters, leafMark, rec};
        for (void *p : tmp)
                if (p) (void)hipFree(p);
        if (!ok) {
                scene_release(s);
                return nullptr;
        }
        return s;
}

// The handle's tree in the reference's BVH2 layout (BVH2Node.swift: depth-first, left child at
// index + 1, right child at `offset`, leaves over [offset, offset + count) of the primIds in leaf
// order). order[j] is the caller's primitive index of leaf-order position j (for a handle from
// gz_gpu_upload_scene, the identity). Returns the node count, or a negative GZ_E_* code
// (GZ_E_INVALID when nodeCapacity is too small: the count needed is 2 x leaves - 1 <= 2 x primCount).
extern "C" int64_t gz_gpu_export_bvh2(void *handle, BVH2Node *nodes, int64_t nodeCapacity, int32_t *order) {
        clear_error();
        if (!handle || nodeCapacity < 0 || (nodeCapacity > 0 && !nodes)) {
                set_error("gz_gpu_export_bvh2: bad arguments");
                return GZ_E_INVALID;
        }
        Scene *s = static_cast<Scene *>(handle);
        std::lock_guard<std::mutex> lock(s->mu);
        if (!enter_device(s)) return GZ_E_HIP;
        if (s->empty) return 0;
        std::vector<NodePair> pairs;
        std::vector<TriRec> tris;
        std::vector<uint32_t> built;
        try {
                // a host upload numbers its records densely; a device build leaves unused slots
                pairs.resize((size_t)s->nodeRecords);
                tris.resize((size_t)s->triCount);
                if (s->dBuildOrder) built.resize((size_t)s->triCount);
        } catch (const std::bad_alloc &) {
                set_error("gz_gpu_export_bvh2: out of host memory");
                return GZ_E_ALLOC;
        }
        if (!hip_ok(hipStreamSynchronize(s->stream), "export sync") ||
            (!pairs.empty() &&
             !hip_ok(hipMemcpy(pairs.data(), s->dNodes, pairs.size() * sizeof(NodePair), hipMemcpyDeviceToHost),
                     "export nodes")) ||
            !hip_ok(hipMemcpy(tris.data(), s->dTris, tris.size() * sizeof(TriRec), hipMemcpyDeviceToHost), "export tris") ||
            (!built.empty() &&
             !hip_ok(hipMemcpy(built.data(), s->dBuildOrder, built.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
                     "export order")))
                return GZ_E_HIP;
        if (order)
                for (size_t j = 0; j < tris.size(); ++j) {
                        int32_t k;
                        std::memcpy(&k, &tris[j].p1.w, 4);
                        order[j] = built.empty() ? k : (int32_t)built[(size_t)k];
                }
        auto leaf_count = [&](int32_t first) {
                int32_t c;
                std::memcpy(&c, &tris[(size_t)first].p0.w, 4);
                return c;
        };
        struct Item {
                int32_t desc;
                float box[6];
                int64_t parentSlot;  // node whose offset becomes this node's index (-1: none)
        };
        std::vector<Item> stack;
        int64_t count = 0;
        // the root's box is the union of its children's (the traversal never tests it)
        float rootBox[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        if (s->rootDesc >= 0 && (int64_t)s->rootDesc < (int64_t)pairs.size()) {
                const NodePair &r = pairs[(size_t)s->rootDesc];
                const float L[6] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y};
                const float R[6] = {r.b.z, r.b.w, r.c.x, r.c.y, r.c.z, r.c.w};
                for (int q = 0; q < 3; ++q) {
                        rootBox[q] = std::min(L[q], R[q]);
                        rootBox[3 + q] = std::max(L[3 + q], R[3 + q]);
                }
        } else if (s->rootDesc < 0 && (int64_t)~s->rootDesc < (int64_t)tris.size()) {
                // a root leaf: the box of its triangles (skipped primitives have none)
                const int32_t first = ~s->rootDesc, count = leaf_count(first);
                for (int32_t j = first; j < first + count && j < (int32_t)tris.size(); ++j) {
                        int32_t skip;
                        std::memcpy(&skip, &tris[(size_t)j].p2.w, 4);
                        if (skip & kTriSkip) continue;
                        const float4 v[3] = {tris[(size_t)j].p0, tris[(size_t)j].p1, tris[(size_t)j].p2};
                        for (const float4 &p : v) {
                                const float c[3] = {p.x, p.y, p.z};
                                for (int q = 0; q < 3; ++q) {
                                        rootBox[q] = std::min(rootBox[q], c[q]);
                                        rootBox[3 + q] = std::max(rootBox[3 + q], c[q]);
                                }
                        }
                }
        }
        stack.push_back(Item{s->rootDesc, {rootBox[0], rootBox[1], rootBox[2], rootBox[3], rootBox[4], rootBox[5]}, -1});
        while (!stack.empty()) {
                const Item it = stack.back();
                stack.pop_back();
                if (count >= nodeCapacity) {
                        set_error("gz_gpu_export_bvh2: more than %lld nodes", (long long)nodeCapacity);
                        return GZ_E_INVALID;
                }
                const int64_t me = count++;
                if (it.parentSlot >= 0) nodes[it.parentSlot].offset = (int32_t)me;
                BVH2Node &b = nodes[me];
                b.boundsMinX = it.box[0];
                b.boundsMinY = it.box[1];
                b.boundsMinZ = it.box[2];
                b.boundsMaxX = it.box[3];
                b.boundsMaxY = it.box[4];
                b.boundsMaxZ = it.box[5];
                if (it.desc < 0) {
                        if ((int64_t)~it.desc >= (int64_t)tris.size()) {
                                set_error("gz_gpu_export_bvh2: leaf record %d outside %zu", ~it.desc, tris.size());
                                return GZ_E_INVALID;
                        }
                        b.offset = ~it.desc;
                        b.count = leaf_count(~it.desc);
                        continue;
                }
                if ((int64_t)it.desc >= (int64_t)pairs.size()) {
                        set_error("gz_gpu_export_bvh2: node record %d outside %zu", it.desc, pairs.size());
                        return GZ_E_INVALID;
                }
                b.count = 0;
                b.offset = -1;
                const NodePair &p = pairs[(size_t)it.desc];
                // right first on the stack so the left child is visited next (index me + 1)
                stack.push_back(Item{p.d.y, {p.b.z, p.b.w, p.c.x, p.c.y, p.c.z, p.c.w}, me});
                stack.push_back(Item{p.d.x, {p.a.x, p.a.y, p.a.z, p.a.w, p.b.x, p.b.y}, -1});
        }
        return count;
}
