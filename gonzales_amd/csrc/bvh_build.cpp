// Host-side SAH BVH2 builder: the Swift host's BoundingHierarchyBuilder, restated in C++ so the
// library can be driven without the Swift toolchain (the reference keeps this step on the host:
// Api/SceneDescription.swift:4-36 builds, then uploads).
//
// Follows Sources/libgonzales/Accelerators/BoundingHierarchyBuilder.swift:
//   per-primitive bounds + centre 0.5*pMin + 0.5*pMax ...... :66-94, Bounds3.swift:168-171
//   build(range): bounds, centroid bounds, longest axis ..... :266-281, Bounds3.swift:75-84
//   leaf if area 0, one primitive or flat centroids ........ :283-288
//   splitSurfaceAreaHeuristic (12 buckets, leaf <= 4) ...... :199-264
//   partition(by: bucket > best) (Hoare partition of the
//   Swift standard library's MutableCollection) ........... :250-256
//   two-primitive ranges sorted by centre .................. :210-212
//   flatten2: depth-first, left = i + 1, right = offset .... :122-156
// Float arithmetic is single precision with no contraction (compiled with -ffp-contract=off),
// as the Swift source evaluates it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "gonzales_amd.h"

namespace gz {
void set_error(const char *fmt, ...);
void clear_error();
void alloc_fault_arm();
bool alloc_fault(const char *what);
}  // namespace gz

namespace {

struct B3 {
        float mn[3];
        float mx[3];
};

inline B3 empty_bounds() {
        const float inf = std::numeric_limits<float>::infinity();
        return B3{{inf, inf, inf}, {-inf, -inf, -inf}};
}

// Swift's min/max on Comparable: min(x, y) = y < x ? y : x; max(x, y) = y >= x ? y : x.
inline float smin(float x, float y) { return y < x ? y : x; }
inline float smax(float x, float y) { return y >= x ? y : x; }

inline B3 unite(const B3 &a, const B3 &b) {  // Bounds3.swift:104-114
        B3 r;
        for (int d = 0; d < 3; ++d) {
                r.mn[d] = smin(a.mn[d], b.mn[d]);
                r.mx[d] = smax(a.mx[d], b.mx[d]);
        }
        // Bounds3f(first:second:) re-sorts the corners (Bounds3.swift:3-12).
        B3 s;
        for (int d = 0; d < 3; ++d) {
                s.mn[d] = smin(r.mn[d], r.mx[d]);
                s.mx[d] = smax(r.mn[d], r.mx[d]);
        }
        return s;
}

inline B3 unite_point(const B3 &a, const float *p) {  // Geometry.swift:51-61
        B3 r;
        for (int d = 0; d < 3; ++d) {
                r.mn[d] = smin(a.mn[d], p[d]);
                r.mx[d] = smax(a.mx[d], p[d]);
        }
        B3 s;
        for (int d = 0; d < 3; ++d) {
                s.mn[d] = smin(r.mn[d], r.mx[d]);
                s.mx[d] = smax(r.mn[d], r.mx[d]);
        }
        return s;
}

inline float surface_area(const B3 &b) {  // Bounds3.swift:98-101
        float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
        float s = dx * dy + dx * dz;
        s = s + dy * dz;
        return 2.0f * s;
}

inline int maximum_extent(const B3 &b) {  // Bounds3.swift:75-84
        float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
        if (dx > dy && dx > dz) return 0;
        if (dy > dz) return 1;
        return 2;
}

struct Prim {
        int64_t index;
        B3 bound;
        float center[3];
};

struct Builder {
        std::vector<Prim> prims;
        BVH2Node *nodes = nullptr;
        int64_t nodeCount = 0;
        int64_t capacity = 0;
        bool failed = false;

        // Bounds3.offset (Bounds3.swift:90-96), one axis.
        static float offset_axis(const B3 &cb, const float *p, int d) {
                float r = p[d] - cb.mn[d];
                if (cb.mx[d] > cb.mn[d]) r /= cb.mx[d] - cb.mn[d];
                return r;
        }

        static int bucket_of(const B3 &cb, const Prim &p, int dim) {
                float off = offset_axis(cb, p.center, dim);
                float f = 12.0f * off;
                // Swift Int(Float) truncates toward zero.
                int b = (int)f;
                if (b == 12) b = 11;
                if (b < 0) b = 0;
                if (b > 11) b = 11;
                return b;
        }

        // Swift MutableCollection.partition(by:) for a bidirectional collection: Hoare scheme,
        // elements satisfying the predicate end up in the second half; returns the pivot.
        template <class Pred>
        int64_t partition(int64_t lo, int64_t hi, Pred second) {
                for (;;) {
                        for (;;) {
                                if (lo == hi) return lo;
                                if (second(prims[lo])) break;
                                ++lo;
                        }
                        for (;;) {
                                --hi;
                                if (lo == hi) return lo;
                                if (!second(prims[hi])) break;
                        }
                        std::swap(prims[lo], prims[hi]);
                        ++lo;
                }
        }

        int64_t emit(const B3 &b) {
                if (nodeCount >= capacity) {
                        failed = true;
                        return -1;
                }
                BVH2Node &n = nodes[nodeCount];
                n.boundsMinX = b.mn[0];
                n.boundsMinY = b.mn[1];
                n.boundsMinZ = b.mn[2];
                n.boundsMaxX = b.mx[0];
                n.boundsMaxY = b.mx[1];
                n.boundsMaxZ = b.mx[2];
                n.offset = 0;
                n.count = 0;
                return nodeCount++;
        }

        // Builds [lo, hi) and emits it in flatten2's depth-first order (the node index is
        // assigned before its children, left subtree before right). Returns the node's bounds.
        B3 build(int64_t lo, int64_t hi) {
                B3 bounds = empty_bounds();
                for (int64_t i = lo; i < hi; ++i) bounds = unite(bounds, prims[i].bound);
                B3 cb = empty_bounds();
                for (int64_t i = lo; i < hi; ++i) cb = unite_point(cb, prims[i].center);
                const int dim = maximum_extent(cb);
                const int64_t count = hi - lo;
                if (surface_area(bounds) == 0.0f || count == 1 || cb.mx[dim] == cb.mn[dim]) {
                        make_leaf(bounds, lo, count);
                        return bounds;
                }
                int64_t mid;
                if (count <= 2) {
                        // sort(by: centre < centre) of two elements (stable insertion step).
                        if (prims[lo + 1].center[dim] < prims[lo].center[dim]) std::swap(prims[lo], prims[lo + 1]);
                        mid = lo + count / 2;
                } else {
                        int cnt[12] = {0};
                        B3 bb[12];
                        for (int k = 0; k < 12; ++k) bb[k] = empty_bounds();
                        for (int64_t i = lo; i < hi; ++i) {
                                int b = bucket_of(cb, prims[i], dim);
                                cnt[b] += 1;
                                bb[b] = unite(bb[b], prims[i].bound);
                        }
                        float costs[11] = {0};
                        int below = 0;
                        B3 bBelow = empty_bounds();
                        for (int k = 0; k < 11; ++k) {
                                bBelow = unite(bBelow, bb[k]);
                                below += cnt[k];
                                costs[k] = costs[k] + (float)below * surface_area(bBelow);
                        }
                        int above = 0;
                        B3 bAbove = empty_bounds();
                        for (int k = 11; k >= 1; --k) {
                                bAbove = unite(bAbove, bb[k]);
                                above += cnt[k];
                                costs[k - 1] = costs[k - 1] + (float)above * surface_area(bAbove);
                        }
                        int best = -1;
                        float minCost = std::numeric_limits<float>::infinity();
                        for (int k = 0; k < 11; ++k)
                                if (costs[k] < minCost) {
                                        minCost = costs[k];
                                        best = k;
                                }
                        const float leafCost = (float)count;
                        minCost = 0.5f + minCost / surface_area(bounds);
                        if (count > 4 || minCost < leafCost) {
                                mid = partition(lo, hi, [&](const Prim &p) { return bucket_of(cb, p, dim) > best; });
                        } else {
                                make_leaf(bounds, lo, count);
                                return bounds;
                        }
                }
                if (mid == lo || mid == hi) {
                        // Unreachable for finite bounds (bucket 0 and 11 are both occupied);
                        // a leaf keeps the build total for non-finite input.
                        make_leaf(bounds, lo, count);
                        return bounds;
                }
                int64_t me = emit(bounds);
                if (me < 0) return bounds;
                B3 l = build(lo, mid);
                int64_t right = nodeCount;
                B3 r = build(mid, hi);
                if (failed) return bounds;
                B3 u = unite(l, r);  // BVHBuildNode(axis:left:right:) bounds
                BVH2Node &n = nodes[me];
                n.boundsMinX = u.mn[0];
                n.boundsMinY = u.mn[1];
                n.boundsMinZ = u.mn[2];
                n.boundsMaxX = u.mx[0];
                n.boundsMaxY = u.mx[1];
                n.boundsMaxZ = u.mx[2];
                n.offset = (int32_t)right;
                n.count = 0;
                return u;
        }

        void make_leaf(const B3 &b, int64_t lo, int64_t count) {
                int64_t me = emit(b);
                if (me < 0) return;
                nodes[me].offset = (int32_t)lo;
                nodes[me].count = (int32_t)count;
        }
};

}  // namespace

extern "C" int64_t gz_bvh2_build(const float *primBounds, int64_t n, BVH2Node *nodesOut, int64_t *primOrderOut) {
        gz::clear_error();
        gz::alloc_fault_arm();
        if (n < 0 || (n > 0 && (!primBounds || !nodesOut || !primOrderOut))) {
                gz::set_error("gz_bvh2_build: bad arguments");
                return GZ_E_INVALID;
        }
        if (n == 0) return 0;
        if (n > (int64_t)INT32_MAX / 2) {  // node offsets are int32
                gz::set_error("gz_bvh2_build: %lld primitives exceed the int32 node offsets", (long long)n);
                return GZ_E_INVALID;
        }
        try {
                Builder b;
                if (gz::alloc_fault("SAH build primitives")) return GZ_E_ALLOC;
                b.prims.resize((size_t)n);
                for (int64_t i = 0; i < n; ++i) {
                        Prim &p = b.prims[(size_t)i];
                        p.index = i;
                        const float *s = primBounds + 6 * i;
                        for (int d = 0; d < 3; ++d) {
                                p.bound.mn[d] = s[d];
                                p.bound.mx[d] = s[3 + d];
                        }
                        for (int d = 0; d < 3; ++d) {
                                float h0 = 0.5f * p.bound.mn[d];
                                float h1 = 0.5f * p.bound.mx[d];
                                p.center[d] = h0 + h1;
                        }
                }
                b.nodes = nodesOut;
                b.capacity = 2 * n - 1;
                b.build(0, n);
                if (b.failed) {
                        gz::set_error("gz_bvh2_build: more nodes than 2n - 1 (non-finite bounds?)");
                        return GZ_E_INVALID;
                }
                for (int64_t k = 0; k < n; ++k) primOrderOut[k] = b.prims[(size_t)k].index;
                return b.nodeCount;
        } catch (const std::bad_alloc &) {
                gz::set_error("gz_bvh2_build: out of host memory");
                return GZ_E_ALLOC;
        }
}
