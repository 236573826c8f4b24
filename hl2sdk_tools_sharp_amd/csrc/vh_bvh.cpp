// Scene build for libvradhip: the reference's intersection-format triangles
// and a binned-SAH BVH2 laid out for the gfx950 traversal kernels.
//
// The BVH replaces the reference's KD tree (Tracing/KdTreeBuilder.cs) for
// traversal.  A BVH changes which triangles a ray is tested against, never
// the test itself, so every answer is the exact test's answer over the
// triangles the ray can reach; the boxes are widened so that no hit the test
// accepts can lie outside its leaf's box.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <system_error>
#include <thread>

#include "vh_internal.h"

namespace vh {
namespace {

// KdTreeBuilder.GetEdgeEquation (Tracing/KdTreeBuilder.cs:1531-1555): a 2-D
// edge line through p1, p2 in coordinates (c1, c2), sign-fixed so the third
// vertex is on the positive side and scaled to read 1 there.
void edge_equation(const float *p1, const float *p2, int c1, int c2, const float *inside,
                   float *out)
{
    float nx = p1[c2] - p2[c2];
    float ny = p2[c1] - p1[c1];
    float d = -((nx * p1[c1]) + (ny * p1[c2]));
    float trial = (inside[c1] * nx) + (inside[c2] * ny) + d;
    if (trial < 0) {
        nx = -nx;
        ny = -ny;
        d = -d;
        trial = -trial;
    }
    out[0] = nx / trial;
    out[1] = ny / trial;
    out[2] = d / trial;
}

struct Tri48 {
    float n[3], d;
    int32_t id;
    float e[6];
    int cs0, cs1;
};

// KdTreeBuilder.ToIntersectionFormat, Correct side (Tracing/KdTreeBuilder.cs:1568-1612):
// normal = exact normalise of (p2-p1) x (p3-p1) (MapFormats/Geometry/Vec3.cs:138-190),
// drop axis = largest |n| component (ties keep the lower), d = n . p1.
// This file is compiled with -ffp-contract=off, so each operation rounds once.
Tri48 convert(const float *v, int32_t id)
{
    const float *p1 = v, *p2 = v + 3, *p3 = v + 6;
    float e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    float e2[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    float c[3] = {(e1[1] * e2[2]) - (e1[2] * e2[1]), (e1[2] * e2[0]) - (e1[0] * e2[2]),
                  (e1[0] * e2[1]) - (e1[1] * e2[0])};
    float len = std::sqrt((c[0] * c[0]) + (c[1] * c[1]) + (c[2] * c[2]));
    Tri48 t{};
    if (len == 0.0f) {
        t.n[0] = t.n[1] = t.n[2] = 0.0f;
    } else {
        t.n[0] = c[0] / len;
        t.n[1] = c[1] / len;
        t.n[2] = c[2] / len;
    }
    int drop = 0;
    for (int k = 1; k < 3; k++)
        if (std::fabs(t.n[k]) > std::fabs(t.n[drop])) drop = k;
    t.cs0 = (drop + 1) % 3;
    t.cs1 = (drop + 2) % 3;
    edge_equation(p1, p2, t.cs0, t.cs1, p3, t.e);
    edge_equation(p2, p3, t.cs0, t.cs1, p1, t.e + 3);
    t.d = (t.n[0] * p1[0]) + (t.n[1] * p1[1]) + (t.n[2] * p1[2]);
    t.id = id;
    return t;
}

struct Box {
    float lo[3], hi[3];
    void reset()
    {
        for (int k = 0; k < 3; k++) {
            lo[k] = INFINITY;
            hi[k] = -INFINITY;
        }
    }
    void grow(const Box &b)
    {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    float area() const
    {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }
};

// A triangle as the build sees it: bounds, centroid, index.  The build
// partitions these in place, so each node's scans read memory in order.
struct Prim {
    Box box;
    float c[3];
    int32_t id;
};

struct BNode {
    Box box;
    int32_t left = -1, right = -1;  // build-node indices, -1 for a leaf
    int32_t first = 0, count = 0;   // leaf range in the order array
};

// Host threads for the build: at most 16, and what the host has.
unsigned build_threads()
{
    return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

// f(begin, end) over [0, n) in contiguous chunks, one per thread; serially
// when no thread can be had.  Chunks write disjoint outputs.
template <class F>
void parallel_chunks(int64_t n, F &&f)
{
    const int64_t nt = std::min<int64_t>(build_threads(), std::max<int64_t>(1, n / 4096));
    std::vector<std::thread> pool;
    int64_t done = 0;  // chunks [0, done) run on threads
    try {
        for (int64_t t = 1; t < nt; t++) {
            pool.emplace_back(f, n * t / nt, n * (t + 1) / nt);
            done = t;
        }
    } catch (const std::system_error &) {
    }
    f(0, n / nt);
    for (int64_t t = done + 1; t < nt; t++) f(n * t / nt, n * (t + 1) / nt);
    for (auto &th : pool) th.join();
}

// Binned SAH, top-down.  The upper levels fork: a node's right subtree is
// built on another thread into its own node list and appended after the
// left subtree, so the nodes come out in the serial build's preorder and the
// tree is the serial build's whatever the timing.  Subtrees own disjoint
// ranges of prims, which is all they write.
struct Builder {
    std::vector<Prim> &prims;
    std::vector<BNode> nodes;
    int max_depth = 0;

    static constexpr int kBins = 64;
    static constexpr float kTraverse = 1.0f;
    static constexpr float kIntersect = 1.5f;
    static constexpr int kLeafMax = 8;
    static constexpr int32_t kForkMin = 8192;  // smallest subtree worth a thread
    static constexpr int32_t kSplitPar = 65536; // smallest node whose scans are shared out

    explicit Builder(std::vector<Prim> &p) : prims(p) {}

    // One scan's partial result over a range of primitives: the bounds of
    // the boxes and of the centroids, or the three axes' bins.  Merging is
    // min/max and integer sums, so any split of the range gives the same
    // result as one serial scan.
    struct Scan {
        Box b, cb;
        int cnt[3][kBins];
        Box bb[3][kBins];
    };

    template <class F>
    static void shared_scan(int32_t first, int32_t count, std::vector<Scan> &part, F &&f)
    {
        const int32_t nt = count >= kSplitPar ? (int32_t)build_threads() : 1;
        part.assign((size_t)nt, Scan{});
        std::vector<std::thread> pool;
        int32_t started = 1;
        try {
            for (int32_t t = 1; t < nt; t++) {
                pool.emplace_back([&, t]() {
                    f(part[t], first + (int32_t)((int64_t)count * t / nt),
                      first + (int32_t)((int64_t)count * (t + 1) / nt));
                });
                started = t + 1;
            }
        } catch (const std::system_error &) {
        }
        f(part[0], first, first + (int32_t)((int64_t)count / nt));
        for (int32_t t = started; t < nt; t++)
            f(part[t], first + (int32_t)((int64_t)count * t / nt),
              first + (int32_t)((int64_t)count * (t + 1) / nt));
        for (auto &th : pool) th.join();
    }

    // fork > 0: this many more levels may hand their right subtree to a thread
    int32_t build(int32_t first, int32_t count, int depth, int fork)
    {
        int32_t me = (int32_t)nodes.size();
        nodes.emplace_back();
        Box b, cb;
        b.reset();
        cb.reset();
        auto bounds = [&](Box &bo, Box &cbo, int32_t lo, int32_t hi) {
            for (int32_t i = lo; i < hi; i++) {
                bo.grow(prims[i].box);
                Box c;
                for (int k = 0; k < 3; k++) c.lo[k] = c.hi[k] = prims[i].c[k];
                cbo.grow(c);
            }
        };
        const bool par = count >= kSplitPar;
        std::vector<Scan> part;
        if (par) {
            shared_scan(first, count, part, [&](Scan &sc, int32_t lo, int32_t hi) {
                sc.b.reset();
                sc.cb.reset();
                bounds(sc.b, sc.cb, lo, hi);
            });
            for (const Scan &sc : part) {
                b.grow(sc.b);
                cb.grow(sc.cb);
            }
        } else {
            bounds(b, cb, first, first + count);
        }
        nodes[me].box = b;
        max_depth = std::max(max_depth, depth);
        float leaf_cost = kIntersect * count;
        int best_axis = -1, best_bin = 0;
        float best_cost = INFINITY;
        if (count > 1 && depth < kMaxDepth) {
            // one pass bins the centroids along all three axes
            float ext[3];
            bool live[3];
            int cnt[3][kBins] = {};
            Box bb[3][kBins];
            for (int ax = 0; ax < 3; ax++) {
                ext[ax] = cb.hi[ax] - cb.lo[ax];
                live[ax] = ext[ax] > 0;
                for (auto &x : bb[ax]) x.reset();
            }
            auto bin_range = [&](int (&c)[3][kBins], Box (&bx)[3][kBins], int32_t lo, int32_t hi) {
                for (int32_t i = lo; i < hi; i++) {
                    const Prim &t = prims[i];
                    for (int ax = 0; ax < 3; ax++) {
                        if (!live[ax]) continue;
                        int bin = (int)((t.c[ax] - cb.lo[ax]) / ext[ax] * kBins);
                        bin = std::min(std::max(bin, 0), kBins - 1);
                        c[ax][bin]++;
                        bx[ax][bin].grow(t.box);
                    }
                }
            };
            if (par) {
                shared_scan(first, count, part, [&](Scan &sc, int32_t lo, int32_t hi) {
                    for (int ax = 0; ax < 3; ax++)
                        for (int bin = 0; bin < kBins; bin++) {
                            sc.cnt[ax][bin] = 0;
                            sc.bb[ax][bin].reset();
                        }
                    bin_range(sc.cnt, sc.bb, lo, hi);
                });
                for (const Scan &sc : part)
                    for (int ax = 0; ax < 3; ax++)
                        for (int bin = 0; bin < kBins; bin++) {
                            cnt[ax][bin] += sc.cnt[ax][bin];
                            bb[ax][bin].grow(sc.bb[ax][bin]);
                        }
            } else {
                bin_range(cnt, bb, first, first + count);
            }
            for (int ax = 0; ax < 3; ax++) {
                if (!live[ax]) continue;
                float rarea[kBins];
                int rcnt[kBins];
                Box acc;
                acc.reset();
                int n = 0;
                for (int bin = kBins - 1; bin > 0; bin--) {
                    acc.grow(bb[ax][bin]);
                    n += cnt[ax][bin];
                    rarea[bin] = acc.area();
                    rcnt[bin] = n;
                }
                acc.reset();
                n = 0;
                float parea = b.area();
                for (int bin = 0; bin < kBins - 1; bin++) {
                    acc.grow(bb[ax][bin]);
                    n += cnt[ax][bin];
                    if (n == 0 || rcnt[bin + 1] == 0) continue;
                    float cost = kTraverse + kIntersect *
                                                 (acc.area() * n + rarea[bin + 1] * rcnt[bin + 1]) /
                                                 (parea > 0 ? parea : 1.0f);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = ax;
                        best_bin = bin;
                    }
                }
            }
        }
        bool make_leaf = count <= 1 || depth >= kMaxDepth ||
                         (count <= kLeafMax && !(best_cost < leaf_cost));
        if (make_leaf) {
            nodes[me].first = first;
            nodes[me].count = count;
            return me;
        }
        int32_t mid;
        if (best_axis < 0) {
            mid = first + count / 2;  // coincident centroids: split the list
        } else {
            float ext = cb.hi[best_axis] - cb.lo[best_axis];
            auto it = std::partition(prims.begin() + first, prims.begin() + first + count,
                                     [&](const Prim &t) {
                                         int bin = (int)((t.c[best_axis] - cb.lo[best_axis]) /
                                                         ext * kBins);
                                         bin = std::min(std::max(bin, 0), kBins - 1);
                                         return bin <= best_bin;
                                     });
            mid = (int32_t)(it - prims.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        const int32_t rcount = first + count - mid;
        if (fork > 0 && rcount >= kForkMin && mid - first >= kForkMin) {
            Builder rb(prims);
            std::future<int32_t> right;
            try {
                right = std::async(std::launch::async, [&rb, mid, rcount, depth, fork]() {
                    return rb.build(mid, rcount, depth + 1, fork - 1);
                });
            } catch (const std::system_error &) {
                fork = 0;  // no thread to be had: this subtree builds serially
            }
            if (fork > 0) {
                int32_t l = build(first, mid - first, depth + 1, fork - 1);
                int32_t r = right.get();
                const int32_t off = (int32_t)nodes.size();
                for (BNode nd : rb.nodes) {
                    if (nd.left >= 0) {
                        nd.left += off;
                        nd.right += off;
                    }
                    nodes.push_back(nd);
                }
                nodes[me].left = l;
                nodes[me].right = r + off;
                max_depth = std::max(max_depth, rb.max_depth);
                return me;
            }
        }
        int32_t l = build(first, mid - first, depth + 1, fork);
        int32_t r = build(mid, rcount, depth + 1, fork);
        nodes[me].left = l;
        nodes[me].right = r;
        return me;
    }
};

}  // namespace

int build_scene(const float *v9, const int32_t *ids, const uint8_t *flags, int32_t n,
                HostScene &out)
{
    (void)flags;  // transparent triangles are opaque here, as in KdRayTracer.TestLines
    if (n <= 0) {
        set_error("a scene needs at least one triangle");
        return VH_EINVAL;
    }
    if (n >= kMaxOrigIndex) {
        set_error("too many triangles for one scene (limit 2^27)");
        return VH_EINVAL;
    }
    std::vector<Prim> prims(n);
    Builder b(prims);
    float m = 0.0f;
    for (int32_t i = 0; i < n; i++) {
        const float *v = v9 + (size_t)i * 9;
        for (int k = 0; k < 3; k++) {
            float a = v[k], c = v[3 + k], e = v[6 + k];
            if (!std::isfinite(a) || !std::isfinite(c) || !std::isfinite(e)) {
                set_error("triangle " + std::to_string(i) + " has a non-finite vertex");
                return VH_EINVAL;
            }
            Prim &q = prims[i];
            q.box.lo[k] = std::min(a, std::min(c, e));
            q.box.hi[k] = std::max(a, std::max(c, e));
            q.c[k] = 0.5f * (q.box.lo[k] + q.box.hi[k]);
            m = std::max(m, std::max(std::fabs(q.box.lo[k]), std::fabs(q.box.hi[k])));
        }
        prims[i].id = i;
    }
    b.nodes.reserve((size_t)n * 2);
    // up to 16 threads at the fourth level, fewer on a smaller host
    const unsigned hw = build_threads();
    int fork = 0;
    while ((2u << fork) <= hw) fork++;
    b.build(0, n, 0, fork);

    // Widening: the exact test's hit point carries a few ulps of the scene's
    // coordinate magnitude; 2^-16 of it is far beyond that.
    float pad = m * (1.0f / 65536.0f) + 1.0e-6f;
    out.pad = pad;

    // Flatten: only internal nodes become GPU nodes.  Root is always internal.
    std::vector<int32_t> gpu_index(b.nodes.size(), -1);
    std::vector<int32_t> internal;
    int32_t root = 0;
    bool root_leaf = b.nodes[root].left < 0;
    // breadth-first numbering keeps the upper levels together
    std::vector<int32_t> queue{root};
    for (size_t q = 0; q < queue.size(); q++) {
        int32_t bn = queue[q];
        if (b.nodes[bn].left < 0 && !(bn == root && root_leaf)) continue;
        gpu_index[bn] = (int32_t)internal.size();
        internal.push_back(bn);
        if (b.nodes[bn].left >= 0) {
            queue.push_back(b.nodes[bn].left);
            queue.push_back(b.nodes[bn].right);
        }
    }
    // triangles in leaf order: each leaf's start in a serial pass, then the
    // conversions, which are independent, in parallel
    out.tris.assign((size_t)n * 12, 0.0f);
    std::vector<int32_t> leaf_start(b.nodes.size(), -1);
    std::vector<int32_t> leaves;
    int32_t cursor = 0;
    for (size_t bn = 0; bn < b.nodes.size(); bn++)
        if (b.nodes[bn].left < 0 && b.nodes[bn].count > 0) {
            leaf_start[bn] = cursor;
            cursor += b.nodes[bn].count;
            leaves.push_back((int32_t)bn);
        }
    parallel_chunks((int64_t)leaves.size(), [&](int64_t lo, int64_t hi) {
        for (int64_t q = lo; q < hi; q++) {
            const BNode &L = b.nodes[leaves[q]];
            for (int32_t i = 0; i < L.count; i++) {
                int32_t orig = prims[L.first + i].id;
                Tri48 t = convert(v9 + (size_t)orig * 9, ids ? ids[orig] : orig);
                float *o = &out.tris[(size_t)(leaf_start[leaves[q]] + i) * 12];
                o[0] = t.n[0];
                o[1] = t.n[1];
                o[2] = t.n[2];
                o[3] = t.d;
                o[4] = t.e[0];
                o[5] = t.e[1];
                o[6] = t.e[2];
                o[7] = t.e[3];
                o[8] = t.e[4];
                o[9] = t.e[5];
                std::memcpy(&o[10], &t.id, 4);
                uint32_t meta = (uint32_t)t.cs0 | ((uint32_t)t.cs1 << 2) |
                                (i == L.count - 1 ? kTriLast : 0u) | ((uint32_t)orig << kOrigShift);
                std::memcpy(&o[11], &meta, 4);
            }
        }
    });

    auto ref_of = [&](int32_t bn) -> int32_t {
        if (bn < 0) return kEmptyRef;
        if (b.nodes[bn].left >= 0) return gpu_index[bn];
        if (b.nodes[bn].count == 0) return kEmptyRef;
        return ~leaf_start[bn];
    };
    out.nnodes = (int32_t)internal.size();
    out.nodes.assign((size_t)out.nnodes * 16, 0.0f);
    // each GPU node is written from its build node alone
    parallel_chunks(out.nnodes, [&](int64_t glo, int64_t ghi) {
    for (int64_t g = glo; g < ghi; g++) {
        int32_t bn = internal[g];
        int32_t c0, c1;
        Box b0, b1;
        if (bn == root && root_leaf) {
            c0 = bn;
            c1 = -1;
            b0 = b.nodes[bn].box;
            b1.reset();
        } else {
            c0 = b.nodes[bn].left;
            c1 = b.nodes[bn].right;
            b0 = b.nodes[c0].box;
            b1 = b.nodes[c1].box;
        }
        auto padded = [&](Box x) {
            if (x.lo[0] <= x.hi[0])
                for (int k = 0; k < 3; k++) {
                    x.lo[k] -= pad;
                    x.hi[k] += pad;
                }
            else
                for (int k = 0; k < 3; k++) {
                    x.lo[k] = INFINITY;
                    x.hi[k] = -INFINITY;
                }
            return x;
        };
        b0 = padded(b0);
        b1 = padded(b1);
        float *o = &out.nodes[(size_t)g * 16];
        o[0] = b0.lo[0];
        o[1] = b0.hi[0];
        o[2] = b0.lo[1];
        o[3] = b0.hi[1];
        o[4] = b1.lo[0];
        o[5] = b1.hi[0];
        o[6] = b1.lo[1];
        o[7] = b1.hi[1];
        o[8] = b0.lo[2];
        o[9] = b0.hi[2];
        o[10] = b1.lo[2];
        o[11] = b1.hi[2];
        int32_t r0 = (bn == root && root_leaf) ? (b.nodes[bn].count ? ~leaf_start[bn] : kEmptyRef)
                                               : ref_of(c0);
        int32_t r1 = (bn == root && root_leaf) ? kEmptyRef : ref_of(c1);
        std::memcpy(&o[12], &r0, 4);
        std::memcpy(&o[13], &r1, 4);
        // the children's order along the axis their centres differ most on:
        // bits 0-1 the axis, bit 2 set when child 0's centre is the higher
        int32_t ord = 0;
        if (!(bn == root && root_leaf)) {
            float best = -1.0f;
            for (int k = 0; k < 3; k++) {
                float c0 = 0.5f * (b0.lo[k] + b0.hi[k]), c1 = 0.5f * (b1.lo[k] + b1.hi[k]);
                float dk = std::fabs(c1 - c0);
                if (dk > best) {
                    best = dk;
                    ord = k | (c0 > c1 ? 4 : 0);
                }
            }
        }
        std::memcpy(&o[14], &ord, 4);
    }
    });
    out.ntri = n;
    out.depth = b.max_depth + (root_leaf ? 1 : 0);
    Box sb = b.nodes[root].box;
    for (int k = 0; k < 3; k++) {
        out.lo[k] = sb.lo[k];
        out.hi[k] = sb.hi[k];
    }
    return VH_OK;
}

}  // namespace vh
