// The scene build on the GPU: the host builder's binned SAH (vh_bvh.cpp) run
// level by level, with the same 64 bins per axis, the same cost and leaf rule
// and the same split predicate, so the tree has the host tree's shape and
// bounds; nodes come out in the same breadth-first order and leaves in the
// same left-to-right order.  Only the order of the triangles inside a leaf
// may differ (the host partitions in place, this build stably), and a node
// whose centroids all coincide, split by position, may split differently.
//
// Per level, over the level's nodes:
//   * nodes of more than kSmall triangles ("big") are binned by blocks over
//     chunks of kChunk triangles, each block's LDS bins merged into the
//     node's global bins with integer atomics (order-independent: counts add,
//     bounds are min/max of order-preserving integer images of the floats);
//   * one wave per node sweeps the 3 x 64 bins in parallel (prefix and suffix
//     unions by shuffles; costs per lane; the lowest cost, ties to the lower
//     axis then bin, as the host's strict `<` in axis-then-bin order), and a
//     small node is binned and partitioned by that wave alone;
//   * a big node is partitioned by its chunks: counts, a scan per node, then
//     a stable scatter, the children's centroid bounds by atomics;
//   * the split nodes are numbered in level order (the host's breadth-first
//     numbering), write their GPU node (children's padded boxes and the
//     near-child axis) and hand their children to the next level, which
//     writes each child's reference into its parent.
// The triangles are converted to the intersection format in leaf order at the
// end (the host's convert, operation for operation).
#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "vh_internal.h"

#pragma clang fp contract(off)

namespace vh {
namespace {

constexpr int kBins = 64;
constexpr float kTraverse = 1.0f;
constexpr float kIntersect = 1.5f;
constexpr int kLeafMax = 8;
constexpr int kSmall = 256;   // nodes of at most this many triangles: one wave does all
constexpr int kChunk = 1024;  // triangles per binning block of a big node
constexpr int kBinInts = 3 * kBins * 7;  // per node: count and 6 bounds per bin and axis

// Order-preserving integer image of a float (for min/max by integer atomics).
__device__ __forceinline__ int f2o(float x)
{
    const int i = __float_as_int(x);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

// std::min / std::max as the host build uses them
__device__ __forceinline__ float mn(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return (a < b) ? b : a; }

struct FBox {
    float lo[3], hi[3];
};

__device__ __forceinline__ void box_reset(FBox &b)
{
    for (int k = 0; k < 3; k++) {
        b.lo[k] = INFINITY;
        b.hi[k] = -INFINITY;
    }
}

__device__ __forceinline__ float box_area(const FBox &b)
{
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
}

// A triangle as the build sees it, 48 bytes: (lo, c.x), (hi, c.y), (c.z, index).
struct Prim {
    float4 a, b, d;
};

// A node of the level being built.  Bounds as order-preserving integers.
struct LNode {
    int32_t first, count, depth, pslot;  // pslot: parent GPU index * 2 + side, -1 for the root
    int32_t box[6], cb[6];               // lo xyz, hi xyz
};

// A node's decision and its children's data.
struct NDec {
    int32_t split, axis, bin, nleft;  // axis -1: split by position
    int32_t lbox[6], rbox[6], lcb[6], rcb[6];
};

struct LevelArgs {
    const Prim *src;
    Prim *dst;
    const LNode *nodes;
    int32_t na;
    NDec *dec;
    const int32_t *task_off;  // per node, exclusive prefix of its binning tasks (na + 1)
    const int32_t *big_rank;  // per node, index among the big nodes
    int32_t *gbin;            // per big node, kBinInts
    int32_t *task_left;       // per task: triangles going left
    int32_t *order;           // final leaf order: original triangle index
    uint8_t *last;            // final leaf order: last triangle of a leaf
};

__device__ __forceinline__ int bin_of(float c, float lo, float ext)
{
    int bin = (int)((c - lo) / ext * kBins);
    return min(max(bin, 0), kBins - 1);
}

__device__ __forceinline__ float prim_c(const Prim &p, int ax) { return ax == 0 ? p.a.w : (ax == 1 ? p.b.w : p.d.x); }

__device__ __forceinline__ void node_bounds(const LNode &nd, FBox &b, FBox &cb)
{
    for (int k = 0; k < 3; k++) {
        b.lo[k] = o2f(nd.box[k]);
        b.hi[k] = o2f(nd.box[3 + k]);
        cb.lo[k] = o2f(nd.cb[k]);
        cb.hi[k] = o2f(nd.cb[3 + k]);
    }
}

// Clears a bin set: counts 0, bounds empty.
__device__ __forceinline__ void bins_clear(int32_t *bins, int t, int nt)
{
    for (int i = t; i < kBinInts; i += nt) {
        const int f = i % 7;
        bins[i] = f == 0 ? 0 : (f <= 3 ? f2o(INFINITY) : f2o(-INFINITY));
    }
}

// Adds one triangle to a bin set (LDS or global) by atomics.
__device__ __forceinline__ void bins_add(int32_t *bins, const Prim &p, const float (&clo)[3], const float (&ext)[3],
                                         const bool (&live)[3])
{
    const float lo[3] = {p.a.x, p.a.y, p.a.z}, hi[3] = {p.b.x, p.b.y, p.b.z};
    for (int ax = 0; ax < 3; ax++) {
        if (!live[ax]) continue;
        int32_t *e = bins + (ax * kBins + bin_of(prim_c(p, ax), clo[ax], ext[ax])) * 7;
        atomicAdd(&e[0], 1);
        for (int k = 0; k < 3; k++) {
            atomicMin(&e[1 + k], f2o(lo[k]));
            atomicMax(&e[4 + k], f2o(hi[k]));
        }
    }
}

// The SAH sweep of one node's bins by one wave (lane = bin): the host's
// best (axis, bin) and cost, or axis -1 with cost +inf.  Also each side's
// box and count at the best split.
__device__ void sweep(const int32_t *bins, const FBox &b, const bool (&live)[3], int &best_axis, int &best_bin,
                      float &best_cost, int &nleft, FBox &lb, FBox &rb)
{
    const int lane = threadIdx.x & 63;
    const float parea = box_area(b);
    float bc = INFINITY;
    int bkey = 1 << 30;
    FBox bl, br;
    box_reset(bl);
    box_reset(br);
    int bnl = 0;
    for (int ax = 0; ax < 3; ax++) {
        if (!live[ax]) continue;
        const int32_t *e = bins + (ax * kBins + lane) * 7;
        const int cnt = e[0];
        FBox mine;
        for (int k = 0; k < 3; k++) {
            mine.lo[k] = o2f(e[1 + k]);
            mine.hi[k] = o2f(e[4 + k]);
        }
        // inclusive prefix (bins 0 .. lane) and suffix (lane .. 63)
        FBox pre = mine, suf = mine;
        int pc = cnt, sc = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            float v[6], w[6];
            for (int k = 0; k < 3; k++) {
                v[k] = __shfl_up(pre.lo[k], off, 64);
                v[3 + k] = __shfl_up(pre.hi[k], off, 64);
                w[k] = __shfl_down(suf.lo[k], off, 64);
                w[3 + k] = __shfl_down(suf.hi[k], off, 64);
            }
            const int vc = __shfl_up(pc, off, 64), wc = __shfl_down(sc, off, 64);
            if (lane >= off) {
                for (int k = 0; k < 3; k++) {
                    pre.lo[k] = mn(pre.lo[k], v[k]);
                    pre.hi[k] = mx(pre.hi[k], v[3 + k]);
                }
                pc += vc;
            }
            if (lane + off < 64) {
                for (int k = 0; k < 3; k++) {
                    suf.lo[k] = mn(suf.lo[k], w[k]);
                    suf.hi[k] = mx(suf.hi[k], w[3 + k]);
                }
                sc += wc;
            }
        }
        // the right side of split `lane` is the suffix from lane + 1
        FBox right;
        for (int k = 0; k < 3; k++) {
            right.lo[k] = __shfl_down(suf.lo[k], 1, 64);
            right.hi[k] = __shfl_down(suf.hi[k], 1, 64);
        }
        const int rc = __shfl_down(sc, 1, 64);
        if (lane < kBins - 1 && pc != 0 && rc != 0) {
            const float cost = kTraverse + kIntersect * (box_area(pre) * pc + box_area(right) * rc) /
                                               (parea > 0 ? parea : 1.0f);
            if (cost < bc) {
                bc = cost;
                bkey = ax * kBins + lane;
                bl = pre;
                br = right;
                bnl = pc;
            }
        }
    }
    // the lowest cost over the wave, ties to the lower key
    float wc = bc;
    int wk = bkey;
    for (int off = 32; off > 0; off >>= 1) {
        const float oc = __shfl_xor(wc, off, 64);
        const int ok = __shfl_xor(wk, off, 64);
        if (oc < wc || (oc == wc && ok < wk)) {
            wc = oc;
            wk = ok;
        }
    }
    best_cost = wc;
    if (wk == (1 << 30)) {
        best_axis = -1;
        best_bin = 0;
        nleft = 0;
        return;
    }
    best_axis = wk / kBins;
    best_bin = wk % kBins;
    const int owner = best_bin;  // the lane holding the winning split (its own best is this one)
    nleft = __shfl(bnl, owner, 64);
    for (int k = 0; k < 3; k++) {
        lb.lo[k] = __shfl(bl.lo[k], owner, 64);
        lb.hi[k] = __shfl(bl.hi[k], owner, 64);
        rb.lo[k] = __shfl(br.lo[k], owner, 64);
        rb.hi[k] = __shfl(br.hi[k], owner, 64);
    }
}

__device__ __forceinline__ void store_box(int32_t *dst, const FBox &b)
{
    for (int k = 0; k < 3; k++) {
        dst[k] = f2o(b.lo[k]);
        dst[3 + k] = f2o(b.hi[k]);
    }
}

// The triangles' build records, the scene bounds, the largest coordinate
// magnitude and the first non-finite triangle.
__global__ __launch_bounds__(256) void k_prims(const float *__restrict__ v9, int32_t n, Prim *__restrict__ out,
                                               int32_t *__restrict__ root, int32_t *__restrict__ maxabs,
                                               int32_t *__restrict__ bad)
{
    __shared__ int32_t red[13];
    if (threadIdx.x < 13) red[threadIdx.x] = threadIdx.x < 3 || (threadIdx.x >= 6 && threadIdx.x < 9)
                                                 ? f2o(INFINITY)
                                                 : (threadIdx.x < 12 ? f2o(-INFINITY) : 0);
    __syncthreads();
    const int32_t i = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    if (i < n) {
        const float *v = v9 + (size_t)i * 9;
        float lo[3], hi[3], c[3], m = 0.0f;
        bool fin = true;
        for (int k = 0; k < 3; k++) {
            const float a = v[k], cc = v[3 + k], e = v[6 + k];
            fin = fin && isfinite(a) && isfinite(cc) && isfinite(e);
            lo[k] = mn(a, mn(cc, e));
            hi[k] = mx(a, mx(cc, e));
            c[k] = 0.5f * (lo[k] + hi[k]);
            m = mx(m, mx(fabsf(lo[k]), fabsf(hi[k])));
        }
        if (!fin) {
            atomicMin(bad, i);
        } else {
            out[i].a = make_float4(lo[0], lo[1], lo[2], c[0]);
            out[i].b = make_float4(hi[0], hi[1], hi[2], c[1]);
            out[i].d = make_float4(c[2], __int_as_float(i), 0.0f, 0.0f);
            for (int k = 0; k < 3; k++) {
                atomicMin(&red[k], f2o(lo[k]));
                atomicMax(&red[3 + k], f2o(hi[k]));
                atomicMin(&red[6 + k], f2o(c[k]));
                atomicMax(&red[9 + k], f2o(c[k]));
            }
            atomicMax(&red[12], __float_as_int(m));  // non-negative: the bits order
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        atomicMin(&root[threadIdx.x], red[threadIdx.x]);
        atomicMax(&root[3 + threadIdx.x], red[3 + threadIdx.x]);
        atomicMin(&root[6 + threadIdx.x], red[6 + threadIdx.x]);
        atomicMax(&root[9 + threadIdx.x], red[9 + threadIdx.x]);
    }
    if (threadIdx.x == 0) atomicMax(maxabs, red[12]);
}

// The root's level record from the reduced bounds.
__global__ void k_root(const int32_t *__restrict__ root, int32_t n, LNode *__restrict__ lv)
{
    LNode r;
    r.first = 0;
    r.count = n;
    r.depth = 0;
    r.pslot = -1;
    for (int k = 0; k < 6; k++) r.box[k] = root[k];
    for (int k = 0; k < 6; k++) r.cb[k] = root[6 + k];
    lv[0] = r;
}

// Per node of a level: its binning tasks (big nodes) and whether it is big;
// entry na is 0 (the scans' total).
__global__ void k_level_counts(const LNode *__restrict__ nodes, int32_t na, int32_t *__restrict__ ntask,
                               int32_t *__restrict__ isbig)
{
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == na) {
        ntask[i] = 0;
        isbig[i] = 0;
    }
    if (i >= na) return;
    const LNode nd = nodes[i];
    const bool big = nd.count > kSmall && nd.depth < kMaxDepth;
    ntask[i] = big ? (nd.count + kChunk - 1) / kChunk : 0;
    isbig[i] = big ? 1 : 0;
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_clear_bins(int32_t *__restrict__ gbin, int32_t nbig)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)nbig * kBinInts) return;
    const int f = (int)(i % kBinInts) % 7;
    gbin[i] = f == 0 ? 0 : (f <= 3 ? f2o(INFINITY) : f2o(-INFINITY));
}

// The node of task t (binary search over the per-node task prefix).
__device__ __forceinline__ int32_t task_node(const int32_t *task_off, int32_t na, int32_t t)
{
    int32_t lo = 0, hi = na;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (task_off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_split_flags(const NDec *__restrict__ dec, int32_t na, int32_t *__restrict__ flags)
{
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < na) flags[i] = dec[i].split;
}

// Big nodes: a block per chunk bins into LDS, then into the node's bins.
__global__ __launch_bounds__(256) void k_big_bin(LevelArgs a)
{
    __shared__ int32_t bins[kBinInts];
    const int32_t t = blockIdx.x;
    const int32_t i = task_node(a.task_off, a.na, t);
    const LNode nd = a.nodes[i];
    FBox b, cb;
    node_bounds(nd, b, cb);
    float ext[3];
    bool live[3];
    for (int k = 0; k < 3; k++) {
        ext[k] = cb.hi[k] - cb.lo[k];
        live[k] = ext[k] > 0;
    }
    bins_clear(bins, threadIdx.x, 256);
    __syncthreads();
    const int32_t c0 = nd.first + (t - a.task_off[i]) * kChunk;
    const int32_t c1 = min(nd.first + nd.count, c0 + kChunk);
    for (int32_t j = c0 + (int32_t)threadIdx.x; j < c1; j += 256) bins_add(bins, a.src[j], cb.lo, ext, live);
    __syncthreads();
    int32_t *g = a.gbin + (size_t)a.big_rank[i] * kBinInts;
    for (int q = threadIdx.x; q < kBinInts; q += 256) {
        const int f = q % 7;
        if (f == 0) {
            if (bins[q]) atomicAdd(&g[q], bins[q]);
        } else if (f <= 3) {
            atomicMin(&g[q], bins[q]);
        } else {
            atomicMax(&g[q], bins[q]);
        }
    }
}

// One wave per node: the split decision (bins: the big node's global ones,
// or the small node's own, made here); a small node is also partitioned here
// (stably) with its children's bounds.  Leaves write the final order.
__global__ __launch_bounds__(256) void k_decide(LevelArgs a)
{
    __shared__ int32_t wbins[4][kBinInts];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t i = blockIdx.x * 4 + w;
    if (i >= a.na) return;  // whole wave
    const LNode nd = a.nodes[i];
    FBox b, cb;
    node_bounds(nd, b, cb);
    float ext[3];
    bool live[3];
    for (int k = 0; k < 3; k++) {
        ext[k] = cb.hi[k] - cb.lo[k];
        live[k] = ext[k] > 0;
    }
    const bool big = nd.count > kSmall && nd.depth < kMaxDepth;
    int best_axis = -1, best_bin = 0, nleft = 0;
    float best_cost = INFINITY;
    FBox lb, rb;
    box_reset(lb);
    box_reset(rb);
    const int32_t *bins = nullptr;
    if (nd.count > 1 && nd.depth < kMaxDepth) {
        if (big) {
            bins = a.gbin + (size_t)a.big_rank[i] * kBinInts;
        } else {
            int32_t *mine = wbins[w];
            bins_clear(mine, lane, 64);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            for (int32_t j = nd.first + lane; j < nd.first + nd.count; j += 64)
                bins_add(mine, a.src[j], cb.lo, ext, live);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            bins = mine;
        }
        sweep(bins, b, live, best_axis, best_bin, best_cost, nleft, lb, rb);
    }
    const float leaf_cost = kIntersect * nd.count;
    const bool leaf = nd.count <= 1 || nd.depth >= kMaxDepth || (nd.count <= kLeafMax && !(best_cost < leaf_cost));
    NDec &d = a.dec[i];
    if (leaf) {
        for (int32_t j = nd.first + lane; j < nd.first + nd.count; j += 64) {
            a.order[j] = __float_as_int(a.src[j].d.y);
            a.last[j] = j == nd.first + nd.count - 1 ? 1 : 0;
        }
        if (lane == 0) {
            d.split = 0;
            d.axis = -1;
            d.bin = 0;
            d.nleft = 0;
        }
        return;
    }
    if (best_axis < 0) nleft = nd.count / 2;  // coincident centroids: split the list
    if (lane == 0) {
        d.split = 1;
        d.axis = best_axis;
        d.bin = best_bin;
        d.nleft = nleft;
        if (best_axis >= 0) {
            store_box(d.lbox, lb);
            store_box(d.rbox, rb);
        }
        FBox e;
        box_reset(e);
        store_box(d.lcb, e);
        store_box(d.rcb, e);
        if (best_axis < 0) {  // the sides' boxes come from the partition below
            store_box(d.lbox, e);
            store_box(d.rbox, e);
        }
    }
    if (big) return;  // partitioned by its chunks
    // the small node's stable partition and its children's bounds
    FBox lcb, rcb, lbx, rbx;
    box_reset(lcb);
    box_reset(rcb);
    box_reset(lbx);
    box_reset(rbx);
    int32_t lbase = 0, rbase = 0;
    for (int32_t j0 = nd.first; j0 < nd.first + nd.count; j0 += 64) {
        const int32_t j = j0 + lane;
        const bool have = j < nd.first + nd.count;
        Prim p{};
        if (have) p = a.src[j];
        bool left;
        if (best_axis >= 0)
            left = bin_of(prim_c(p, best_axis), cb.lo[best_axis], ext[best_axis]) <= best_bin;
        else
            left = (j - nd.first) < nleft;
        const uint64_t bl = __builtin_amdgcn_ballot_w64(have && left);
        const uint64_t br = __builtin_amdgcn_ballot_w64(have && !left);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        if (have) {
            const int32_t pos = left ? nd.first + lbase + __builtin_popcountll(bl & below)
                                     : nd.first + nleft + rbase + __builtin_popcountll(br & below);
            a.dst[pos] = p;
            const float pc[3] = {p.a.w, p.b.w, p.d.x};
            const float plo[3] = {p.a.x, p.a.y, p.a.z}, phi[3] = {p.b.x, p.b.y, p.b.z};
            for (int k = 0; k < 3; k++) {
                if (left) {
                    lcb.lo[k] = mn(lcb.lo[k], pc[k]);
                    lcb.hi[k] = mx(lcb.hi[k], pc[k]);
                    lbx.lo[k] = mn(lbx.lo[k], plo[k]);
                    lbx.hi[k] = mx(lbx.hi[k], phi[k]);
                } else {
                    rcb.lo[k] = mn(rcb.lo[k], pc[k]);
                    rcb.hi[k] = mx(rcb.hi[k], pc[k]);
                    rbx.lo[k] = mn(rbx.lo[k], plo[k]);
                    rbx.hi[k] = mx(rbx.hi[k], phi[k]);
                }
            }
        }
        lbase += __builtin_popcountll(bl);
        rbase += __builtin_popcountll(br);
    }
    // wave min/max of the sides' bounds
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 3; k++) {
            lcb.lo[k] = mn(lcb.lo[k], __shfl_xor(lcb.lo[k], off, 64));
            lcb.hi[k] = mx(lcb.hi[k], __shfl_xor(lcb.hi[k], off, 64));
            rcb.lo[k] = mn(rcb.lo[k], __shfl_xor(rcb.lo[k], off, 64));
            rcb.hi[k] = mx(rcb.hi[k], __shfl_xor(rcb.hi[k], off, 64));
            lbx.lo[k] = mn(lbx.lo[k], __shfl_xor(lbx.lo[k], off, 64));
            lbx.hi[k] = mx(lbx.hi[k], __shfl_xor(lbx.hi[k], off, 64));
            rbx.lo[k] = mn(rbx.lo[k], __shfl_xor(rbx.lo[k], off, 64));
            rbx.hi[k] = mx(rbx.hi[k], __shfl_xor(rbx.hi[k], off, 64));
        }
    }
    if (lane == 0) {
        store_box(d.lcb, lcb);
        store_box(d.rcb, rcb);
        store_box(d.lbox, lbx);
        store_box(d.rbox, rbx);
    }
}

// Big split nodes, per chunk: how many of its triangles go left.
__global__ __launch_bounds__(256) void k_big_count(LevelArgs a)
{
    __shared__ int32_t cnt;
    const int32_t t = blockIdx.x;
    const int32_t i = task_node(a.task_off, a.na, t);
    const LNode nd = a.nodes[i];
    const NDec d = a.dec[i];
    if (!d.split) return;  // uniform over the block
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    FBox b, cb;
    node_bounds(nd, b, cb);
    const int32_t c0 = nd.first + (t - a.task_off[i]) * kChunk;
    const int32_t c1 = min(nd.first + nd.count, c0 + kChunk);
    int32_t mine = 0;
    for (int32_t j = c0 + (int32_t)threadIdx.x; j < c1; j += 256) {
        bool left;
        if (d.axis >= 0) {
            const float ext = cb.hi[d.axis] - cb.lo[d.axis];
            left = bin_of(prim_c(a.src[j], d.axis), cb.lo[d.axis], ext) <= d.bin;
        } else {
            left = (j - nd.first) < d.nleft;
        }
        mine += left ? 1 : 0;
    }
    atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) a.task_left[t] = cnt;
}

// Per big node: the exclusive prefix of its chunks' left counts (in place).
__global__ void k_big_scan(const int32_t *__restrict__ task_off, int32_t na, const NDec *__restrict__ dec,
                           const int32_t *__restrict__ isbig, int32_t *__restrict__ task_left)
{
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= na || !isbig[i] || !dec[i].split) return;
    int32_t run = 0;
    for (int32_t t = task_off[i]; t < task_off[i + 1]; t++) {
        const int32_t v = task_left[t];
        task_left[t] = run;
        run += v;
    }
}

// Big split nodes, per chunk: the stable scatter, and the children's
// centroid bounds (and, for a split by position, their boxes) by atomics.
__global__ __launch_bounds__(256) void k_big_scatter(LevelArgs a)
{
    __shared__ int32_t wl[4], wr[4];
    __shared__ int32_t red[24];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t t = blockIdx.x;
    const int32_t i = task_node(a.task_off, a.na, t);
    const LNode nd = a.nodes[i];
    NDec &d = a.dec[i];
    const int32_t axis = d.axis, bin = d.bin, nleft = d.nleft;
    if (!d.split) return;  // uniform over the block
    if (threadIdx.x < 24) {
        const int k = threadIdx.x % 6;
        red[threadIdx.x] = k < 3 ? f2o(INFINITY) : f2o(-INFINITY);
    }
    FBox b, cb;
    node_bounds(nd, b, cb);
    const int32_t ci = t - a.task_off[i];
    const int32_t c0 = nd.first + ci * kChunk;
    const int32_t c1 = min(nd.first + nd.count, c0 + kChunk);
    const int32_t lstart = a.task_left[t];                  // lefts before this chunk
    const int32_t rstart = (c0 - nd.first) - lstart;         // rights before this chunk
    __syncthreads();
    int32_t lrun = 0, rrun = 0;
    float lc[6], rc[6], lx[6], rx[6];
    for (int k = 0; k < 3; k++) {
        lc[k] = rc[k] = lx[k] = rx[k] = INFINITY;
        lc[3 + k] = rc[3 + k] = lx[3 + k] = rx[3 + k] = -INFINITY;
    }
    for (int32_t j0 = c0; j0 < c1; j0 += 256) {
        const int32_t j = j0 + (int32_t)threadIdx.x;
        const bool have = j < c1;
        Prim p{};
        if (have) p = a.src[j];
        bool left = false;
        if (have) {
            if (axis >= 0)
                left = bin_of(prim_c(p, axis), cb.lo[axis], cb.hi[axis] - cb.lo[axis]) <= bin;
            else
                left = (j - nd.first) < nleft;
        }
        const uint64_t bl = __builtin_amdgcn_ballot_w64(have && left);
        const uint64_t br = __builtin_amdgcn_ballot_w64(have && !left);
        if (lane == 0) {
            wl[w] = __builtin_popcountll(bl);
            wr[w] = __builtin_popcountll(br);
        }
        __syncthreads();
        int32_t lpre = 0, rpre = 0, ltot = 0, rtot = 0;
        for (int q = 0; q < 4; q++) {
            if (q < w) {
                lpre += wl[q];
                rpre += wr[q];
            }
            ltot += wl[q];
            rtot += wr[q];
        }
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        if (have) {
            const int32_t pos = left ? nd.first + lstart + lrun + lpre + __builtin_popcountll(bl & below)
                                     : nd.first + nleft + rstart + rrun + rpre + __builtin_popcountll(br & below);
            a.dst[pos] = p;
            const float pc[3] = {p.a.w, p.b.w, p.d.x};
            const float plo[3] = {p.a.x, p.a.y, p.a.z}, phi[3] = {p.b.x, p.b.y, p.b.z};
            for (int k = 0; k < 3; k++) {
                if (left) {
                    lc[k] = mn(lc[k], pc[k]);
                    lc[3 + k] = mx(lc[3 + k], pc[k]);
                    lx[k] = mn(lx[k], plo[k]);
                    lx[3 + k] = mx(lx[3 + k], phi[k]);
                } else {
                    rc[k] = mn(rc[k], pc[k]);
                    rc[3 + k] = mx(rc[3 + k], pc[k]);
                    rx[k] = mn(rx[k], plo[k]);
                    rx[3 + k] = mx(rx[3 + k], phi[k]);
                }
            }
        }
        lrun += ltot;
        rrun += rtot;
        __syncthreads();
    }
    for (int k = 0; k < 6; k++) {
        if (k < 3) {
            atomicMin(&red[k], f2o(lc[k]));
            atomicMin(&red[6 + k], f2o(rc[k]));
            atomicMin(&red[12 + k], f2o(lx[k]));
            atomicMin(&red[18 + k], f2o(rx[k]));
        } else {
            atomicMax(&red[k], f2o(lc[k]));
            atomicMax(&red[6 + k], f2o(rc[k]));
            atomicMax(&red[12 + k], f2o(lx[k]));
            atomicMax(&red[18 + k], f2o(rx[k]));
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        if (k < 3) {
            atomicMin(&d.lcb[k], red[k]);
            atomicMin(&d.rcb[k], red[6 + k]);
        } else {
            atomicMax(&d.lcb[k], red[k]);
            atomicMax(&d.rcb[k], red[6 + k]);
        }
        if (axis < 0) {  // split by position: the sides' boxes from the triangles
            if (k < 3) {
                atomicMin(&d.lbox[k], red[12 + k]);
                atomicMin(&d.rbox[k], red[18 + k]);
            } else {
                atomicMax(&d.lbox[k], red[12 + k]);
                atomicMax(&d.rbox[k], red[18 + k]);
            }
        }
    }
}

// Per node: its reference into the parent; for a split node its GPU node
// (padded child boxes, near-child axis) and its children at the next level.
// `rank` is the exclusive prefix of the split flags.
__global__ void k_children(const LNode *__restrict__ nodes, int32_t na, const NDec *__restrict__ dec,
                           const int32_t *__restrict__ rank, int32_t gbase, const int32_t *__restrict__ maxabs,
                           float *__restrict__ gnodes, LNode *__restrict__ next)
{
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= na) return;
    const LNode nd = nodes[i];
    const NDec d = dec[i];
    const int32_t gi = gbase + rank[i];
    if (nd.pslot >= 0) {
        int32_t ref = d.split ? gi : ~nd.first;
        gnodes[(size_t)(nd.pslot >> 1) * 16 + 12 + (nd.pslot & 1)] = __int_as_float(ref);
    }
    if (!d.split) return;
    // the host's widening: 2^-16 of the largest coordinate magnitude, plus 1e-6
    const float m = __int_as_float(*maxabs);
    const float pad = m * (1.0f / 65536.0f) + 1.0e-6f;
    FBox b0, b1;
    for (int k = 0; k < 3; k++) {
        b0.lo[k] = o2f(d.lbox[k]) - pad;
        b0.hi[k] = o2f(d.lbox[3 + k]) + pad;
        b1.lo[k] = o2f(d.rbox[k]) - pad;
        b1.hi[k] = o2f(d.rbox[3 + k]) + pad;
    }
    float *o = gnodes + (size_t)gi * 16;
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
    int32_t ord = 0;
    float best = -1.0f;
    for (int k = 0; k < 3; k++) {
        const float c0 = 0.5f * (b0.lo[k] + b0.hi[k]), c1 = 0.5f * (b1.lo[k] + b1.hi[k]);
        const float dk = fabsf(c1 - c0);
        if (dk > best) {
            best = dk;
            ord = k | (c0 > c1 ? 4 : 0);
        }
    }
    o[14] = __int_as_float(ord);
    o[15] = 0.0f;
    LNode l, r;
    l.first = nd.first;
    l.count = d.nleft;
    r.first = nd.first + d.nleft;
    r.count = nd.count - d.nleft;
    l.depth = r.depth = nd.depth + 1;
    l.pslot = gi * 2;
    r.pslot = gi * 2 + 1;
    for (int k = 0; k < 6; k++) {
        l.box[k] = d.lbox[k];
        r.box[k] = d.rbox[k];
        l.cb[k] = d.lcb[k];
        r.cb[k] = d.rcb[k];
    }
    next[2 * rank[i]] = l;
    next[2 * rank[i] + 1] = r;
}

// KdTreeBuilder.GetEdgeEquation and ToIntersectionFormat as the host's
// convert (vh_bvh.cpp), operation for operation.
__device__ __forceinline__ void edge_eq(const float *p1, const float *p2, int c1, int c2, const float *inside,
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

__global__ __launch_bounds__(256) void k_tris(const float *__restrict__ v9, const int32_t *__restrict__ ids,
                                              const int32_t *__restrict__ order, const uint8_t *__restrict__ last,
                                              int32_t n, float4 *__restrict__ tris, int32_t *__restrict__ unplaced)
{
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int32_t orig = order[j];
    if (orig < 0 || orig >= n) {  // a slot no leaf filled: reported, never read through
        atomicOr(unplaced, 1);
        for (int q = 0; q < 3; q++) tris[(size_t)j * 3 + q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    float v[9];
    for (int k = 0; k < 9; k++) v[k] = v9[(size_t)orig * 9 + k];
    const float *p1 = v, *p2 = v + 3, *p3 = v + 6;
    const float e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const float e2[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    const float c[3] = {(e1[1] * e2[2]) - (e1[2] * e2[1]), (e1[2] * e2[0]) - (e1[0] * e2[2]),
                        (e1[0] * e2[1]) - (e1[1] * e2[0])};
    const float len = sqrtf((c[0] * c[0]) + (c[1] * c[1]) + (c[2] * c[2]));
    float nn[3];
    if (len == 0.0f) {
        nn[0] = nn[1] = nn[2] = 0.0f;
    } else {
        nn[0] = c[0] / len;
        nn[1] = c[1] / len;
        nn[2] = c[2] / len;
    }
    int drop = 0;
    for (int k = 1; k < 3; k++)
        if (fabsf(nn[k]) > fabsf(nn[drop])) drop = k;
    const int cs0 = (drop + 1) % 3, cs1 = (drop + 2) % 3;
    float e[6];
    edge_eq(p1, p2, cs0, cs1, p3, e);
    edge_eq(p2, p3, cs0, cs1, p1, e + 3);
    const float dd = (nn[0] * p1[0]) + (nn[1] * p1[1]) + (nn[2] * p1[2]);
    const int32_t id = ids ? ids[orig] : orig;
    const uint32_t meta = (uint32_t)cs0 | ((uint32_t)cs1 << 2) | (last[j] ? kTriLast : 0u) |
                          ((uint32_t)orig << kOrigShift);
    tris[(size_t)j * 3 + 0] = make_float4(nn[0], nn[1], nn[2], dd);
    tris[(size_t)j * 3 + 1] = make_float4(e[0], e[1], e[2], e[3]);
    tris[(size_t)j * 3 + 2] = make_float4(e[4], e[5], __int_as_float(id), __uint_as_float(meta));
}

inline uint32_t grid(int64_t n, int per = 256) { return (uint32_t)((n + per - 1) / per); }

// An exclusive prefix sum of n ints (out has n + 1 entries: the total last).
int exclusive_scan(vh_ctx *c, DevBuf &tmp, const int32_t *in, int32_t *out, int32_t n, hipStream_t s)
{
    size_t bytes = 0;
    VH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n + 1, s));
    int rc = tmp.ensure(bytes, "scan scratch");
    if (rc) return rc;
    VH_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, in, out, n + 1, s));
    (void)c;
    return VH_OK;
}

}  // namespace

int build_scene_device(vh_ctx *c, const float *v9, const int32_t *ids, int32_t n)
{
    hipStream_t s = c->stream;
    DevBuf dv, dids, pa, pb, lva, lvb, dec, ntask, isbig, toff, brank, flags, srank, gbin, tleft, order, last,
        red, tmp;
    auto release_all = [&]() {
        for (DevBuf *b : {&dv, &dids, &pa, &pb, &lva, &lvb, &dec, &ntask, &isbig, &toff, &brank, &flags, &srank,
                          &gbin, &tleft, &order, &last, &red, &tmp})
            b->release();
    };
    struct Guard {
        std::function<void()> f;
        ~Guard() { f(); }
    } cleanup{release_all};
    const size_t N = (size_t)n;
    int rc;
    if ((rc = dv.ensure(N * 9 * 4, "scene vertices")) || (rc = pa.ensure(N * sizeof(Prim), "build records")) ||
        (rc = pb.ensure(N * sizeof(Prim), "build records")) || (rc = lva.ensure(N * sizeof(LNode), "level nodes")) ||
        (rc = lvb.ensure(N * sizeof(LNode), "level nodes")) || (rc = dec.ensure(N * sizeof(NDec), "split decisions")) ||
        (rc = ntask.ensure((N + 1) * 4, "binning tasks")) || (rc = isbig.ensure((N + 1) * 4, "big nodes")) ||
        (rc = toff.ensure((N + 1) * 4, "task offsets")) || (rc = brank.ensure((N + 1) * 4, "big ranks")) ||
        (rc = flags.ensure((N + 1) * 4, "split flags")) || (rc = srank.ensure((N + 1) * 4, "split ranks")) ||
        (rc = tleft.ensure((N / kChunk + N / kSmall + 2) * 4, "chunk counts")) ||
        (rc = order.ensure(N * 4, "leaf order")) || (rc = last.ensure(N, "leaf ends")) ||
        (rc = red.ensure(64 * 4, "scene bounds")) ||
        (rc = c->nodes.ensure((N > 1 ? N - 1 : 1) * 64, "BVH nodes")) || (rc = c->tris.ensure(N * 48, "triangles")))
        return rc;
    if (ids && (rc = dids.ensure(N * 4, "triangle ids"))) return rc;
    VH_HIP(hipMemcpyAsync(dv.ptr, v9, N * 9 * 4, hipMemcpyHostToDevice, s));
    if (ids) VH_HIP(hipMemcpyAsync(dids.ptr, ids, N * 4, hipMemcpyHostToDevice, s));
    int32_t *rd = (int32_t *)red.ptr;  // [0,12) root box and centroid bounds, [12] max |coord|,
                                       // [13] first non-finite triangle, [14] unplaced triangles
    {
        int32_t init[15];
        for (int k = 0; k < 12; k++) {
            const bool lo = (k % 6) < 3;
            const float f = lo ? INFINITY : -INFINITY;
            const int32_t bits = *(const int32_t *)&f;
            init[k] = bits >= 0 ? bits : bits ^ 0x7FFFFFFF;
        }
        init[12] = 0;
        init[13] = INT32_MAX;
        init[14] = 0;
        VH_HIP(hipMemcpyAsync(rd, init, sizeof init, hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(grid(n)), dim3(256), 0, s, (int32_t *)order.ptr, (int64_t)n, -1);
    hipLaunchKernelGGL(k_prims, dim3(grid(n)), dim3(256), 0, s, (const float *)dv.ptr, n, (Prim *)pa.ptr, rd, rd + 12,
                       rd + 13);
    VH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, s, (const int32_t *)rd, n, (LNode *)lva.ptr);
    int32_t host_red[14];
    VH_HIP(hipMemcpyAsync(host_red, rd, sizeof host_red, hipMemcpyDeviceToHost, s));
    VH_WAIT(c, s);
    if (host_red[13] != INT32_MAX) {
        set_error("triangle " + std::to_string(host_red[13]) + " has a non-finite vertex");
        return VH_EINVAL;
    }
    Prim *src = (Prim *)pa.ptr, *dst = (Prim *)pb.ptr;
    LNode *cur = (LNode *)lva.ptr, *nxt = (LNode *)lvb.ptr;
    int32_t na = 1, gbase = 0, depth = 0;
    size_t gbin_cap = 0;
    for (int level = 0; na > 0; level++) {
        depth = level;
        // the level's big nodes and their chunks
        hipLaunchKernelGGL(k_level_counts, dim3(grid(na + 1)), dim3(256), 0, s, cur, na, (int32_t *)ntask.ptr,
                           (int32_t *)isbig.ptr);
        if ((rc = exclusive_scan(c, tmp, (const int32_t *)ntask.ptr, (int32_t *)toff.ptr, na, s)) ||
            (rc = exclusive_scan(c, tmp, (const int32_t *)isbig.ptr, (int32_t *)brank.ptr, na, s)))
            return rc;
        int32_t tot[2];
        VH_HIP(hipMemcpyAsync(&tot[0], (int32_t *)toff.ptr + na, 4, hipMemcpyDeviceToHost, s));
        VH_HIP(hipMemcpyAsync(&tot[1], (int32_t *)brank.ptr + na, 4, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
        const int32_t ntasks = tot[0], nbig = tot[1];
        if (nbig > 0) {
            const size_t need = (size_t)nbig * kBinInts * 4;
            if (need > gbin_cap) {
                if ((rc = gbin.ensure(need, "node bins"))) return rc;
                gbin_cap = need;
            }
            hipLaunchKernelGGL(k_clear_bins, dim3(grid((int64_t)nbig * kBinInts)), dim3(256), 0, s,
                               (int32_t *)gbin.ptr, nbig);
        }
        LevelArgs la{};
        la.src = src;
        la.dst = dst;
        la.nodes = cur;
        la.na = na;
        la.dec = (NDec *)dec.ptr;
        la.task_off = (const int32_t *)toff.ptr;
        la.big_rank = (const int32_t *)brank.ptr;
        la.gbin = (int32_t *)gbin.ptr;
        la.task_left = (int32_t *)tleft.ptr;
        la.order = (int32_t *)order.ptr;
        la.last = (uint8_t *)last.ptr;
        if (ntasks > 0) hipLaunchKernelGGL(k_big_bin, dim3((uint32_t)ntasks), dim3(256), 0, s, la);
        hipLaunchKernelGGL(k_decide, dim3(grid(na, 4)), dim3(256), 0, s, la);
        if (ntasks > 0) {
            hipLaunchKernelGGL(k_big_count, dim3((uint32_t)ntasks), dim3(256), 0, s, la);
            hipLaunchKernelGGL(k_big_scan, dim3(grid(na)), dim3(256), 0, s, (const int32_t *)toff.ptr, na,
                               (const NDec *)dec.ptr, (const int32_t *)isbig.ptr, (int32_t *)tleft.ptr);
            hipLaunchKernelGGL(k_big_scatter, dim3((uint32_t)ntasks), dim3(256), 0, s, la);
        }
        VH_HIP(hipGetLastError());
        // split flags -> level-order numbering and the next level
        hipLaunchKernelGGL(k_fill_i32, dim3(grid(na + 1)), dim3(256), 0, s, (int32_t *)flags.ptr, (int64_t)na + 1, 0);
        hipLaunchKernelGGL(k_split_flags, dim3(grid(na)), dim3(256), 0, s, (const NDec *)dec.ptr, na,
                           (int32_t *)flags.ptr);
        if ((rc = exclusive_scan(c, tmp, (const int32_t *)flags.ptr, (int32_t *)srank.ptr, na, s))) return rc;
        hipLaunchKernelGGL(k_children, dim3(grid(na)), dim3(256), 0, s, (const LNode *)cur, na, (const NDec *)dec.ptr,
                           (const int32_t *)srank.ptr, gbase, (const int32_t *)(rd + 12), (float *)c->nodes.ptr, nxt);
        VH_HIP(hipGetLastError());
        int32_t nsplit = 0;
        VH_HIP(hipMemcpyAsync(&nsplit, (int32_t *)srank.ptr + na, 4, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
        gbase += nsplit;
        na = 2 * nsplit;
        std::swap(src, dst);
        std::swap(cur, nxt);
    }
    hipLaunchKernelGGL(k_tris, dim3(grid(n)), dim3(256), 0, s, (const float *)dv.ptr,
                       ids ? (const int32_t *)dids.ptr : nullptr, (const int32_t *)order.ptr,
                       (const uint8_t *)last.ptr, n, (float4 *)c->tris.ptr, rd + 14);
    VH_HIP(hipGetLastError());
    int32_t unplaced = 0;
    VH_HIP(hipMemcpyAsync(&unplaced, rd + 14, 4, hipMemcpyDeviceToHost, s));
    VH_WAIT(c, s);
    if (unplaced) {
        set_error("the device build left triangles outside every leaf");
        return VH_EDEVICE;
    }
    c->nnodes = gbase;
    c->ntri = n;
    c->depth = depth;
    for (int k = 0; k < 3; k++) {
        int32_t lo = host_red[k], hi = host_red[3 + k];
        lo = lo >= 0 ? lo : lo ^ 0x7FFFFFFF;
        hi = hi >= 0 ? hi : hi ^ 0x7FFFFFFF;
        c->lo[k] = *(const float *)&lo;
        c->hi[k] = *(const float *)&hi;
    }
    c->have_scene = true;
    std::swap(c->verts, dv);  // the caller's triangles stay on the device (dv takes the old copy to release)
    c->have_verts = true;
    c->scene_gen++;
    return VH_OK;
}

}  // namespace vh
