// Device-side traversal shared by the ray kernels and the transfer build:
// the reference KD tracer's exact triangle test over a BVH2.
//
// Numerics.  The triangle test is the reference's IntersectTriangle
// (Tracing/KdRayTracer.cs:1194-1267) operation for operation: every file that
// includes this is built with -ffp-contract=off and `fp contract(off)` below,
// so no multiply and add fuse, and the divide is IEEE (hipcc's default
// correctly rounded f32 divide).  Box tests only cull, and use explicit fmaf
// with a widened interval, so a hit the exact test accepts is never culled.
//
// The traversal stack lives in LDS, one column per lane ([depth][lane] with a
// row stride of the block size, so a wave's pushes hit 64 different banks),
// kMaxDepth entries deep, which the scene builder guarantees is enough.
#pragma once

#include <type_traits>

#include "vh_internal.h"

#pragma clang fp contract(off)

namespace vh {

constexpr float kParallelEps = 1.0e-10f;  // stock's FourZeros (KdRayTracer.cs:1211)
constexpr float kZeroSub = 8.6736174e-19f;  // 2^-60, KdRayTracer.cs:193
constexpr float kSlack = 6.1035156e-5f;     // 2^-14 relative widening of box intervals

enum Mode { kAny = 0, kClosestInclusive = 1, kClosestShort = 2 };

struct Ray {
    float ox, oy, oz, dx, dy, dz, reach;
};

__device__ __forceinline__ bool ray_finite(const Ray &r)
{
    return isfinite(r.ox) && isfinite(r.oy) && isfinite(r.oz) && isfinite(r.dx) &&
           isfinite(r.dy) && isfinite(r.dz);
}

__device__ __forceinline__ float inv_saturate(float d)
{
    // zeros become +-2^-60 with their sign kept (ReciprocalSaturateExact,
    // KdRayTracer.cs:1446-1451), and so do nonzero components smaller than
    // that, so the inverse is always finite; the box test tolerates an
    // approximate inverse.
    if (fabsf(d) < kZeroSub) d = copysignf(kZeroSub, d);
    return __builtin_amdgcn_rcpf(d);
}

__device__ __forceinline__ float sel3(int c, float x, float y, float z)
{
    return c == 0 ? x : (c == 1 ? y : z);
}

// The exact test.  Returns true with t when the ray hits the triangle.
__device__ __forceinline__ bool tri_test(const float4 t0, const float4 t1, const float4 t2,
                                         uint32_t meta, const Ray &r, float &t)
{
    float ddotn = r.dx * t0.x;
    ddotn = (r.dy * t0.y) + ddotn;
    ddotn = (r.dz * t0.z) + ddotn;
    float odotn = r.ox * t0.x;
    odotn = (r.oy * t0.y) + odotn;
    odotn = (r.oz * t0.z) + odotn;
    t = (t0.w - odotn) / ddotn;
    int cs0 = meta & 3, cs1 = (meta >> 2) & 3;
    float h1 = sel3(cs0, r.ox, r.oy, r.oz) + (t * sel3(cs0, r.dx, r.dy, r.dz));
    float h2 = sel3(cs1, r.ox, r.oy, r.oz) + (t * sel3(cs1, r.dx, r.dy, r.dz));
    float b0 = t1.x * h1;
    b0 = b0 + (t1.y * h2);
    b0 = b0 + t1.z;
    float b1 = t1.w * h1;
    b1 = b1 + (t2.x * h2);
    b1 = b1 + t2.y;
    return (ddotn > kParallelEps || ddotn < -kParallelEps) && (t > kParallelEps) &&
           (b0 >= kParallelEps) && (b1 >= kParallelEps) && ((b1 + b0) <= 1.0f);
}

// tri_test's answer with every condition evaluated (no short circuit), for
// the packet walks, where all lanes run the test together anyway: the same
// operations in the same order, so the same t and the same answer.
__device__ __forceinline__ bool tri_test_flat(const float4 t0, const float4 t1, const float4 t2,
                                              uint32_t meta, const Ray &r, float &t)
{
    float ddotn = r.dx * t0.x;
    ddotn = (r.dy * t0.y) + ddotn;
    ddotn = (r.dz * t0.z) + ddotn;
    float odotn = r.ox * t0.x;
    odotn = (r.oy * t0.y) + odotn;
    odotn = (r.oz * t0.z) + odotn;
    t = (t0.w - odotn) / ddotn;
    int cs0 = meta & 3, cs1 = (meta >> 2) & 3;
    float h1 = sel3(cs0, r.ox, r.oy, r.oz) + (t * sel3(cs0, r.dx, r.dy, r.dz));
    float h2 = sel3(cs1, r.ox, r.oy, r.oz) + (t * sel3(cs1, r.dx, r.dy, r.dz));
    float b0 = t1.x * h1;
    b0 = b0 + (t1.y * h2);
    b0 = b0 + t1.z;
    float b1 = t1.w * h1;
    b1 = b1 + (t2.x * h2);
    b1 = b1 + t2.y;
    return ((ddotn > kParallelEps) | (ddotn < -kParallelEps)) & (t > kParallelEps) &
           (b0 >= kParallelEps) & (b1 >= kParallelEps) & ((b1 + b0) <= 1.0f);
}

// tri_test_flat with the two kept axes fixed at compile time (C0, C1 are the
// triangle's cs0, cs1): the same operations on the same operands, without the
// per-lane selects.
template <int C0, int C1>
__device__ __forceinline__ bool tri_test_axes(const float4 t0, const float4 t1, const float4 t2,
                                              const Ray &r, float &t)
{
    float ddotn = r.dx * t0.x;
    ddotn = (r.dy * t0.y) + ddotn;
    ddotn = (r.dz * t0.z) + ddotn;
    float odotn = r.ox * t0.x;
    odotn = (r.oy * t0.y) + odotn;
    odotn = (r.oz * t0.z) + odotn;
    t = (t0.w - odotn) / ddotn;
    const float o0 = C0 == 0 ? r.ox : (C0 == 1 ? r.oy : r.oz);
    const float d0 = C0 == 0 ? r.dx : (C0 == 1 ? r.dy : r.dz);
    const float o1 = C1 == 0 ? r.ox : (C1 == 1 ? r.oy : r.oz);
    const float d1 = C1 == 0 ? r.dx : (C1 == 1 ? r.dy : r.dz);
    float h1 = o0 + (t * d0);
    float h2 = o1 + (t * d1);
    float b0 = t1.x * h1;
    b0 = b0 + (t1.y * h2);
    b0 = b0 + t1.z;
    float b1 = t1.w * h1;
    b1 = b1 + (t2.x * h2);
    b1 = b1 + t2.y;
    return ((ddotn > kParallelEps) | (ddotn < -kParallelEps)) & (t > kParallelEps) &
           (b0 >= kParallelEps) & (b1 >= kParallelEps) & ((b1 + b0) <= 1.0f);
}

// The exact test for a wave-uniform triangle: the scene builder only makes
// (cs0, cs1) = (1, 2), (2, 0) or (0, 1), each taken by a uniform branch.
__device__ __forceinline__ bool tri_test_uniform(const float4 t0, const float4 t1, const float4 t2,
                                                 uint32_t meta, const Ray &r, float &t)
{
    switch (meta & 15u) {
    case 1u | (2u << 2): return tri_test_axes<1, 2>(t0, t1, t2, r, t);
    case 2u | (0u << 2): return tri_test_axes<2, 0>(t0, t1, t2, r, t);
    case 0u | (1u << 2): return tri_test_axes<0, 1>(t0, t1, t2, r, t);
    default: return tri_test_flat(t0, t1, t2, meta, r, t);
    }
}

// One ray through the BVH.  MODE kAny returns at the first hit strictly short
// of the reach; the closest modes keep the nearest (ties to the lower
// original triangle index).  Returns the original triangle index or -1.
// STRIDE is the LDS stack's row stride: the block size of the caller.
//
// The walk is "while-while": a lane descends through internal nodes until it
// reaches a leaf or runs out of nodes, and only then tests the leaf's
// triangles, so the lanes of a wave test leaves together instead of one
// lane's leaf stalling the others' node steps.
template <int MODE, bool COUNT, int STRIDE>
__device__ __forceinline__ int32_t trace_one(const float4 *__restrict__ nodes,
                                             const float4 *__restrict__ tris, int *stk,
                                             const Ray &r, float tmin, int32_t skip_id,
                                             float &t_out, int32_t &id_out,
                                             uint32_t *nvisit = nullptr, uint32_t *ntest = nullptr)
{
    float ix = inv_saturate(r.dx), iy = inv_saturate(r.dy), iz = inv_saturate(r.dz);
    float nox = -(r.ox * ix), noy = -(r.oy * iy), noz = -(r.oz * iz);
    float tlo = tmin > 0.0f ? tmin * (1.0f - kSlack) : 0.0f;
    float best_t = r.reach;
    int32_t best_orig = -1, best_id = -1;
    int sp = 0;
    int32_t ref = 0;
    for (;;) {
        while (ref >= 0) {
            if (COUNT) ++*nvisit;
            const float4 *nd = nodes + (size_t)ref * 4;
            float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3];
            float thi = MODE == kAny ? r.reach : best_t;
            float ax0 = __builtin_fmaf(q0.x, ix, nox), ax1 = __builtin_fmaf(q0.y, ix, nox);
            float ay0 = __builtin_fmaf(q0.z, iy, noy), ay1 = __builtin_fmaf(q0.w, iy, noy);
            float az0 = __builtin_fmaf(q2.x, iz, noz), az1 = __builtin_fmaf(q2.y, iz, noz);
            float bx0 = __builtin_fmaf(q1.x, ix, nox), bx1 = __builtin_fmaf(q1.y, ix, nox);
            float by0 = __builtin_fmaf(q1.z, iy, noy), by1 = __builtin_fmaf(q1.w, iy, noy);
            float bz0 = __builtin_fmaf(q2.z, iz, noz), bz1 = __builtin_fmaf(q2.w, iz, noz);
            float tn0 = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fmaxf(fminf(az0, az1), tlo));
            float tf0 = fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fminf(fmaxf(az0, az1), thi));
            float tn1 = fmaxf(fmaxf(fminf(bx0, bx1), fminf(by0, by1)), fmaxf(fminf(bz0, bz1), tlo));
            float tf1 = fminf(fminf(fmaxf(bx0, bx1), fmaxf(by0, by1)), fminf(fmaxf(bz0, bz1), thi));
            int32_t r0 = __float_as_int(q3.x), r1 = __float_as_int(q3.y);
            bool h0 = tn0 <= __builtin_fmaf(fabsf(tf0), kSlack, tf0) && r0 != kEmptyRef;
            bool h1 = tn1 <= __builtin_fmaf(fabsf(tf1), kSlack, tf1) && r1 != kEmptyRef;
            if (h0 && h1) {
                bool swap = tn1 < tn0;
                stk[sp * STRIDE] = swap ? r0 : r1;
                sp++;
                ref = swap ? r1 : r0;
                continue;
            }
            if (h0 | h1) {
                ref = h0 ? r0 : r1;
                continue;
            }
            if (sp == 0) {
                ref = kEmptyRef;  // done: no leaf to test
                break;
            }
            sp--;
            ref = stk[sp * STRIDE];
        }
        if (ref == kEmptyRef) break;
        {
            int32_t k = ~ref;
            for (;;) {
                if (COUNT) ++*ntest;
                const float4 *tp = tris + (size_t)k * 3;
                float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
                uint32_t meta = __float_as_uint(t2.w);
                int32_t id = __float_as_int(t2.z);
                float t;
                bool hit = tri_test(t0, t1, t2, meta, r, t) && t >= tmin &&
                           !(skip_id != -1 && id == skip_id);
                if (hit) {
                    int32_t orig = (int32_t)(meta >> kOrigShift);
                    if (MODE == kAny) {
                        if (t < r.reach) {
                            t_out = t;
                            id_out = id;
                            return orig;
                        }
                    } else {
                        bool within = MODE == kClosestInclusive ? (t <= r.reach) : (t < r.reach);
                        bool better = best_orig < 0
                                          ? within
                                          : (t < best_t || (t == best_t && orig < best_orig));
                        if (better) {
                            best_t = t;
                            best_orig = orig;
                            best_id = id;
                        }
                    }
                }
                if (meta & kTriLast) break;
                k++;
            }
        }
        if (sp == 0) break;
        sp--;
        ref = stk[sp * STRIDE];
    }
    t_out = best_t;
    id_out = best_id;
    return best_orig;
}

// The wave's lane index and the wave-uniform form of a value every lane holds.
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ int32_t uniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Both children's parametric spans [tn, tf] for one lane's ray, from a node's
// boxes: q0 = child 0's x and y slabs, q1 = child 1's, q2 = both z slabs.
// OCT < 0 takes each slab's near and far plane by min/max.  OCT 0..7 is the
// direction octant every lane of the wave shares (bit 0 set: x component
// negative, bit 1: y, bit 2: z); the planes are then picked at compile time,
// which leaves 12 of the general form's 24 min/max.  fmaf(q, ix, nox) is
// monotonic in q for a finite inverse of fixed sign, so with lo <= hi the
// picked plane is the one min/max would take and every span, and so every
// lane mask, is bit-identical to the general form's.  (An empty child's box
// is inverted, lo = +inf > hi = -inf, so its spans differ; its mask is
// cleared by its empty ref in both forms.)
template <int OCT>
__device__ __forceinline__ void child_spans(const float4 q0, const float4 q1, const float4 q2,
                                            float ix, float iy, float iz, float nox, float noy,
                                            float noz, float tlo, float thi, float &tn0,
                                            float &tf0, float &tn1, float &tf1)
{
    if constexpr (OCT < 0) {
        const float ax0 = __builtin_fmaf(q0.x, ix, nox), ax1 = __builtin_fmaf(q0.y, ix, nox);
        const float ay0 = __builtin_fmaf(q0.z, iy, noy), ay1 = __builtin_fmaf(q0.w, iy, noy);
        const float az0 = __builtin_fmaf(q2.x, iz, noz), az1 = __builtin_fmaf(q2.y, iz, noz);
        const float bx0 = __builtin_fmaf(q1.x, ix, nox), bx1 = __builtin_fmaf(q1.y, ix, nox);
        const float by0 = __builtin_fmaf(q1.z, iy, noy), by1 = __builtin_fmaf(q1.w, iy, noy);
        const float bz0 = __builtin_fmaf(q2.z, iz, noz), bz1 = __builtin_fmaf(q2.w, iz, noz);
        tn0 = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fmaxf(fminf(az0, az1), tlo));
        tf0 = fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fminf(fmaxf(az0, az1), thi));
        tn1 = fmaxf(fmaxf(fminf(bx0, bx1), fminf(by0, by1)), fmaxf(fminf(bz0, bz1), tlo));
        tf1 = fminf(fminf(fmaxf(bx0, bx1), fmaxf(by0, by1)), fminf(fmaxf(bz0, bz1), thi));
    } else {
        constexpr bool nx = (OCT & 1) != 0, ny = (OCT & 2) != 0, nz = (OCT & 4) != 0;
        const float ax_n = __builtin_fmaf(nx ? q0.y : q0.x, ix, nox);
        const float ax_f = __builtin_fmaf(nx ? q0.x : q0.y, ix, nox);
        const float ay_n = __builtin_fmaf(ny ? q0.w : q0.z, iy, noy);
        const float ay_f = __builtin_fmaf(ny ? q0.z : q0.w, iy, noy);
        const float az_n = __builtin_fmaf(nz ? q2.y : q2.x, iz, noz);
        const float az_f = __builtin_fmaf(nz ? q2.x : q2.y, iz, noz);
        const float bx_n = __builtin_fmaf(nx ? q1.y : q1.x, ix, nox);
        const float bx_f = __builtin_fmaf(nx ? q1.x : q1.y, ix, nox);
        const float by_n = __builtin_fmaf(ny ? q1.w : q1.z, iy, noy);
        const float by_f = __builtin_fmaf(ny ? q1.z : q1.w, iy, noy);
        const float bz_n = __builtin_fmaf(nz ? q2.w : q2.z, iz, noz);
        const float bz_f = __builtin_fmaf(nz ? q2.z : q2.w, iz, noz);
        tn0 = fmaxf(fmaxf(ax_n, ay_n), fmaxf(az_n, tlo));
        tf0 = fminf(fminf(ax_f, ay_f), fminf(az_f, thi));
        tn1 = fmaxf(fmaxf(bx_n, by_n), fmaxf(bz_n, tlo));
        tf1 = fminf(fminf(bx_f, by_f), fminf(bz_f, thi));
    }
}

// The direction octant shared by every lane with `want` set (the OCT of
// child_spans), or -1 when the lanes' signs differ on some axis.  The sign of
// a component is that of its inverse (inv_saturate keeps it).  It is also -1
// when some lane's ray is not tame: a direction component of 2^100 or more
// (whose inverse could be zero) or an origin component of 2^60 or more (whose
// product with an inverse up to 2^60 could overflow).  For a tame ray an empty
// child's inverted box gives tn = +inf and a NaN widened tf, so it fails the
// box test by itself and the octant walks skip the empty-ref check.
__device__ __forceinline__ int wave_octant(bool want, const Ray &r)
{
    const float dmax = fmaxf(fmaxf(fabsf(r.dx), fabsf(r.dy)), fabsf(r.dz));
    const float omax = fmaxf(fmaxf(fabsf(r.ox), fabsf(r.oy)), fabsf(r.oz));
    const bool tame = dmax < 0x1p100f && omax < 0x1p60f;
    if (__builtin_amdgcn_ballot_w64(want && !tame)) return -1;
    const float dx = r.dx, dy = r.dy, dz = r.dz;
    const uint64_t w = __builtin_amdgcn_ballot_w64(want);
    const uint64_t sx = __builtin_amdgcn_ballot_w64(want && signbit(dx));
    const uint64_t sy = __builtin_amdgcn_ballot_w64(want && signbit(dy));
    const uint64_t sz = __builtin_amdgcn_ballot_w64(want && signbit(dz));
    if ((sx != 0 && sx != w) || (sy != 0 && sy != w) || (sz != 0 && sz != w)) return -1;
    return uniform((sx != 0 ? 1 : 0) | (sy != 0 ? 2 : 0) | (sz != 0 ? 4 : 0));
}

// Calls f(std::integral_constant<int, OCT>) with the octant as a compile-time
// constant (-1 for mixed signs), so each octant gets its own walk.
template <typename F>
__device__ __forceinline__ auto by_octant(int oct, F &&f)
{
    switch (oct) {
    case 0: return f(std::integral_constant<int, 0>{});
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 6: return f(std::integral_constant<int, 6>{});
    case 7: return f(std::integral_constant<int, 7>{});
    default: return f(std::integral_constant<int, -1>{});
    }
}

// Any hit for the 64 rays of a wave that travel together (one light from
// nearby starts, one sky or sun direction, one AO direction of one face).
// Every lane of the wave calls it at once; `want` is false for a lane with no
// ray.  The wave walks one node sequence: each node is tested by every lane
// whose ray entered its parent's box (the node's lane mask); a child is
// visited when any of those lanes' rays enters it, with the lanes that do; a
// lane leaves the walk when its ray is blocked.  A lane therefore tests
// exactly the nodes and triangles its own traversal would (up to when it
// stops), so the answer is the one-ray answer, while the control flow is the
// wave's, not the lane's: no divergence, node fetches shared by the wave.
// The stack is the wave's (`wstk`, 3 ints a level in LDS: node, lane mask).
// The near child is, for an octant walk, the one the shared direction meets
// first along the node's separating axis, else the one the mask's first lane
// enters first.  The loop is "descend, then leaf, then pop", like trace_one's
// while-while walk: the inner descent has one exit, which keeps the scalar
// control flow of a visit short.
template <bool COUNT, int OCT = -1>
__device__ __forceinline__ bool packet_any(const float4 *__restrict__ nodes,
                                           const float4 *__restrict__ tris, int *wstk, const Ray &r,
                                           float tmin, bool want, uint32_t *nvisit = nullptr,
                                           uint32_t *ntest = nullptr, int32_t *blocker = nullptr,
                                           uint32_t *slots = nullptr)
{
    const int lane = lane_id();
    const float ix = inv_saturate(r.dx), iy = inv_saturate(r.dy), iz = inv_saturate(r.dz);
    const float nox = -(r.ox * ix), noy = -(r.oy * iy), noz = -(r.oz * iz);
    const float tlo = fmaxf(tmin * (1.0f - kSlack), 0.0f);  // 0 for tmin <= 0 or NaN
    const float thi = r.reach;
    uint64_t mask = uniform64(__builtin_amdgcn_ballot_w64(want));
    uint64_t done = 0;  // lanes already blocked
    // Per-lane state is kept in integers and every lane runs every test, with
    // the lane masks applied arithmetically: a wave's VALU instruction costs
    // the same whatever lanes are on, and divergent branches would cost
    // scalar exec-mask bookkeeping on every visit.
    uint32_t blk = 0;  // this lane's ray is blocked
    int sp = 0;
    int32_t ref = 0;
    if (mask == 0) return false;
    for (;;) {
        // descend through internal nodes until a leaf, or until no child is entered
        while (ref >= 0) {
            if (COUNT) {
                *nvisit += (uint32_t)(mask >> lane) & 1u;
                if (slots) ++slots[0];
            }
            const float4 *nd = nodes + (size_t)ref * 4;
            const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3];
            float tn0, tf0, tn1, tf1;
            child_spans<OCT>(q0, q1, q2, ix, iy, iz, nox, noy, noz, tlo, thi, tn0, tf0, tn1, tf1);
            const int32_t r0 = uniform(__float_as_int(q3.x)), r1 = uniform(__float_as_int(q3.y));
            const uint64_t m0 = uniform64(__builtin_amdgcn_ballot_w64(
                                    tn0 <= __builtin_fmaf(fabsf(tf0), kSlack, tf0))) &
                                mask & (OCT >= 0 || r0 != kEmptyRef ? ~0ull : 0ull);
            const uint64_t m1 = uniform64(__builtin_amdgcn_ballot_w64(
                                    tn1 <= __builtin_fmaf(fabsf(tf1), kSlack, tf1))) &
                                mask & (OCT >= 0 || r1 != kEmptyRef ? ~0ull : 0ull);
            if (m0 != 0) {
                if (m1 != 0) {
                    bool swap;
                    if constexpr (OCT >= 0) {
                        const int32_t ord = uniform(__float_as_int(q3.z));
                        swap = (((OCT >> (ord & 3)) ^ (ord >> 2)) & 1) != 0;
                    } else {
                        const uint64_t both = m0 & m1;
                        const int fl = __builtin_ctzll(both != 0 ? both : mask);
                        const float n0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tn0), fl));
                        const float n1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tn1), fl));
                        swap = n1 < n0;
                    }
                    const int32_t far = swap ? r0 : r1;
                    const uint64_t fm = swap ? m0 : m1;
                    wstk[3 * sp] = far;
                    wstk[3 * sp + 1] = (int32_t)(uint32_t)fm;
                    wstk[3 * sp + 2] = (int32_t)(uint32_t)(fm >> 32);
                    sp++;
                    ref = swap ? r1 : r0;
                    mask = swap ? m1 : m0;
                } else {
                    ref = r0;
                    mask = m0;
                }
            } else if (m1 != 0) {
                ref = r1;
                mask = m1;
            } else {
                ref = kEmptyRef;
            }
        }
        if (ref != kEmptyRef) {
            const uint32_t act = (uint32_t)(mask >> lane) & 1u;
            for (int32_t k = ~ref;; k++) {
                const float4 *tp = tris + (size_t)k * 3;
                const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
                const uint32_t meta = (uint32_t)uniform(__float_as_int(t2.w));
                const uint32_t live = act & (blk ^ 1u);
                if (COUNT) {
                    if (slots) ++slots[1];
                    *ntest += live;
                }
                float t;
                const bool hit = tri_test_uniform(t0, t1, t2, meta, r, t) & (t >= tmin) & (t < r.reach);
                const uint32_t take = live & (uint32_t)hit;
                if (blocker) *blocker = take ? k : *blocker;
                blk |= take;
                if (meta & kTriLast) break;
            }
            done = uniform64(__builtin_amdgcn_ballot_w64(blk != 0));
        }
        mask = 0;
        while (sp > 0) {
            sp--;
            ref = uniform(wstk[3 * sp]);
            const uint64_t lo = (uint32_t)uniform(wstk[3 * sp + 1]);
            const uint64_t hi = (uint32_t)uniform(wstk[3 * sp + 2]);
            mask = ((hi << 32) | lo) & ~done;
            if (mask != 0) break;
        }
        if (mask == 0) break;
    }
    return blk != 0;
}

// Whether triangle k (leaf order) blocks the ray strictly short of its reach:
// the exact test for one remembered occluder.
__device__ __forceinline__ bool occluder_blocks(const float4 *__restrict__ tris, int32_t k, const Ray &r,
                                                float tmin)
{
    const float4 *tp = tris + (size_t)k * 3;
    const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
    float t;
    return tri_test(t0, t1, t2, __float_as_uint(t2.w), r, t) && t >= tmin && t < r.reach;
}

// The closest hit (MODE kClosestShort: t < reach; kClosestInclusive:
// t <= reach) for the 64 rays of a wave that travel together, walked as in
// packet_any.  A lane's box tests are bounded by its own best hit so far, so
// masks shrink as lanes find hits; ties go to the lower original triangle
// index, so the answer is the one-ray answer.  Returns the lane's original
// triangle index or -1, with its distance and caller id.
template <int MODE, int OCT = -1, bool COUNT = false>
__device__ __forceinline__ int32_t packet_closest(const float4 *__restrict__ nodes,
                                                  const float4 *__restrict__ tris, int *wstk,
                                                  const Ray &r, float tmin, bool want, float &t_out,
                                                  int32_t &id_out, uint32_t *nvisit = nullptr,
                                                  uint32_t *ntest = nullptr, uint32_t *slots = nullptr)
{
    const int lane = lane_id();
    const float ix = inv_saturate(r.dx), iy = inv_saturate(r.dy), iz = inv_saturate(r.dz);
    const float nox = -(r.ox * ix), noy = -(r.oy * iy), noz = -(r.oz * iz);
    const float tlo = fmaxf(tmin * (1.0f - kSlack), 0.0f);  // 0 for tmin <= 0 or NaN
    float best_t = r.reach;
    int32_t best_orig = -1, best_id = -1;
    uint64_t mask = uniform64(__builtin_amdgcn_ballot_w64(want));
    int sp = 0;
    int32_t ref = 0;
    if (mask == 0) {
        t_out = best_t;
        id_out = best_id;
        return best_orig;
    }
    for (;;) {
        // descend (see packet_any)
        while (ref >= 0) {
            if (COUNT) {
                *nvisit += (uint32_t)(mask >> lane) & 1u;
                ++slots[0];
            }
            const float4 *nd = nodes + (size_t)ref * 4;
            const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3];
            const float thi = best_t;
            float tn0, tf0, tn1, tf1;
            child_spans<OCT>(q0, q1, q2, ix, iy, iz, nox, noy, noz, tlo, thi, tn0, tf0, tn1, tf1);
            const int32_t r0 = uniform(__float_as_int(q3.x)), r1 = uniform(__float_as_int(q3.y));
            // a child's lanes: the node's lanes whose rays enter its box
            const uint64_t m0 = uniform64(__builtin_amdgcn_ballot_w64(
                                    tn0 <= __builtin_fmaf(fabsf(tf0), kSlack, tf0))) &
                                mask & (OCT >= 0 || r0 != kEmptyRef ? ~0ull : 0ull);
            const uint64_t m1 = uniform64(__builtin_amdgcn_ballot_w64(
                                    tn1 <= __builtin_fmaf(fabsf(tf1), kSlack, tf1))) &
                                mask & (OCT >= 0 || r1 != kEmptyRef ? ~0ull : 0ull);
            if (m0 != 0) {
                if (m1 != 0) {
                    const uint64_t both = m0 & m1;
                    const int fl = __builtin_ctzll(both != 0 ? both : mask);
                    const float n0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tn0), fl));
                    const float n1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tn1), fl));
                    const bool swap = n1 < n0;
                    const int32_t far = swap ? r0 : r1;
                    const uint64_t fm = swap ? m0 : m1;
                    wstk[3 * sp] = far;
                    wstk[3 * sp + 1] = (int32_t)(uint32_t)fm;
                    wstk[3 * sp + 2] = (int32_t)(uint32_t)(fm >> 32);
                    sp++;
                    ref = swap ? r1 : r0;
                    mask = swap ? m1 : m0;
                } else {
                    ref = r0;
                    mask = m0;
                }
            } else if (m1 != 0) {
                ref = r1;
                mask = m1;
            } else {
                ref = kEmptyRef;
            }
        }
        if (ref != kEmptyRef) {
            // every lane runs the test; the mask and the order rule are
            // applied by selects (see packet_any)
            const bool act = ((mask >> lane) & 1ull) != 0;
            for (int32_t k = ~ref;; k++) {
                const float4 *tp = tris + (size_t)k * 3;
                const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
                const uint32_t meta = (uint32_t)uniform(__float_as_int(t2.w));
                const int32_t orig = (int32_t)(meta >> kOrigShift);
                if (COUNT) {
                    *ntest += act ? 1u : 0u;
                    ++slots[1];
                }
                float t;
                const bool h = tri_test_uniform(t0, t1, t2, meta, r, t) & (t >= tmin);
                const bool within = MODE == kClosestInclusive ? (t <= r.reach) : (t < r.reach);
                const bool better = best_orig < 0
                                        ? within
                                        : ((t < best_t) | ((t == best_t) & (orig < best_orig)));
                const bool take = act & h & better;
                best_t = take ? t : best_t;
                best_orig = take ? orig : best_orig;
                best_id = take ? __float_as_int(t2.z) : best_id;
                if (meta & kTriLast) break;
            }
        }
        mask = 0;
        while (sp > 0) {
            sp--;
            ref = uniform(wstk[3 * sp]);
            const uint64_t lo = (uint32_t)uniform(wstk[3 * sp + 1]);
            const uint64_t hi = (uint32_t)uniform(wstk[3 * sp + 2]);
            mask = (hi << 32) | lo;
            if (mask != 0) break;
        }
        if (mask == 0) break;
    }
    t_out = best_t;
    id_out = best_id;
    return best_orig;
}

}  // namespace vh
