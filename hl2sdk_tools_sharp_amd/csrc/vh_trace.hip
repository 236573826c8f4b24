// Ray traversal kernels for gfx950: BVH2 any-hit and closest-hit, with the
// reference KD tracer's exact triangle test.
//
// Numerics.  The triangle test is the reference's IntersectTriangle
// (Tracing/KdRayTracer.cs:1194-1267) operation for operation: this file is
// built with -ffp-contract=off and `fp contract(off)` below, so no multiply
// and add fuse, and the divide is IEEE (hipcc's default correctly rounded
// f32 divide).  Box tests only cull, and use explicit fmaf with a widened
// interval, so a hit the exact test accepts is never culled.
//
// Layout.  One lane traces one ray.  The traversal stack lives in LDS, one
// column per lane ([depth][lane], so a wave's pushes hit 64 different banks),
// kMaxDepth entries deep, which the scene builder guarantees is enough.
#include <hipcub/hipcub.hpp>

#include "vh_internal.h"

#pragma clang fp contract(off)

namespace vh {
namespace {

constexpr int kBlock = 128;
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
    // KdRayTracer.cs:1446-1451); the box test tolerates an approximate inverse.
    if (d == 0.0f) d = copysignf(kZeroSub, d);
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

// One ray through the BVH.  MODE kAny returns at the first hit strictly short
// of the reach; the closest modes keep the nearest (ties to the lower
// original triangle index).  Returns the original triangle index or -1.
template <int MODE, bool COUNT = false>
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
        if (ref >= 0) {
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
                stk[sp * kBlock] = swap ? r0 : r1;
                sp++;
                ref = swap ? r1 : r0;
                continue;
            }
            if (h0 | h1) {
                ref = h0 ? r0 : r1;
                continue;
            }
        } else {
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
        ref = stk[sp * kBlock];
    }
    t_out = best_t;
    id_out = best_id;
    return best_orig;
}

__device__ __forceinline__ Ray load_ray(const float *__restrict__ rays, int rw, float shared_reach,
                                        int64_t i)
{
    const float *p = rays + i * rw;
    Ray r;
    r.ox = p[0];
    r.oy = p[1];
    r.oz = p[2];
    r.dx = p[3];
    r.dy = p[4];
    r.dz = p[5];
    r.reach = rw == 7 ? p[6] : shared_reach;
    return r;
}

// Visibility over caller rays: one ballot per wave is exactly one LSB-first
// word of the seam's hit bits (Tracing/IRayTracer.cs:150-153).
template <bool SKY>
__global__ __launch_bounds__(kBlock) void k_visibility(const float4 *__restrict__ nodes,
                                                       const float4 *__restrict__ tris,
                                                       const float *__restrict__ rays, int rw,
                                                       float shared_reach, int64_t n, float tmin,
                                                       int32_t skip_id,
                                                       uint64_t *__restrict__ bits)
{
    __shared__ int stack_mem[kMaxDepth * kBlock];
    int *stk = stack_mem + threadIdx.x;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool blocked = false;
    if (i < n) {
        Ray r = load_ray(rays, rw, shared_reach, i);
        if (ray_finite(r)) {
            float t;
            int32_t id;
            if (SKY) {
                int32_t h = trace_one<kClosestShort>(nodes, tris, stk, r, tmin, skip_id, t, id);
                blocked = h >= 0 && (id & VH_TRACE_ID_SKY) == 0;
            } else {
                blocked = trace_one<kAny>(nodes, tris, stk, r, tmin, skip_id, t, id) >= 0;
            }
        }
    }
    uint64_t word = __ballot(blocked);
    int64_t first = i - (threadIdx.x & 63);
    if ((threadIdx.x & 63) == 0 && first < n) bits[first >> 6] = word;
}

__global__ __launch_bounds__(kBlock) void k_closest(const float4 *__restrict__ nodes,
                                                    const float4 *__restrict__ tris,
                                                    const float *__restrict__ rays, int rw,
                                                    float shared_reach, int64_t n, float tmin,
                                                    int32_t skip_id, int32_t *__restrict__ surface,
                                                    float *__restrict__ fraction)
{
    __shared__ int stack_mem[kMaxDepth * kBlock];
    int *stk = stack_mem + threadIdx.x;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Ray r = load_ray(rays, rw, shared_reach, i);
    int32_t sid = -1;
    float frac = 1.0f;
    if (ray_finite(r)) {
        float t;
        int32_t id;
        int32_t h = trace_one<kClosestInclusive>(nodes, tris, stk, r, tmin, skip_id, t, id);
        if (h >= 0) {
            sid = id;
            // SlabBatcher's fraction rule: the divide by a reach of 1 is skipped
            // (MapTools.Gpu/SlabBatcher.cs:1369).
            frac = r.reach == 1.0f ? t : t / r.reach;
        }
    }
    surface[i] = sid;
    fraction[i] = frac;
}

// Morton key of each point inside the scene box (10 bits an axis), for the
// spatial sort that makes a wave's 64 shadow rays start close together.
__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton(const float *__restrict__ pts, int64_t n, float lx, float ly, float lz,
                         float sx, float sy, float sz, uint32_t *__restrict__ keys,
                         int32_t *__restrict__ vals)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = (pts[i * 3 + 0] - lx) * sx, y = (pts[i * 3 + 1] - ly) * sy,
          z = (pts[i * 3 + 2] - lz) * sz;
    uint32_t qx = (uint32_t)fminf(fmaxf(x, 0.0f), 1023.0f);
    uint32_t qy = (uint32_t)fminf(fmaxf(y, 0.0f), 1023.0f);
    uint32_t qz = (uint32_t)fminf(fmaxf(z, 0.0f), 1023.0f);
    keys[i] = spread10(qx) | (spread10(qy) << 1) | (spread10(qz) << 2);
    vals[i] = (int32_t)i;
}

// Adds a per-lane count into counters[slot] once per wave.
__device__ __forceinline__ void wave_add(unsigned long long *counters, int slot, uint32_t v)
{
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&counters[slot], x);
}

// Shadow rays from points to point lights.  One lane owns one point (visited
// in Morton order) and walks every light; all lanes of a wave trace the same
// light at once, so the wave's 64 rays converge on one light from nearby
// starts.  The lane assembles its own hit word: bit l of word (p, l/64).
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_point_lights(
    const float4 *__restrict__ nodes, const float4 *__restrict__ tris,
    const float *__restrict__ pts, const int32_t *__restrict__ order, int64_t npoints,
    const float *__restrict__ lights, int32_t nlights, float tmin, uint64_t *__restrict__ bits,
    unsigned long long *__restrict__ counters)
{
    __shared__ int stack_mem[kMaxDepth * kBlock];
    int *stk = stack_mem + threadIdx.x;
    int64_t slot = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t nvisit = 0, ntest = 0;
    if (slot < npoints) {
        int64_t p = order ? order[slot] : slot;
        float px = pts[p * 3 + 0], py = pts[p * 3 + 1], pz = pts[p * 3 + 2];
        bool whole_words = (nlights & 63) == 0;
        for (int32_t g = 0; g < nlights; g += 64) {
            int32_t cnt = nlights - g < 64 ? nlights - g : 64;
            uint64_t w = 0;
            for (int32_t l = 0; l < cnt; l++) {
                const float *L = lights + (size_t)(g + l) * 3;
                Ray r;
                r.ox = px;
                r.oy = py;
                r.oz = pz;
                r.dx = L[0] - px;
                r.dy = L[1] - py;
                r.dz = L[2] - pz;
                r.reach = 1.0f;
                bool blocked = false;
                if (ray_finite(r)) {
                    float t;
                    int32_t id;
                    blocked = trace_one<kAny, COUNT>(nodes, tris, stk, r, tmin, -1, t, id, &nvisit,
                                                     &ntest) >= 0;
                }
                w |= (uint64_t)blocked << l;
            }
            if (whole_words) {
                bits[p * (nlights >> 6) + (g >> 6)] = w;
            } else {
                for (int32_t l = 0; l < cnt; l++) {
                    if ((w >> l) & 1) {
                        int64_t rix = p * nlights + g + l;
                        atomicOr((unsigned long long *)&bits[rix >> 6], 1ull << (rix & 63));
                    }
                }
            }
        }
    }
    // every lane of the wave reaches the counter reduction together
    if (COUNT) {
        wave_add(counters, 0, nvisit);
        wave_add(counters, 1, ntest);
    }
}

}  // namespace

int launch_visibility(vh_ctx *c, const float *d_rays, int32_t rw, float shared_reach, int64_t n,
                      float tmin, int32_t skip_id, int32_t sky_passes, uint64_t *d_bits,
                      hipStream_t s)
{
    if (n <= 0) return VH_OK;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    const float4 *nodes = (const float4 *)c->nodes.ptr;
    const float4 *tris = (const float4 *)c->tris.ptr;
    KernelTimer timer(c, VH_KERNEL_VISIBILITY, s);
    if (sky_passes)
        hipLaunchKernelGGL(k_visibility<true>, dim3((uint32_t)blocks), dim3(kBlock), 0, s, nodes,
                           tris, d_rays, rw, shared_reach, n, tmin, skip_id, d_bits);
    else
        hipLaunchKernelGGL(k_visibility<false>, dim3((uint32_t)blocks), dim3(kBlock), 0, s, nodes,
                           tris, d_rays, rw, shared_reach, n, tmin, skip_id, d_bits);
    timer.done();
    VH_HIP(hipGetLastError());
    return VH_OK;
}

int launch_closest(vh_ctx *c, const float *d_rays, int32_t rw, float shared_reach, int64_t n,
                   float tmin, int32_t skip_id, int32_t *d_surface, float *d_fraction,
                   hipStream_t s)
{
    if (n <= 0) return VH_OK;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    KernelTimer timer(c, VH_KERNEL_CLOSEST, s);
    hipLaunchKernelGGL(k_closest, dim3((uint32_t)blocks), dim3(kBlock), 0, s,
                       (const float4 *)c->nodes.ptr, (const float4 *)c->tris.ptr, d_rays, rw,
                       shared_reach, n, tmin, skip_id, d_surface, d_fraction);
    timer.done();
    VH_HIP(hipGetLastError());
    return VH_OK;
}

int launch_point_lights(vh_ctx *c, const float *d_points, int64_t npoints, const float *d_lights,
                        int32_t nlights, float tmin, uint64_t *d_bits, hipStream_t s)
{
    if (npoints <= 0 || nlights <= 0) return VH_OK;
    if (npoints > INT32_MAX) {
        set_error("at most 2^31-1 points per call");
        return VH_EINVAL;
    }
    int rc;
    size_t n = (size_t)npoints;
    if ((rc = c->keys_a.ensure(n * 4, "morton keys")) || (rc = c->keys_b.ensure(n * 4, "morton keys")) ||
        (rc = c->perm_a.ensure(n * 4, "point order")) || (rc = c->perm_b.ensure(n * 4, "point order")))
        return rc;
    float ext[3], sc[3];
    for (int k = 0; k < 3; k++) {
        ext[k] = c->hi[k] - c->lo[k];
        sc[k] = ext[k] > 0 ? 1024.0f / ext[k] : 0.0f;
    }
    hipLaunchKernelGGL(k_morton, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, d_points,
                       npoints, c->lo[0], c->lo[1], c->lo[2], sc[0], sc[1], sc[2],
                       (uint32_t *)c->keys_a.ptr, (int32_t *)c->perm_a.ptr);
    VH_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    VH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (uint32_t *)c->keys_a.ptr,
                                              (uint32_t *)c->keys_b.ptr, (int32_t *)c->perm_a.ptr,
                                              (int32_t *)c->perm_b.ptr, (int)n, 0, 30, s));
    if ((rc = c->sort_tmp.ensure(tmp_bytes, "sort scratch"))) return rc;
    VH_HIP(hipcub::DeviceRadixSort::SortPairs(c->sort_tmp.ptr, tmp_bytes, (uint32_t *)c->keys_a.ptr,
                                              (uint32_t *)c->keys_b.ptr, (int32_t *)c->perm_a.ptr,
                                              (int32_t *)c->perm_b.ptr, (int)n, 0, 30, s));
    if ((nlights & 63) != 0) {
        size_t words = (size_t)((npoints * (int64_t)nlights + 63) / 64);
        VH_HIP(hipMemsetAsync(d_bits, 0, words * 8, s));
    }
    int64_t blocks = (npoints + kBlock - 1) / kBlock;
    KernelTimer timer(c, VH_KERNEL_POINT_LIGHTS, s);
    if (c->counting)
        hipLaunchKernelGGL(k_point_lights<true>, dim3((uint32_t)blocks), dim3(kBlock), 0, s,
                           (const float4 *)c->nodes.ptr, (const float4 *)c->tris.ptr, d_points,
                           (const int32_t *)c->perm_b.ptr, npoints, d_lights, nlights, tmin,
                           d_bits, (unsigned long long *)c->counters.ptr);
    else
        hipLaunchKernelGGL(k_point_lights<false>, dim3((uint32_t)blocks), dim3(kBlock), 0, s,
                           (const float4 *)c->nodes.ptr, (const float4 *)c->tris.ptr, d_points,
                           (const int32_t *)c->perm_b.ptr, npoints, d_lights, nlights, tmin,
                           d_bits, nullptr);
    timer.done();
    VH_HIP(hipGetLastError());
    return VH_OK;
}

}  // namespace vh
