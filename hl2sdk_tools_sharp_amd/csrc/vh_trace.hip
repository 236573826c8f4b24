// Ray traversal kernels for gfx950: BVH2 any-hit and closest-hit over the
// caller's rays, and shadow rays generated in-kernel from points and lights.
// The traversal itself (exact triangle test, LDS stack) is in vh_traverse.h.
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "vh_traverse.h"

#pragma clang fp contract(off)

namespace vh {
// The no-allocation point-light entry point's working memory: a counting sort
// of up to kNoAllocMaxPoints points by the top 15 bits of their Morton key
// (each point's bucket, the sorted point order, the bucket offsets).
constexpr int64_t kNoAllocMaxPoints = VH_NOALLOC_MAX_POINTS;  // 2^21
constexpr int kBucketBits = 15;
constexpr int kBuckets = 1 << kBucketBits;
constexpr size_t kScratchBytes = (size_t)kNoAllocMaxPoints * 8 + (size_t)kBuckets * 4;
}  // namespace vh

// One copy per GPU, held until the process ends.  Host code never names it
// except in hipGetSymbolAddress: its device address (per device) is kept in
// g_scratch_ptr and passed to the kernels.
static __device__ __align__(256) unsigned char scratch[vh::kScratchBytes];

namespace vh {
namespace {

constexpr int kBlock = 64;
constexpr int kMaxScratchDevices = 64;
void *g_scratch_ptr[kMaxScratchDevices] = {};
std::mutex g_scratch_mu;
constexpr int kOccluders = 2;  // occluders a point-light lane remembers

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
                int32_t h = trace_one<kClosestShort, false, kBlock>(nodes, tris, stk, r, tmin, skip_id, t, id);
                blocked = h >= 0 && (id & VH_TRACE_ID_SKY) == 0;
            } else {
                blocked = trace_one<kAny, false, kBlock>(nodes, tris, stk, r, tmin, skip_id, t, id) >= 0;
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
        int32_t h = trace_one<kClosestInclusive, false, kBlock>(nodes, tris, stk, r, tmin, skip_id, t, id);
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

// The no-allocation path's counting sort: each point's bucket (the top
// kBucketBits of its Morton key) and the bucket sizes; an exclusive scan of
// the sizes (one block); each point placed at its bucket's next slot.  The
// order inside a bucket depends on atomic timing, which changes only which
// lanes walk together: every point's hit bits are a function of that point.
__global__ void k_fill_u32(uint32_t *__restrict__ p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_morton_bucket(const float *__restrict__ pts, int64_t n, float lx, float ly, float lz,
                                float sx, float sy, float sz, uint32_t *__restrict__ bucket,
                                uint32_t *__restrict__ count)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = (pts[i * 3 + 0] - lx) * sx, y = (pts[i * 3 + 1] - ly) * sy,
          z = (pts[i * 3 + 2] - lz) * sz;
    uint32_t qx = (uint32_t)fminf(fmaxf(x, 0.0f), 1023.0f);
    uint32_t qy = (uint32_t)fminf(fmaxf(y, 0.0f), 1023.0f);
    uint32_t qz = (uint32_t)fminf(fmaxf(z, 0.0f), 1023.0f);
    uint32_t key = spread10(qx) | (spread10(qy) << 1) | (spread10(qz) << 2);
    uint32_t b = key >> (30 - kBucketBits);
    bucket[i] = b;
    atomicAdd(&count[b], 1u);
}

constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void k_bucket_scan(uint32_t *__restrict__ count)
{
    __shared__ uint32_t part[kScanThreads];
    constexpr int per = kBuckets / kScanThreads;
    const int t = threadIdx.x;
    uint32_t local[per];
    uint32_t sum = 0;
    for (int k = 0; k < per; k++) {
        local[k] = count[t * per + k];
        sum += local[k];
    }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t base = part[t] - sum;
    for (int k = 0; k < per; k++) {
        count[t * per + k] = base;
        base += local[k];
    }
}

__global__ void k_bucket_scatter(const uint32_t *__restrict__ bucket, int64_t n,
                                 uint32_t *__restrict__ next, int32_t *__restrict__ order)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t pos = atomicAdd(&next[bucket[i]], 1u);
    order[pos] = (int32_t)i;
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
// starts, and they walk the tree together (packet_any: one node sequence per
// wave, lane masks per node, the wave's stack in LDS).  Each lane first tests
// the few triangles that blocked its earlier rays (a point's own face blocks
// every light behind it): a hit is an exact hit short of the end, so the bit
// is the traversal's, and a lane it settles leaves the packet.  The lane
// assembles its own hit word: bit l of word (p, l/64).
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_point_lights(
    const float4 *__restrict__ nodes, const float4 *__restrict__ tris,
    const float *__restrict__ pts, const int32_t *__restrict__ order, int64_t npoints,
    const float *__restrict__ lights, int32_t nlights, float tmin, uint64_t *__restrict__ bits,
    unsigned long long *__restrict__ counters, int32_t depth)
{
    extern __shared__ int stack_mem[];
    int *wstk = stack_mem + (threadIdx.x >> 6) * 3 * depth;
    const int64_t slot = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t nvisit = 0, ntest = 0, slots[2] = {0, 0};
    const bool have = slot < npoints;
    const int64_t p = have ? (order ? order[slot] : slot) : 0;
    const float px = have ? pts[p * 3 + 0] : 0.0f, py = have ? pts[p * 3 + 1] : 0.0f,
                pz = have ? pts[p * 3 + 2] : 0.0f;
    const bool whole_words = (nlights & 63) == 0;
    int32_t occ[kOccluders];  // the lane's remembered occluders (leaf-order triangles)
    for (int c = 0; c < kOccluders; c++) occ[c] = -1;
    // every lane of the wave runs every light together
    for (int32_t g = 0; g < nlights; g += 64) {
        const int32_t cnt = nlights - g < 64 ? nlights - g : 64;
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
            const bool fin = have && ray_finite(r);
            bool blocked = false;
            // the remembered occluders, kept as a transpose list: a hit moves
            // one place forward, so the often-blocking ones stay in front
            for (int c = 0; c < kOccluders && fin && !blocked; c++) {
                if (occ[c] < 0) break;
                if (COUNT) ++ntest;
                if (occluder_blocks(tris, occ[c], r, tmin)) {
                    blocked = true;
                    if (c > 0) {
                        const int32_t x = occ[c - 1];
                        occ[c - 1] = occ[c];
                        occ[c] = x;
                    }
                }
            }
            int32_t hit = -1;
            const bool want = fin && !blocked;
            const bool any = by_octant(wave_octant(want, r), [&](auto oct) {
                return packet_any<COUNT, decltype(oct)::value>(nodes, tris, wstk, r, tmin, want,
                                                               &nvisit, &ntest, &hit, slots);
            });
            if (any) {
                blocked = true;
                int c = 0;
                while (c < kOccluders - 1 && occ[c] >= 0) c++;
                occ[c] = hit;  // an empty slot, else the last
            }
            w |= (uint64_t)blocked << l;
        }
        if (!have) continue;
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
    // every lane of the wave reaches the counter reduction together
    if (COUNT) {
        wave_add(counters, 0, nvisit);
        wave_add(counters, 1, ntest);
        wave_add(counters, 2, slots[0]);
        wave_add(counters, 3, slots[1]);
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
    // light buffers pay a per-light build over every triangle, which a large
    // batch of points amortises; small batches walk the BVH
    const bool lb = c->pl_method == VH_POINT_LIGHTS_LIGHT_BUFFER ||
                    (c->pl_method == VH_POINT_LIGHTS_AUTO && npoints >= 65536 && npoints >= c->ntri);
    if (lb)
        return launch_point_lights_lb(c, d_points, (const int32_t *)c->perm_b.ptr, npoints, d_lights,
                                      nlights, tmin, d_bits, s);
    int64_t blocks = (npoints + kBlock - 1) / kBlock;
    // each wave's stack holds one (node, lane mask) entry per internal level
    int32_t depth = c->depth + 1 < kMaxDepth ? c->depth + 1 : kMaxDepth;
    size_t lds = (size_t)(kBlock / 64) * 3 * depth * sizeof(int);
    KernelTimer timer(c, VH_KERNEL_POINT_LIGHTS, s);
    if (c->counting)
        hipLaunchKernelGGL(k_point_lights<true>, dim3((uint32_t)blocks), dim3(kBlock), lds, s,
                           (const float4 *)c->nodes.ptr, (const float4 *)c->tris.ptr, d_points,
                           (const int32_t *)c->perm_b.ptr, npoints, d_lights, nlights, tmin,
                           d_bits, (unsigned long long *)c->counters.ptr, depth);
    else
        hipLaunchKernelGGL(k_point_lights<false>, dim3((uint32_t)blocks), dim3(kBlock), lds, s,
                           (const float4 *)c->nodes.ptr, (const float4 *)c->tris.ptr, d_points,
                           (const int32_t *)c->perm_b.ptr, npoints, d_lights, nlights, tmin,
                           d_bits, nullptr, depth);
    timer.done();
    VH_HIP(hipGetLastError());
    return VH_OK;
}

int noalloc_init(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) {
        set_error("device index " + std::to_string(device) + " is past the no-allocation table");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return VH_OK;
    void *p = nullptr;
    VH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
    return VH_OK;
}

int launch_point_lights_noalloc(vh_ctx *c, const float *d_points, int64_t npoints,
                                const float *d_lights, int32_t nlights, float tmin,
                                uint64_t *d_bits, hipStream_t s)
{
    if (npoints > kNoAllocMaxPoints) {
        set_error("vh_trace_point_lights_noalloc takes at most " + std::to_string(kNoAllocMaxPoints) +
                  " points per call (its working memory is fixed); " + std::to_string(npoints) +
                  " were given: split the batch or call vh_trace_point_lights_dev");
        return VH_EINVAL;
    }
    if (npoints <= 0 || nlights <= 0) return VH_OK;
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c->device >= 0 && c->device < kMaxScratchDevices) base = g_scratch_ptr[c->device];
    }
    if (!base) {
        set_error("the no-allocation working memory of this device was not set up at vh_open");
        return VH_ESTATE;
    }
    uint32_t *bucket = (uint32_t *)base;
    int32_t *order = (int32_t *)((char *)base + (size_t)kNoAllocMaxPoints * 4);
    uint32_t *next = (uint32_t *)((char *)base + (size_t)kNoAllocMaxPoints * 8);
    float sc[3];
    for (int k = 0; k < 3; k++) {
        float ext = c->hi[k] - c->lo[k];
        sc[k] = ext > 0 ? 1024.0f / ext : 0.0f;
    }
    const uint32_t pblocks = (uint32_t)((npoints + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(kBuckets / 256), dim3(256), 0, s, next, (int64_t)kBuckets, 0u);
    hipLaunchKernelGGL(k_morton_bucket, dim3(pblocks), dim3(256), 0, s, d_points, npoints, c->lo[0],
                       c->lo[1], c->lo[2], sc[0], sc[1], sc[2], bucket, next);
    hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(kScanThreads), 0, s, next);
    hipLaunchKernelGGL(k_bucket_scatter, dim3(pblocks), dim3(256), 0, s, bucket, npoints, next, order);
    VH_HIP(hipGetLastError());
    if ((nlights & 63) != 0) {
        int64_t words = (npoints * (int64_t)nlights + 63) / 64;
        uint32_t wb = (uint32_t)std::min<int64_t>((words + 255) / 256, 65536);
        hipLaunchKernelGGL(k_fill_u64, dim3(wb), dim3(256), 0, s, (unsigned long long *)d_bits, words);
    }
    int32_t depth = c->depth + 1 < kMaxDepth ? c->depth + 1 : kMaxDepth;
    size_t lds = (size_t)(kBlock / 64) * 3 * depth * sizeof(int);
    hipLaunchKernelGGL(k_point_lights<false>, dim3((uint32_t)((npoints + kBlock - 1) / kBlock)),
                       dim3(kBlock), lds, s, (const float4 *)c->nodes.ptr, (const float4 *)c->tris.ptr,
                       d_points, (const int32_t *)order, npoints, d_lights, nlights, tmin, d_bits,
                       nullptr, depth);
    VH_HIP(hipGetLastError());
    return VH_OK;
}

}  // namespace vh
