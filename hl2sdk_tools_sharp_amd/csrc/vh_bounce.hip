// The transfer build and the radiosity bounce on gfx950.
//
// Transfer build (VisMatrix.BuildAsync, Rad/Bounce/VisMatrix.cs:219-620).
// One lane per receiver enumerates its candidate sources in the reference's
// order (PVS faces, the face-plane test, hierarchical descent), traces each
// candidate's visibility ray inline through the BVH (vh_traverse.h), and
// applies MakeTransfer and MakeScales.  A count pass sizes each receiver's
// staging range; the make pass fills it; a compaction lays the rows out as CSR
// by patch index and derives FP16 weights with a per-row scale.
//
// Bounce (Radiosity.BounceAsync, Rad/Bounce/Radiosity.cs:132-316), per
// iteration: k_gather (one wave per receiver row: sum shoot[src] * w with FP32
// accumulation), k_collect (leaves take the gathered light, FP64 partial sums
// of the residual and the emission), k_pullup (parents from their children,
// one launch per tree level, deepest first), k_reduce (the FP64 partials in a
// fixed order).  The iteration is captured once as a hipGraph and replayed.
#include <hip/hip_fp16.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vh_traverse.h"

#pragma clang fp contract(off)

struct BounceData {
    int32_t npatch = 0, nface = 0, ncluster = 0, nrecv = 0, nleaves = 0, height = 0;
    // patch system
    vh::DevBuf origin_area, normal_pd, children, next_parent, face, cluster, winding, refl, sky;
    vh::DevBuf face_sky, face_parent, row_faces_ptr, row_faces, receivers, leaves, levels;
    std::vector<int32_t> level_ptr;  // parents by height: level_ptr[h-1]..level_ptr[h]
    std::vector<int32_t> h_receivers;
    std::vector<int32_t> h_child1;
    // transfers
    bool have_transfers = false;
    int64_t ntransfers = 0;
    vh::DevBuf row_ptr, src, w32, w16, rscale, gather_rows;
    // bounce state
    vh::DevBuf e0, emit, shoot, add, total, partials, result;
    // bumped patches: flags, the three bump normals (float4 each), and the
    // gathered light and totals in the three bump directions
    bool has_bump = false;
    vh::DevBuf bumped, bnorm, add_b, total_b;
    // the faces' direct light handed to the patches (vh_patch_lights)
    vh::DevBuf direct, sample_area;
    bool have_direct = false;
    // the last bounce's emission and totals are valid; a predicted start
    bool have_e0 = false, have_total = false, have_pred = false;
    uint64_t generation = 0;  // distinct for every vh_patches_load in the process
    vh::DevBuf pred;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounds of the patch origins
    double *h_result = nullptr;  // pinned
};

namespace vh {
namespace {

constexpr int kTB = 64;          // transfer build: one wave per block, a lane per receiver
constexpr int kPatchStack = 32;  // hierarchical-descent stack per lane
constexpr double kPlaneTestEps = 0.01;     // VisMatrix.PlaneTestEpsilon
constexpr double kTransferEps = 0.0000001; // VisMatrix.TransferEpsilon
constexpr int64_t kMaxPatchesPerRow = 4 * 65536;  // VisMatrix.MaxPatches
constexpr int kGB = 256;         // gather: four rows per block
constexpr int kCB = 256;         // collect
constexpr float kF16Top = 16384.0f;  // a row's largest weight maps here in FP16

struct PatchDev {
    const float4 *origin_area;  // origin xyz, area
    const float4 *normal_pd;    // plane normal xyz, plane distance
    const int2 *children;
    const int32_t *next_parent;
    const int32_t *face;
    const int32_t *cluster;
    const float *winding;  // 12 per patch
    const uint8_t *face_sky;
    const int32_t *face_parent;
    const int32_t *row_faces_ptr;
    const int32_t *row_faces;
};

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx) + (ay * by) + (az * bz);
}

// Vec3.Normalise (MapFormats/Geometry/Vec3.cs:178-190).
__device__ __forceinline__ float normalise3(float x, float y, float z, float &nx, float &ny,
                                            float &nz)
{
    float len = sqrtf((x * x) + (y * y) + (z * z));
    if (len == 0.0f) {
        nx = ny = nz = 0.0f;
        return 0.0f;
    }
    nx = x / len;
    ny = y / len;
    nz = z / len;
    return len;
}

// FormFactors.DiffToDiff (Rad/Bounce/FormFactors.cs), Correct normalise.
__device__ __forceinline__ float diff_to_diff(float4 o1, float4 n1, float4 o2, float4 n2)
{
    float dx, dy, dz;
    float length = normalise3(o1.x - o2.x, o1.y - o2.y, o1.z - o2.z, dx, dy, dz);
    return -dot3(dx, dy, dz, n1.x, n1.y, n1.z) * dot3(dx, dy, dz, n2.x, n2.y, n2.z) /
           (length * length);
}

// FormFactors.PolyToDiff, Correct side (a sine above 1 is clamped).  The
// arcsine is evaluated in double and rounded once to float, the reference's
// correctly rounded DetMathF.Asin except where that double rounds twice.
__device__ float poly_to_diff(const float *poly, float area, float4 o, float4 n)
{
    float ff = 0.0f;
    for (int p = 0; p < 4; p++) {
        int q = p < 3 ? p + 1 : 0;
        float v1x, v1y, v1z, v2x, v2y, v2z;
        normalise3(poly[3 * p] - o.x, poly[3 * p + 1] - o.y, poly[3 * p + 2] - o.z, v1x, v1y, v1z);
        normalise3(poly[3 * q] - o.x, poly[3 * q + 1] - o.y, poly[3 * q + 2] - o.z, v2x, v2y, v2z);
        float cx = (v1y * v2z) - (v1z * v2y);
        float cy = (v1z * v2x) - (v1x * v2z);
        float cz = (v1x * v2y) - (v1y * v2x);
        float gx, gy, gz;
        float sin_alpha = normalise3(cx, cy, cz, gx, gy, gz);
        if (sin_alpha < -1.0f || sin_alpha > 1.0f) sin_alpha = sin_alpha > 1.0f ? 1.0f : -1.0f;
        float as = (float)asin((double)sin_alpha);
        gx = gx * as;
        gy = gy * as;
        gz = gz * as;
        ff += dot3(gx, gy, gz, n.x, n.y, n.z);
    }
    ff *= 0.5f / area;
    return ff;
}

// VisMatrix.MakeTransfer (Correct side).
__device__ __forceinline__ bool make_transfer(const PatchDev &pd, float4 ro, float4 rn, int32_t s,
                                              float4 so, float &trans)
{
    if (pd.face_sky[pd.face[s]]) return false;
    float area2 = so.w;
    if (area2 <= 0) return false;
    float4 sn = pd.normal_pd[s];
    float scale = diff_to_diff(so, sn, ro, rn);
    if (scale <= 0) return false;
    float dx = ro.x - so.x, dy = ro.y - so.y, dz = ro.z - so.z;
    float threshold = (float)((M_PI * 0.04) * (double)dot3(dx, dy, dz, dx, dy, dz));
    if (threshold < area2) {
        scale = poly_to_diff(pd.winding + 12 * (size_t)s, area2, ro, rn);
        if ((double)scale <= 0.0) return false;
    }
    trans = area2 * scale;
    return (double)trans > kTransferEps;
}

// The candidates one PVS face gives a receiver, in the reference's order
// (the inner loops of VisMatrix.BuildAsync): the face's patch trees
// descended while the receiver is within four times a parent's size; a leaf
// (or a patch not descended into) is a candidate when it is in front of the
// receiver's plane.  f(s, so) sees each candidate.  The descent stack is the
// lane's LDS column (pdepth rows); returns false if a descent outgrew it.
template <class F>
__device__ __forceinline__ bool face_candidates(const PatchDev &pd, float4 ro, float4 rn, int32_t head,
                                                int *pstk, int32_t pdepth, F &&f)
{
    bool ok = true;
    for (int32_t p = head; p != -1; p = pd.next_parent[p]) {
        int sp = 0;
        pstk[0] = p;
        sp = 1;
        while (sp > 0) {
            sp--;
            int32_t sc = pstk[sp * kTB];
            float4 so = pd.origin_area[sc];
            int2 ch = pd.children[sc];
            if (ch.x != -1) {
                float tx = ro.x - so.x, ty = ro.y - so.y, tz = ro.z - so.z;
                if ((double)dot3(tx, ty, tz, tx, ty, tz) * 0.0625 < (double)so.w) {
                    if (sp + 2 > pdepth) {
                        ok = false;
                        continue;
                    }
                    pstk[sp * kTB] = ch.y;
                    pstk[(sp + 1) * kTB] = ch.x;  // child1 first, as the recursion
                    sp += 2;
                    continue;
                }
            }
            if (!((double)dot3(so.x, so.y, so.z, rn.x, rn.y, rn.z) > (double)rn.w + kPlaneTestEps))
                continue;
            f(sc, so);
        }
    }
    return ok;
}

// One wave per receiver, a lane per PVS face of its cluster (64 faces at a
// time, in row order): the faces it skips are its own and those whose plane
// it is not in front of.  LIST = false counts the receiver's candidates;
// LIST = true writes their sources in the reference's order from the
// receiver's offset (each lane's face's candidates after those of the faces
// before it, by a prefix sum over the wave).
template <bool LIST>
__global__ __launch_bounds__(kTB) void k_candidates(PatchDev pd, const int32_t *__restrict__ recv,
                                                    int32_t nrecv, int64_t *__restrict__ counts,
                                                    const int64_t *__restrict__ cand_off,
                                                    int32_t *__restrict__ cand_src, int *__restrict__ err,
                                                    int32_t pdepth)
{
    extern __shared__ int stacks[];  // the lanes' descent stacks, [pdepth][lane]
    int *pstk = stacks + threadIdx.x;
    const int lane = lane_id();
    const int64_t k = blockIdx.x;
    if (k >= nrecv) return;  // whole wave
    const int32_t r = recv[k];
    const float4 ro = pd.origin_area[r];
    const float4 rn = pd.normal_pd[r];
    const int32_t cl = pd.cluster[r], myface = pd.face[r];
    const int64_t base = LIST ? cand_off[k] : 0;
    const int64_t limit = LIST ? cand_off[k + 1] - base : 0;
    int64_t cnt = 0;  // the receiver's candidates so far (the same in every lane)
    bool ok = true, over = false;
    const int32_t q_end = pd.row_faces_ptr[cl + 1];
    for (int32_t q0 = pd.row_faces_ptr[cl]; q0 < q_end; q0 += 64) {
        const int32_t q = q0 + lane;
        int32_t head = -1;
        if (q < q_end) {
            const int32_t fc = pd.row_faces[q];
            if (fc != myface) {
                head = pd.face_parent[fc];
                if (head >= 0) {
                    const float4 fn = pd.normal_pd[head];
                    if (!((double)dot3(ro.x, ro.y, ro.z, fn.x, fn.y, fn.z) > (double)fn.w + kPlaneTestEps))
                        head = -1;
                }
            }
        }
        int32_t n = 0;
        if (head >= 0) ok &= face_candidates(pd, ro, rn, head, pstk, pdepth, [&](int32_t, float4) { n++; });
        // exclusive prefix of the lanes' counts
        int32_t incl = n;
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (LIST && head >= 0 && n > 0) {
            int64_t at = cnt + (incl - n);
            face_candidates(pd, ro, rn, head, pstk, pdepth, [&](int32_t sc, float4) {
                if (at < limit)
                    cand_src[base + at] = sc;
                else
                    over = true;
                at++;
            });
        }
        cnt += __shfl(incl, 63, 64);
    }
    if (!ok) atomicOr(err, 1);
    if (LIST && (over || cnt != limit)) atomicOr(err, 2);
    if (!LIST && lane == 0) counts[k] = cnt;
}

constexpr int kRayBatches = 32;  // 64-candidate batches a transfer-ray wave takes

// The candidate rays, one lane per candidate over all receivers' candidates
// in order (VisMatrix.TransferRay and MakeTransfer).  A wave takes
// kRayBatches x 64 consecutive candidates, mostly one receiver's rays from one
// start.  It first sorts them by direction (octant, major axis, the larger
// minor axis: 48 bins) with LDS counters, so a batch of 64 walks the tree as
// a packet of like directions, mostly in one octant (the specialised walk);
// each candidate's answer depends only on its own ray, so the order of the
// walk changes nothing.  cand_w[c] gets the transfer's weight, or 0 when the
// ray is blocked or MakeTransfer drops it.
__global__ __launch_bounds__(64) void k_transfer_rays(PatchDev pd, const int32_t *__restrict__ recv,
                                                      int32_t nrecv, const int64_t *__restrict__ cand_off,
                                                      const int32_t *__restrict__ cand_src, int64_t ncand,
                                                      const float4 *__restrict__ nodes,
                                                      const float4 *__restrict__ tris,
                                                      float *__restrict__ cand_w,
                                                      unsigned long long *__restrict__ stats)
{
    extern __shared__ int wstk[];  // the wave's walk stack, 3 ints a level
    __shared__ uint16_t order[kRayBatches * 64];  // the wave's candidates by bin
    __shared__ int32_t start[49];                 // per bin: count, then next slot
    const int lane = lane_id();
    const int64_t first = (int64_t)blockIdx.x * (kRayBatches * 64);
    int32_t lo = 0, hi = nrecv;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (cand_off[mid] <= first) lo = mid; else hi = mid;
    }
    const int32_t kb0 = uniform(lo);
    if (lane < 49) start[lane] = 0;
    __syncthreads();
    // pass 1: each candidate's direction bin (48: no ray), counted
    int key[kRayBatches];
    {
        int32_t kb = kb0;
#pragma unroll
        for (int b = 0; b < kRayBatches; b++) {
            const int64_t c = first + (int64_t)b * 64 + lane;
            int kk = 48;
            int32_t k = kb;
            if (c < ncand) {
                while (cand_off[k + 1] <= c) k++;
                const int32_t r = recv[k];
                const float4 ro = pd.origin_area[r], rn = pd.normal_pd[r];
                const int32_t sc = cand_src[c];
                const float4 so = pd.origin_area[sc], sn = pd.normal_pd[sc];
                const float dx = (so.x + sn.x) - (ro.x + rn.x);
                const float dy = (so.y + sn.y) - (ro.y + rn.y);
                const float dz = (so.z + sn.z) - (ro.z + rn.z);
                const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
                const int major = ax >= ay ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
                const float m1 = major == 0 ? ay : ax, m2 = major == 2 ? ay : az;
                const int sub = major * 2 + (m1 >= m2);
                kk = ((dx < 0.0f) | ((dy < 0.0f) << 1) | ((dz < 0.0f) << 2)) * 6 + sub;
            }
            kb = uniform(__builtin_amdgcn_readlane(k, 63));
            key[b] = kk;
            atomicAdd(&start[kk], 1);
        }
    }
    __syncthreads();
    if (lane == 0) {
        int32_t run = 0;
        for (int q = 0; q < 49; q++) {
            const int32_t n = start[q];
            start[q] = run;
            run += n;
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kRayBatches; b++) {
        const int32_t pos = atomicAdd(&start[key[b]], 1);
        order[pos] = (uint16_t)(b * 64 + lane);
    }
    __syncthreads();
    uint32_t nblocked = 0;
    for (int b = 0; b < kRayBatches; b++) {
        const int64_t c = first + order[b * 64 + lane];
        const bool valid = c < ncand;
        int32_t k = kb0;
        if (valid)
            while (cand_off[k + 1] <= c) k++;
        Ray ray{};
        float4 ro{}, rn{}, so{};
        int32_t sc = 0;
        bool fin = false;
        if (valid) {
            const int32_t r = recv[k];
            ro = pd.origin_area[r];
            rn = pd.normal_pd[r];
            sc = cand_src[c];
            so = pd.origin_area[sc];
            // the transfer ray starts one unit off the receiver and ends one unit
            // off the source (VisMatrix.MakeRays, TransferRay), exact divide
            const float p1x = ro.x + rn.x, p1y = ro.y + rn.y, p1z = ro.z + rn.z;
            const float4 sn = pd.normal_pd[sc];
            const float p2x = so.x + sn.x, p2y = so.y + sn.y, p2z = so.z + sn.z;
            const float dx = p2x - p1x, dy = p2y - p1y, dz = p2z - p1z;
            const float len = sqrtf((dx * dx) + (dy * dy) + (dz * dz));
            const float safe = len == 0.0f ? 1.1920929E-07f : len;
            ray.ox = p1x;
            ray.oy = p1y;
            ray.oz = p1z;
            ray.dx = dx / safe;
            ray.dy = dy / safe;
            ray.dz = dz / safe;
            ray.reach = len;
            fin = isfinite(ray.dx) && isfinite(ray.dy) && isfinite(ray.dz);
        }
        const bool blocked = by_octant(wave_octant(fin, ray), [&](auto oct) {
            return packet_any<false, decltype(oct)::value>(nodes, tris, wstk, ray, 0.0f, fin);
        });
        if (valid) {
            float w = 0.0f;
            if (blocked) {
                nblocked++;
            } else {
                float t;
                if (make_transfer(pd, ro, rn, sc, so, t)) w = t;
            }
            cand_w[c] = w;
        }
    }
    unsigned long long x = nblocked;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) atomicAdd(&stats[0], x);
}

// One wave per receiver: the kept candidates (weight > 0; MakeTransfer keeps
// only weights above 1e-7), in order and at most kMaxPatchesPerRow of them,
// packed to the front of the row's staging range, and VisMatrix.MakeScales'
// factor from their float sum taken left to right.
__global__ __launch_bounds__(kGB) void k_rowscan(int32_t nrecv, const int64_t *__restrict__ cand_off,
                                                 const int32_t *__restrict__ cand_src,
                                                 const float *__restrict__ cand_w,
                                                 int32_t *__restrict__ st_src, float *__restrict__ st_w,
                                                 int64_t *__restrict__ made, float *__restrict__ msc)
{
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * (kGB / 64) + (threadIdx.x >> 6);
    if (k >= nrecv) return;  // whole wave
    const int64_t base = cand_off[k], end = cand_off[k + 1];
    int64_t kept = 0;
    float tot = 0.0f;  // the same value in every lane
    for (int64_t j0 = base; j0 < end && kept < kMaxPatchesPerRow; j0 += 64) {
        const int64_t j = j0 + lane;
        const float w = j < end ? cand_w[j] : 0.0f;
        uint64_t bal = __builtin_amdgcn_ballot_w64(w > 0.0f);
        const int64_t room = kMaxPatchesPerRow - kept;
        if ((int64_t)__builtin_popcountll(bal) > room) {
            // keep only the first `room` of this batch
            uint64_t m = bal;
            for (int64_t i = 0; i < room; i++) m &= m - 1;
            bal &= ~m;
        }
        const int pos = __builtin_popcountll(bal & ((1ull << lane) - 1ull));
        if ((bal >> lane) & 1ull) {
            st_src[base + kept + pos] = cand_src[j];
            st_w[base + kept + pos] = w;
        }
        for (uint64_t m = bal; m != 0; m &= m - 1)
            tot += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w), __builtin_ctzll(m)));
        kept += __builtin_popcountll(bal);
    }
    if (lane == 0) {
        made[k] = kept;
        msc[k] = kept > 0 ? ((double)tot > M_PI ? 1.0f / tot : (float)(1.0f / M_PI)) : 1.0f;
    }
}

// One wave per receiver: its staged row to its CSR place, and the FP16 copy
// with the row's scale (the largest weight maps to kF16Top).
__global__ __launch_bounds__(kGB) void k_compact(const int32_t *__restrict__ recv, int32_t nrecv,
                                                 const int64_t *__restrict__ cand_off,
                                                 const int64_t *__restrict__ row_ptr,
                                                 const int32_t *__restrict__ st_src,
                                                 const float *__restrict__ st_w,
                                                 const float *__restrict__ msc,
                                                 int32_t *__restrict__ src, float *__restrict__ w32,
                                                 __half *__restrict__ w16,
                                                 float *__restrict__ rscale)
{
    int lane = threadIdx.x & 63;
    int64_t k = (int64_t)blockIdx.x * (kGB / 64) + (threadIdx.x >> 6);
    if (k >= nrecv) return;
    int32_t r = recv[k];
    int64_t from = cand_off[k];
    int64_t to = row_ptr[r];
    int64_t n = row_ptr[r + 1] - to;
    const float f = msc[k];  // VisMatrix.MakeScales' factor for the row
    float mx = 0.0f;
    for (int64_t j = lane; j < n; j += 64) {
        float x = st_w[from + j] * f;
        src[to + j] = st_src[from + j];
        w32[to + j] = x;
        mx = fmaxf(mx, fabsf(x));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    float scale = mx > 0.0f ? mx / kF16Top : 1.0f;
    for (int64_t j = lane; j < n; j += 64) w16[to + j] = __float2half_rn((st_w[from + j] * f) / scale);
    if (lane == 0) rscale[r] = scale;
}

// The transfers as 8-byte records, receiver after receiver: the source patch,
// then the weight's bits (TransferSetCodec.Chunks' record,
// Rad/Bounce/TransferSetCodec.cs:85-98; the device is little-endian).
__global__ __launch_bounds__(256) void k_records(const int32_t *__restrict__ src,
                                                 const float *__restrict__ w, int64_t n,
                                                 uint2 *__restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j < n) out[j] = make_uint2((uint32_t)src[j], __float_as_uint(w[j]));
}

// A device buffer released with its scope.
struct LocalRecords {
    DevBuf buf;
    ~LocalRecords() { buf.release(); }
};

// Scales weights given as FP32 CSR into FP16 + row scale (for loaded transfers).
__global__ __launch_bounds__(kGB) void k_half_rows(const int32_t *__restrict__ rows, int32_t nrows,
                                                   const int64_t *__restrict__ row_ptr,
                                                   const float *__restrict__ w32,
                                                   __half *__restrict__ w16,
                                                   float *__restrict__ rscale)
{
    int lane = threadIdx.x & 63;
    int64_t k = (int64_t)blockIdx.x * (kGB / 64) + (threadIdx.x >> 6);
    if (k >= nrows) return;
    int32_t r = rows[k];
    int64_t b = row_ptr[r], e = row_ptr[r + 1];
    float mx = 0.0f;
    for (int64_t j = b + lane; j < e; j += 64) mx = fmaxf(mx, fabsf(w32[j]));
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    float scale = mx > 0.0f ? mx / kF16Top : 1.0f;
    for (int64_t j = b + lane; j < e; j += 64) w16[j] = __float2half_rn(w32[j] / scale);
    if (lane == 0) rscale[r] = scale;
}

// GatherLight (Radiosity.cs:216-252), flat: one wave per receiver row, FP32
// accumulation of shoot[src] * weight, the row reduced across the wave.  Lane
// l takes transfers l, l + 64, ... so each load instruction of the wave reads
// 64 consecutive transfers: 256 contiguous bytes of sources and of weights,
// and shooters of about 64 consecutive patches (rows are sorted by source),
// a few cache lines.  Four such steps are in flight at once.
// GatherBumpedScalar / GatherBumpedFour (Radiosity.cs:406-512) for one bumped
// receiver row, one wave: per transfer the direction to the source
// normalised (exact length, true divides, zero for a zero length), the
// receiver's cosine divided out, and shoot * (w * scale) added along each of
// the four normals (the plane normal, then the bump basis) the source is not
// behind (Accumulate, :523-531).  Partial sums per lane, reduced across the
// wave.
template <bool F16>
__device__ void gather_bumped(int32_t r, int64_t b, int64_t e, const int32_t *__restrict__ src,
                              const float *__restrict__ w32, const __half *__restrict__ w16,
                              const float *__restrict__ rscale, const float4 *__restrict__ shoot,
                              const float4 *__restrict__ origin_area,
                              const float4 *__restrict__ normal_pd,
                              const float4 *__restrict__ bnorm, float4 *__restrict__ add,
                              float4 *__restrict__ add_b)
{
    const int lane = threadIdx.x & 63;
    const float4 ro = origin_area[r];
    float4 nrm[4];
    nrm[0] = normal_pd[r];
    nrm[1] = bnorm[(size_t)r * 3];
    nrm[2] = bnorm[(size_t)r * 3 + 1];
    nrm[3] = bnorm[(size_t)r * 3 + 2];
    float sum[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++) sum[n][0] = sum[n][1] = sum[n][2] = 0.0f;
    for (int64_t j = b + lane; j < e; j += 64) {
        const int32_t s = src[j];
        const float w = F16 ? __half2float(w16[j]) : w32[j];
        const float4 so = origin_area[s];
        float dx = so.x - ro.x, dy = so.y - ro.y, dz = so.z - ro.z;
        const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
        if (len != 0.0f) {
            dx = dx / len;
            dy = dy / len;
            dz = dz / len;
        } else {
            dx = dy = dz = 0.0f;
        }
        const float scale = 1.0f / ((dx * nrm[0].x + dy * nrm[0].y) + dz * nrm[0].z);
        const float ws = w * scale;
        const float4 h = shoot[s];
        const float vx = h.x * ws, vy = h.y * ws, vz = h.z * ws;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const float d = (dx * nrm[n].x + dy * nrm[n].y) + dz * nrm[n].z;
            if (!(d <= 0.0f)) {
                sum[n][0] = sum[n][0] + vx * d;
                sum[n][1] = sum[n][1] + vy * d;
                sum[n][2] = sum[n][2] + vz * d;
            }
        }
    }
#pragma unroll
    for (int n = 0; n < 4; n++)
        for (int c = 0; c < 3; c++)
            for (int off = 32; off > 0; off >>= 1) sum[n][c] += __shfl_xor(sum[n][c], off, 64);
    if (lane == 0) {
        const float sc = F16 ? rscale[r] : 1.0f;
        add[r] = make_float4(sum[0][0] * sc, sum[0][1] * sc, sum[0][2] * sc, 0.0f);
#pragma unroll
        for (int n = 1; n < 4; n++)
            add_b[(size_t)r * 3 + n - 1] = make_float4(sum[n][0] * sc, sum[n][1] * sc, sum[n][2] * sc, 0.0f);
    }
}

template <bool F16>
__global__ __launch_bounds__(kGB) void k_gather(const int32_t *__restrict__ rows, int32_t nrows,
                                                const int64_t *__restrict__ row_ptr,
                                                const int32_t *__restrict__ src,
                                                const float *__restrict__ w32,
                                                const __half *__restrict__ w16,
                                                const float *__restrict__ rscale,
                                                const float4 *__restrict__ shoot,
                                                float4 *__restrict__ add,
                                                const uint8_t *__restrict__ bumped,
                                                const float4 *__restrict__ origin_area,
                                                const float4 *__restrict__ normal_pd,
                                                const float4 *__restrict__ bnorm,
                                                float4 *__restrict__ add_b)
{
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * (kGB / 64) + (threadIdx.x >> 6);
    if (k >= nrows) return;
    const int32_t r = rows[k];
    const int64_t b = row_ptr[r], e = row_ptr[r + 1];
    if (bumped && bumped[r]) {
        gather_bumped<F16>(r, b, e, src, w32, w16, rscale, shoot, origin_area, normal_pd, bnorm, add,
                           add_b);
        return;
    }
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    auto acc = [&](const float4 &h, float w) {
        ax = __builtin_fmaf(h.x, w, ax);
        ay = __builtin_fmaf(h.y, w, ay);
        az = __builtin_fmaf(h.z, w, az);
    };
    int64_t j = b + lane;
    // whole steps of four: the four loads of sources and weights, then the four
    // shooter gathers, are all in flight before the first is used
    for (; j + 192 < e; j += 256) {
        const int32_t s0 = src[j], s1 = src[j + 64], s2 = src[j + 128], s3 = src[j + 192];
        float w0, w1, w2, w3;
        if (F16) {
            w0 = __half2float(w16[j]);
            w1 = __half2float(w16[j + 64]);
            w2 = __half2float(w16[j + 128]);
            w3 = __half2float(w16[j + 192]);
        } else {
            w0 = w32[j];
            w1 = w32[j + 64];
            w2 = w32[j + 128];
            w3 = w32[j + 192];
        }
        const float4 h0 = shoot[s0], h1 = shoot[s1], h2 = shoot[s2], h3 = shoot[s3];
        acc(h0, w0);
        acc(h1, w1);
        acc(h2, w2);
        acc(h3, w3);
    }
    // the last, partial step: up to three slots, each read only inside the row
    if (j < e) {
        float4 h[3];
        float w[3];
        _Pragma("unroll") for (int u = 0; u < 3; u++) {
            const int64_t ju = j + 64 * u;
            h[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            w[u] = 0.0f;
            if (ju < e) {
                w[u] = F16 ? __half2float(w16[ju]) : w32[ju];
                h[u] = shoot[src[ju]];
            }
        }
        _Pragma("unroll") for (int u = 0; u < 3; u++) acc(h[u], w[u]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        ax += __shfl_xor(ax, off, 64);
        ay += __shfl_xor(ay, off, 64);
        az += __shfl_xor(az, off, 64);
    }
    if (lane == 0) {
        float sc = F16 ? rscale[r] : 1.0f;
        add[r] = make_float4(ax * sc, ay * sc, az * sc, 0.0f);
    }
}

__device__ __forceinline__ double block_sum(double v, double *lds)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += lds[i];
    return s;
}

// CollectLight for the leaves (Radiosity.cs:268-316): sky emits nothing; a
// leaf takes the gathered light.  Block partials (FP64): the residual's
// numerator and denominator and the summed emission per channel.
__global__ __launch_bounds__(kCB) void k_collect(const int32_t *__restrict__ leaves, int32_t n,
                                                 const uint8_t *__restrict__ sky, int mode,
                                                 const float4 *__restrict__ add,
                                                 float4 *__restrict__ total,
                                                 float4 *__restrict__ emit,
                                                 float4 *__restrict__ shoot,
                                                 const float4 *__restrict__ e0,
                                                 const float4 *__restrict__ refl,
                                                 double *__restrict__ partials,
                                                 const uint8_t *__restrict__ bumped,
                                                 const float4 *__restrict__ add_b,
                                                 float4 *__restrict__ total_b)
{
    __shared__ double lds[kCB / 64];
    int64_t i = (int64_t)blockIdx.x * kCB + threadIdx.x;
    double num = 0, den = 0, ex = 0, ey = 0, ez = 0;
    if (i < n) {
        int32_t p = leaves[i];
        if (sky[p]) {
            emit[p] = make_float4(0, 0, 0, 0);
            shoot[p] = make_float4(0, 0, 0, 0);
        } else {
            float4 a = add[p], t = total[p], em;
            if (mode == 0) {
                num = fabs((double)a.x) + fabs((double)a.y) + fabs((double)a.z);
                t = make_float4(t.x + a.x, t.y + a.y, t.z + a.z, 0.0f);
                em = a;
            } else {
                num = fabs((double)a.x - (double)t.x) + fabs((double)a.y - (double)t.y) +
                      fabs((double)a.z - (double)t.z);
                t = a;
                float4 b = e0[p];
                em = make_float4(b.x + a.x, b.y + a.y, b.z + a.z, 0.0f);
            }
            den = fabs((double)t.x) + fabs((double)t.y) + fabs((double)t.z);
            ex = em.x;
            ey = em.y;
            ez = em.z;
            float4 rf = refl[p];
            if (bumped && bumped[p]) {
                // the bump directions collect as the flat slot does
                for (int n = 0; n < 3; n++) {
                    const float4 ab = add_b[(size_t)p * 3 + n];
                    float4 tb = total_b[(size_t)p * 3 + n];
                    tb = mode == 0 ? make_float4(tb.x + ab.x, tb.y + ab.y, tb.z + ab.z, 0.0f) : ab;
                    total_b[(size_t)p * 3 + n] = tb;
                }
            }
            total[p] = t;
            emit[p] = make_float4(em.x, em.y, em.z, 0.0f);
            shoot[p] = make_float4(em.x * rf.x, em.y * rf.y, em.z * rf.z, 0.0f);
        }
    }
    double v[5] = {num, den, ex, ey, ez};
    for (int j = 0; j < 5; j++) {
        double s = block_sum(v[j], lds);
        if (threadIdx.x == 0) partials[(size_t)blockIdx.x * 5 + j] = s;
    }
}

// Parents of one tree level from their children: v = v[c1]*s1 + v[c2]*s2 with
// s = area ratio (Radiosity.cs:292-311); WITH_SHOOT also refreshes shoot.
template <bool WITH_SHOOT>
__global__ __launch_bounds__(kCB) void k_pullup(const int32_t *__restrict__ list, int32_t n,
                                                const int2 *__restrict__ children,
                                                const float4 *__restrict__ origin_area,
                                                float4 *__restrict__ v,
                                                float4 *__restrict__ shoot,
                                                const float4 *__restrict__ refl)
{
    int64_t i = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (i >= n) return;
    int32_t p = list[i];
    int2 ch = children[p];
    float a1 = origin_area[ch.x].w, a2 = origin_area[ch.y].w;
    float s1 = a1 / (a1 + a2);
    float s2 = a2 / (a1 + a2);
    float4 x = v[ch.x], y = v[ch.y];
    float4 r = make_float4((x.x * s1) + (y.x * s2), (x.y * s1) + (y.y * s2),
                           (x.z * s1) + (y.z * s2), 0.0f);
    v[p] = r;
    if (WITH_SHOOT) {
        float4 rf = refl[p];
        shoot[p] = make_float4(r.x * rf.x, r.y * rf.y, r.z * rf.z, 0.0f);
    }
}

// The bump directions' totals of bumped parents from their children, one tree
// level (Radiosity.cs:292-311: a parent blends every normal it has).
__global__ __launch_bounds__(kCB) void k_pullup_bump(const int32_t *__restrict__ list, int32_t n,
                                                     const int2 *__restrict__ children,
                                                     const float4 *__restrict__ origin_area,
                                                     const uint8_t *__restrict__ bumped,
                                                     float4 *__restrict__ tb)
{
    int64_t i = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (i >= n) return;
    int32_t p = list[i];
    if (!bumped[p]) return;
    int2 ch = children[p];
    float a1 = origin_area[ch.x].w, a2 = origin_area[ch.y].w;
    float s1 = a1 / (a1 + a2);
    float s2 = a2 / (a1 + a2);
    for (int k = 0; k < 3; k++) {
        float4 x = tb[(size_t)ch.x * 3 + k], y = tb[(size_t)ch.y * 3 + k];
        tb[(size_t)p * 3 + k] = make_float4((x.x * s1) + (y.x * s2), (x.y * s1) + (y.y * s2),
                                            (x.z * s1) + (y.z * s2), 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_zero4(float4 *__restrict__ v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The FP64 partials summed in a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ partials, int32_t nblk,
                                                double *__restrict__ out)
{
    __shared__ double lds[4];
    for (int j = 0; j < 5; j++) {
        double s = 0.0;
        for (int32_t b = threadIdx.x; b < nblk; b += 256) s += partials[(size_t)b * 5 + j];
        s = block_sum(s, lds);
        if (threadIdx.x == 0) out[j] = s;
    }
}

// Start of the bounce.  Progressive: emit = direct light as given for every
// patch (MoveDirectLightToEmission, Radiosity.cs:177-187).  From a start x0:
// leaves total = x0, emit = direct + x0 (sky 0); parents are pulled up after.
__global__ __launch_bounds__(kCB) void k_init(int32_t P, int mode, const float4 *__restrict__ e0,
                                              const float4 *__restrict__ start,
                                              const int2 *__restrict__ children,
                                              const uint8_t *__restrict__ sky,
                                              const float4 *__restrict__ refl,
                                              float4 *__restrict__ emit,
                                              float4 *__restrict__ shoot,
                                              float4 *__restrict__ total,
                                              float4 *__restrict__ add)
{
    int64_t p = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (p >= P) return;
    float4 z = make_float4(0, 0, 0, 0);
    float4 d = e0[p];
    float4 em = d, t = z;
    if (mode == 1) {
        float4 x = start ? start[p] : z;
        t = x;
        em = make_float4(d.x + x.x, d.y + x.y, d.z + x.z, 0.0f);
        if (sky[p]) em = z;
    }
    float4 rf = refl[p];
    emit[p] = em;
    shoot[p] = make_float4(em.x * rf.x, em.y * rf.y, em.z * rf.z, 0.0f);
    total[p] = t;
    add[p] = z;
    (void)children;
}

int check_loaded(vh_ctx *c, bool need_transfers)
{
    if (!c->bounce || c->bounce->npatch == 0) {
        set_error("no patches loaded: call vh_patches_load first");
        return VH_ESTATE;
    }
    if (need_transfers && !c->bounce->have_transfers) {
        set_error("no transfers: call vh_transfers_build or vh_transfers_load first");
        return VH_ESTATE;
    }
    return VH_OK;
}

template <typename T>
int upload(DevBuf &b, const T *host, size_t count, const char *what, hipStream_t s)
{
    int rc = b.ensure(count * sizeof(T) + 16, what);
    if (rc) return rc;
    if (count) VH_HIP(hipMemcpyAsync(b.ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    return VH_OK;
}

PatchDev patch_dev(BounceData *B)
{
    PatchDev pd;
    pd.origin_area = (const float4 *)B->origin_area.ptr;
    pd.normal_pd = (const float4 *)B->normal_pd.ptr;
    pd.children = (const int2 *)B->children.ptr;
    pd.next_parent = (const int32_t *)B->next_parent.ptr;
    pd.face = (const int32_t *)B->face.ptr;
    pd.cluster = (const int32_t *)B->cluster.ptr;
    pd.winding = (const float *)B->winding.ptr;
    pd.face_sky = (const uint8_t *)B->face_sky.ptr;
    pd.face_parent = (const int32_t *)B->face_parent.ptr;
    pd.row_faces_ptr = (const int32_t *)B->row_faces_ptr.ptr;
    pd.row_faces = (const int32_t *)B->row_faces.ptr;
    return pd;
}

// Allocates the CSR for `total` transfers and the gather's row list.
int alloc_transfers(BounceData *B, int64_t total)
{
    int rc;
    size_t T = (size_t)(total > 0 ? total : 1);
    if ((rc = B->src.ensure(T * 4, "transfer sources")) ||
        (rc = B->w32.ensure(T * 4, "transfer weights")) ||
        (rc = B->w16.ensure(T * 2, "FP16 transfer weights")) ||
        (rc = B->rscale.ensure((size_t)B->npatch * 4, "row scales")))
        return rc;
    return VH_OK;
}

// AddSampleToPatch and the leaf step of BuildPatchLights
// (Rad/Light/PatchLighting.cs:65-130, 194-249): a lane per leaf patch walks
// its face's samples in sample order.  A sample whose average light is under 1
// is dropped; one whose position grown by sqrt(area)/2 (a float root, a double
// halving narrowed) overlaps the patch winding's bounds adds its area and
// area * light.  The leaf's direct light is SampleLight * (float)(1/SampleArea).
// Sky patches take nothing.  Parents follow by k_pullup, level by level.
__global__ __launch_bounds__(kCB) void k_patch_lights(const int32_t *__restrict__ leaves, int32_t nleaves,
                                                      const int32_t *__restrict__ face,
                                                      const uint8_t *__restrict__ sky,
                                                      const float *__restrict__ winding,
                                                      const int32_t *__restrict__ face_ptr,
                                                      const float *__restrict__ pos,
                                                      const float *__restrict__ sarea,
                                                      const float4 *__restrict__ light,
                                                      float4 *__restrict__ direct)
{
    int64_t i = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (i >= nleaves) return;
    int32_t p = leaves[i];
    float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (!sky[p]) {
        const float *w = winding + (size_t)p * 12;
        float mn[3], mx[3];
        for (int k = 0; k < 3; k++) mn[k] = mx[k] = w[k];
        for (int q = 1; q < 4; q++)
            for (int k = 0; k < 3; k++) {
                float v = w[3 * q + k];
                mn[k] = v < mn[k] ? v : mn[k];
                mx[k] = v > mx[k] ? v : mx[k];
            }
        int32_t f = face[p];
        float sa = 0.0f, lx = 0.0f, ly = 0.0f, lz = 0.0f;
        for (int32_t s = face_ptr[f]; s < face_ptr[f + 1]; s++) {
            float4 L = light[s];
            if (((L.x + L.y) + L.z) / 3.0f < 1.0f) continue;
            float a = sarea[s];
            float r = (float)((double)sqrtf(a) / 2.0);
            bool inside = true;
            for (int k = 0; k < 3; k++) {
                float x = pos[3 * (size_t)s + k];
                if (mn[k] > x + r || mx[k] < x - r) inside = false;
            }
            if (!inside) continue;
            sa += a;
            lx = lx + (a * L.x);
            ly = ly + (a * L.y);
            lz = lz + (a * L.z);
        }
        if (sa != 0.0f) {
            float sc = (float)(1.0 / (double)sa);
            out = make_float4(lx * sc, ly * sc, lz * sc, 0.0f);
        }
    }
    direct[p] = out;
}

int upload_gather_rows(BounceData *B, const std::vector<int64_t> &row_ptr, hipStream_t s)
{
    // rows with transfers, in patch order so the CSR streams sequentially
    std::vector<int32_t> rows;
    for (int32_t p = 0; p < B->npatch; p++)
        if (row_ptr[p + 1] > row_ptr[p]) rows.push_back(p);
    int rc = upload(B->gather_rows, rows.data(), rows.size(), "gather rows", s);
    if (rc) return rc;
    B->nrecv = (int32_t)rows.size();
    return VH_OK;
}

}  // namespace

void bounce_release(vh_ctx *c)
{
    BounceData *B = c->bounce;
    if (!B) return;
    for (DevBuf *b : {&B->origin_area, &B->normal_pd, &B->children, &B->next_parent, &B->face,
                      &B->cluster, &B->winding, &B->refl, &B->sky, &B->face_sky, &B->face_parent,
                      &B->row_faces_ptr, &B->row_faces, &B->receivers, &B->leaves, &B->levels,
                      &B->row_ptr, &B->src, &B->w32, &B->w16, &B->rscale, &B->gather_rows,
                      &B->e0, &B->emit, &B->shoot, &B->add, &B->total, &B->partials, &B->result,
                      &B->direct, &B->sample_area, &B->pred, &B->bumped, &B->bnorm, &B->add_b,
                      &B->total_b})
        b->release();
    if (B->h_result) hipHostFree(B->h_result);
    delete B;
    c->bounce = nullptr;
}

bool bounce_view(vh_ctx *c, BounceView *v)
{
    BounceData *B = c->bounce;
    if (!B || B->npatch == 0) return false;
    v->npatch = B->npatch;
    v->nleaves = B->nleaves;
    v->origin_area = (const float4 *)B->origin_area.ptr;
    v->normal_pd = (const float4 *)B->normal_pd.ptr;
    v->e0 = B->have_e0 ? (const float4 *)B->e0.ptr : nullptr;
    v->direct = B->have_direct ? (const float4 *)B->direct.ptr : nullptr;
    v->total = B->have_total ? (const float4 *)B->total.ptr : nullptr;
    v->total_b = B->have_total && B->has_bump ? (const float4 *)B->total_b.ptr : nullptr;
    v->leaves = (const int32_t *)B->leaves.ptr;
    v->sky = (const uint8_t *)B->sky.ptr;
    v->cluster = (const int32_t *)B->cluster.ptr;
    v->generation = B->generation;
    for (int k = 0; k < 3; k++) {
        v->lo[k] = B->lo[k];
        v->hi[k] = B->hi[k];
    }
    return true;
}

float4 *bounce_pred_buffer(vh_ctx *c)
{
    BounceData *B = c->bounce;
    if (!B || B->pred.ensure((size_t)B->npatch * 16 + 16, "predicted start")) return nullptr;
    B->have_pred = false;
    return (float4 *)B->pred.ptr;
}

void bounce_pred_ready(vh_ctx *c) { c->bounce->have_pred = true; }

void bounce_pull_up(vh_ctx *c, float4 *v, hipStream_t s)
{
    BounceData *B = c->bounce;
    const int32_t *levels = (const int32_t *)B->levels.ptr;
    for (int32_t lv = 0; lv < B->height; lv++) {
        int32_t a = B->level_ptr[lv], n = B->level_ptr[lv + 1] - a;
        if (n > 0)
            hipLaunchKernelGGL(k_pullup<false>, dim3((uint32_t)((n + kCB - 1) / kCB)), dim3(kCB), 0, s,
                               levels + a, n, (const int2 *)B->children.ptr,
                               (const float4 *)B->origin_area.ptr, v, nullptr, (const float4 *)B->refl.ptr);
    }
}

}  // namespace vh

using namespace vh;

extern "C" {

int vh_patches_load(vh_ctx *c, const vh_patch_desc *d)
{
    clear_error();
    if (!c || !d) {
        set_error("NULL context or descriptor");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    VH_HIP(hipSetDevice(c->device));
    int32_t P = d->npatch, F = d->nface, C = d->ncluster, R = d->nreceivers;
    if (P <= 0 || F <= 0 || C <= 0 || R < 0 || !d->origin || !d->normal || !d->plane_dist ||
        !d->area || !d->child1 || !d->child2 || !d->next_parent || !d->face || !d->cluster ||
        !d->winding || !d->reflectivity || !d->face_sky || !d->face_parent || !d->row_faces_ptr ||
        !d->row_faces || (R > 0 && !d->receivers) || (d->bumped && !d->bump_normals)) {
        set_error("incomplete patch descriptor");
        return VH_EINVAL;
    }
    // validate the structure the kernels walk
    for (int32_t p = 0; p < P; p++) {
        int32_t a = d->child1[p], b = d->child2[p];
        if ((a == -1) != (b == -1) || (a != -1 && (a <= p || b <= p || a >= P || b >= P))) {
            set_error("patch " + std::to_string(p) + ": children must both be -1 or both later patches");
            return VH_EINVAL;
        }
        if (d->face[p] < 0 || d->face[p] >= F || d->cluster[p] < -1 || d->cluster[p] >= C ||
            d->next_parent[p] < -1 || d->next_parent[p] >= P) {
            set_error("patch " + std::to_string(p) + " has an out-of-range face, cluster or chain");
            return VH_EINVAL;
        }
    }
    for (int32_t f = 0; f < F; f++)
        if (d->face_parent[f] < -1 || d->face_parent[f] >= P) {
            set_error("face " + std::to_string(f) + " has an out-of-range first patch");
            return VH_EINVAL;
        }
    int32_t nrow = d->row_faces_ptr[C];
    for (int32_t k = 0; k < C; k++)
        if (d->row_faces_ptr[k] < 0 || d->row_faces_ptr[k] > d->row_faces_ptr[k + 1]) {
            set_error("row_faces_ptr is not monotone");
            return VH_EINVAL;
        }
    for (int32_t k = 0; k < nrow; k++)
        if (d->row_faces[k] < 0 || d->row_faces[k] >= F) {
            set_error("row_faces holds an out-of-range face");
            return VH_EINVAL;
        }
    for (int32_t k = 0; k < R; k++) {
        int32_t r = d->receivers[k];
        if (r < 0 || r >= P || d->child1[r] != -1 || d->cluster[r] < 0) {
            set_error("receiver " + std::to_string(k) + " is not a leaf patch with a cluster");
            return VH_EINVAL;
        }
    }
    // heights; the descent stack must hold the deepest chain of children
    std::vector<int32_t> h(P, 0);
    int32_t H = 0;
    for (int32_t p = P - 1; p >= 0; p--)
        if (d->child1[p] != -1) {
            h[p] = 1 + std::max(h[d->child1[p]], h[d->child2[p]]);
            H = std::max(H, h[p]);
        }
    if (H + 2 > kPatchStack) {
        set_error("patch trees deeper than " + std::to_string(kPatchStack - 2) + " levels");
        return VH_EINVAL;
    }

    bounce_release(c);
    BounceData *B = new BounceData();
    static std::atomic<uint64_t> generation{0};
    B->generation = ++generation;
    c->bounce = B;
    B->npatch = P;
    B->nface = F;
    B->ncluster = C;
    B->height = H;
    B->h_child1.assign(d->child1, d->child1 + P);
    B->h_receivers.assign(d->receivers, d->receivers + R);
    std::vector<float> oa((size_t)P * 4), np((size_t)P * 4), rf((size_t)P * 4);
    std::vector<int32_t> ch((size_t)P * 2);
    std::vector<uint8_t> sky(P);
    for (int32_t p = 0; p < P; p++) {
        for (int k = 0; k < 3; k++) {
            oa[4 * p + k] = d->origin[3 * p + k];
            np[4 * p + k] = d->normal[3 * p + k];
            rf[4 * p + k] = d->reflectivity[3 * p + k];
        }
        oa[4 * p + 3] = d->area[p];
        np[4 * p + 3] = d->plane_dist[p];
        rf[4 * p + 3] = 0.0f;
        ch[2 * p] = d->child1[p];
        ch[2 * p + 1] = d->child2[p];
        sky[p] = d->face_sky[d->face[p]];
    }
    std::vector<int32_t> leaves, levels;
    for (int32_t p = 0; p < P; p++)
        if (d->child1[p] == -1) leaves.push_back(p);
    for (int k = 0; k < 3; k++) {
        B->lo[k] = P > 0 ? d->origin[k] : 0.0f;
        B->hi[k] = B->lo[k];
    }
    for (int32_t p = 0; p < P; p++)
        for (int k = 0; k < 3; k++) {
            float v = d->origin[3 * p + k];
            B->lo[k] = v < B->lo[k] ? v : B->lo[k];
            B->hi[k] = v > B->hi[k] ? v : B->hi[k];
        }
    // the parents by height, in patch order within a height (a counting sort)
    B->level_ptr.assign((size_t)H + 1, 0);
    for (int32_t p = 0; p < P; p++)
        if (h[p] > 0) B->level_ptr[h[p]]++;
    for (int32_t lv = 1; lv <= H; lv++) B->level_ptr[lv] += B->level_ptr[lv - 1];
    levels.resize((size_t)B->level_ptr[H]);
    {
        std::vector<int32_t> fill(B->level_ptr.begin(), B->level_ptr.end() - 1);
        for (int32_t p = 0; p < P; p++)
            if (h[p] > 0) levels[fill[h[p] - 1]++] = p;
    }
    B->nleaves = (int32_t)leaves.size();
    hipStream_t s = c->stream;
    int rc;
    if ((rc = upload(B->origin_area, oa.data(), oa.size(), "patch origins", s)) ||
        (rc = upload(B->normal_pd, np.data(), np.size(), "patch normals", s)) ||
        (rc = upload(B->refl, rf.data(), rf.size(), "reflectivity", s)) ||
        (rc = upload(B->children, ch.data(), ch.size(), "children", s)) ||
        (rc = upload(B->sky, sky.data(), sky.size(), "sky flags", s)) ||
        (rc = upload(B->next_parent, d->next_parent, (size_t)P, "face chains", s)) ||
        (rc = upload(B->face, d->face, (size_t)P, "patch faces", s)) ||
        (rc = upload(B->cluster, d->cluster, (size_t)P, "patch clusters", s)) ||
        (rc = upload(B->winding, d->winding, (size_t)P * 12, "windings", s)) ||
        (rc = upload(B->face_sky, d->face_sky, (size_t)F, "face sky flags", s)) ||
        (rc = upload(B->face_parent, d->face_parent, (size_t)F, "face parents", s)) ||
        (rc = upload(B->row_faces_ptr, d->row_faces_ptr, (size_t)C + 1, "PVS rows", s)) ||
        (rc = upload(B->row_faces, d->row_faces, (size_t)nrow, "PVS faces", s)) ||
        (rc = upload(B->receivers, d->receivers, (size_t)R, "receivers", s)) ||
        (rc = upload(B->leaves, leaves.data(), leaves.size(), "leaves", s)) ||
        (rc = upload(B->levels, levels.data(), levels.size(), "levels", s)))
        return rc;
    size_t v4 = (size_t)P * 16;
    int nblk = (B->nleaves + kCB - 1) / kCB;
    if ((rc = B->e0.ensure(v4, "emission")) || (rc = B->emit.ensure(v4, "emission")) ||
        (rc = B->shoot.ensure(v4, "shooters")) || (rc = B->add.ensure(v4, "gathered light")) ||
        (rc = B->total.ensure(v4, "total light")) ||
        (rc = B->partials.ensure((size_t)(nblk > 0 ? nblk : 1) * 5 * 8, "partial sums")) ||
        (rc = B->result.ensure(64, "residuals")))
        return rc;
    if (!B->h_result) VH_HIP(hipHostMalloc((void **)&B->h_result, 64, hipHostMallocDefault));
    B->has_bump = false;
    if (d->bumped) {
        std::vector<uint8_t> bf(d->bumped, d->bumped + P);
        std::vector<float4> bn((size_t)P * 3);
        for (int32_t p = 0; p < P; p++) {
            bf[p] = bf[p] != 0;
            B->has_bump = B->has_bump || bf[p];
            for (int k = 0; k < 3; k++)
                bn[(size_t)p * 3 + k] = make_float4(d->bump_normals[9 * p + 3 * k], d->bump_normals[9 * p + 3 * k + 1],
                                                    d->bump_normals[9 * p + 3 * k + 2], 0.0f);
        }
        if ((rc = upload(B->bumped, bf.data(), bf.size(), "bump flags", s)) ||
            (rc = upload(B->bnorm, bn.data(), bn.size(), "bump normals", s)) ||
            (rc = B->add_b.ensure(v4 * 3, "bump gathered light")) ||
            (rc = B->total_b.ensure(v4 * 3, "bump totals")))
            return rc;
    }
    VH_WAIT(c, s);
    return VH_OK;
}

int vh_transfers_build(vh_ctx *c, int64_t *transfers, int64_t *candidates, int64_t *blocked)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_loaded(c, false);
    if (rc) return rc;
    if (!c->have_scene) {
        set_error("no scene loaded: the transfer build traces against vh_scene_load's casters");
        return VH_ESTATE;
    }
    VH_HIP(hipSetDevice(c->device));
    BounceData *B = c->bounce;
    hipStream_t s = c->stream;
    int32_t R = (int32_t)B->h_receivers.size();
    PatchDev pd = patch_dev(B);
    DevBuf counts, cand_off, cand_src, cand_w, st_src, st_w, made, msc, stats, err;
    auto cleanup = [&]() {
        for (DevBuf *b : {&counts, &cand_off, &cand_src, &cand_w, &st_src, &st_w, &made, &msc, &stats, &err})
            b->release();
    };
    if ((rc = counts.ensure((size_t)(R + 1) * 8, "candidate counts")) ||
        (rc = cand_off.ensure((size_t)(R + 1) * 8, "candidate offsets")) ||
        (rc = made.ensure((size_t)(R + 1) * 8, "row lengths")) ||
        (rc = msc.ensure((size_t)(R + 1) * 4, "row scale factors")) ||
        (rc = stats.ensure(16, "build statistics")) || (rc = err.ensure(16, "build errors"))) {
        cleanup();
        return rc;
    }
    VH_HIP(hipMemsetAsync(stats.ptr, 0, 16, s));
    VH_HIP(hipMemsetAsync(err.ptr, 0, 16, s));
    // stack rows: the tree's depth bounds a traversal stack; a patch descent
    // holds at most height + 1 entries
    const int32_t tdepth = c->depth + 1 < kMaxDepth ? c->depth + 1 : kMaxDepth;
    const int32_t pdepth = B->height + 2 < kPatchStack ? B->height + 2 : kPatchStack;
    KernelTimer timer(c, VH_KERNEL_TRANSFERS, s);
    // 1. candidates per receiver
    if (R > 0)
        hipLaunchKernelGGL(k_candidates<false>, dim3((uint32_t)R), dim3(kTB), (size_t)pdepth * kTB * 4, s, pd,
                           (const int32_t *)B->receivers.ptr, R, (int64_t *)counts.ptr, nullptr, nullptr,
                           (int *)err.ptr, pdepth);
    VH_HIP(hipGetLastError());
    std::vector<int64_t> hcount(R + 1, 0), hoff(R + 1, 0);
    VH_WAIT(c, s);  // bounded: the copy below then finds the kernels done
    if (R > 0)
        VH_HIP(hipMemcpyAsync(hcount.data(), counts.ptr, (size_t)R * 8, hipMemcpyDeviceToHost, s));
    VH_WAIT(c, s);
    for (int32_t k = 0; k < R; k++) hoff[k + 1] = hoff[k] + hcount[k];
    int64_t ncand = hoff[R];
    const size_t nslot = (size_t)(ncand > 0 ? ncand : 1);
    if ((rc = upload(cand_off, hoff.data(), hoff.size(), "candidate offsets", s)) ||
        (rc = cand_src.ensure(nslot * 4, "candidate sources")) ||
        (rc = cand_w.ensure(nslot * 4, "candidate weights")) ||
        (rc = st_src.ensure(nslot * 4, "staged sources")) ||
        (rc = st_w.ensure(nslot * 4, "staged weights"))) {
        cleanup();
        return rc;
    }
    if (R > 0 && ncand > 0) {
        // 2. their sources, in the reference's order
        hipLaunchKernelGGL(k_candidates<true>, dim3((uint32_t)R), dim3(kTB), (size_t)pdepth * kTB * 4, s, pd,
                           (const int32_t *)B->receivers.ptr, R, nullptr, (const int64_t *)cand_off.ptr,
                           (int32_t *)cand_src.ptr, (int *)err.ptr, pdepth);
        // 3. every candidate's ray and weight, a lane each
        const int64_t waves = (ncand + kRayBatches * 64 - 1) / (kRayBatches * 64);
        hipLaunchKernelGGL(k_transfer_rays, dim3((uint32_t)waves), dim3(64), (size_t)3 * tdepth * 4, s, pd,
                           (const int32_t *)B->receivers.ptr, R, (const int64_t *)cand_off.ptr,
                           (const int32_t *)cand_src.ptr, ncand, (const float4 *)c->nodes.ptr,
                           (const float4 *)c->tris.ptr, (float *)cand_w.ptr,
                           (unsigned long long *)stats.ptr);
    }
    // 4. each row's kept transfers, packed in order, and its MakeScales factor
    if (R > 0)
        hipLaunchKernelGGL(k_rowscan, dim3((uint32_t)((R + 3) / 4)), dim3(kGB), 0, s, R,
                           (const int64_t *)cand_off.ptr, (const int32_t *)cand_src.ptr,
                           (const float *)cand_w.ptr, (int32_t *)st_src.ptr, (float *)st_w.ptr,
                           (int64_t *)made.ptr, (float *)msc.ptr);
    timer.done();
    VH_HIP(hipGetLastError());
    std::vector<int64_t> hmade(R + 1, 0);
    int herr = 0;
    unsigned long long hstats[2] = {0, 0};
    VH_WAIT(c, s);  // bounded: the copy below then finds the kernels done
    if (R > 0)
        VH_HIP(hipMemcpyAsync(hmade.data(), made.ptr, (size_t)R * 8, hipMemcpyDeviceToHost, s));
    VH_HIP(hipMemcpyAsync(&herr, err.ptr, 4, hipMemcpyDeviceToHost, s));
    VH_HIP(hipMemcpyAsync(hstats, stats.ptr, 16, hipMemcpyDeviceToHost, s));
    VH_WAIT(c, s);
    if (herr) {
        cleanup();
        set_error(herr & 1 ? "patch descent deeper than the stack"
                           : "candidate enumeration disagreed between its passes");
        return VH_EDEVICE;
    }
    std::vector<int64_t> rp((size_t)B->npatch + 1, 0), len((size_t)B->npatch, 0);
    for (int32_t k = 0; k < R; k++) len[B->h_receivers[k]] = hmade[k];
    for (int32_t p = 0; p < B->npatch; p++) rp[p + 1] = rp[p] + len[p];
    int64_t total = rp[B->npatch];
    if ((rc = upload(B->row_ptr, rp.data(), rp.size(), "row offsets", s)) ||
        (rc = alloc_transfers(B, total))) {
        cleanup();
        return rc;
    }
    if (R > 0)
        hipLaunchKernelGGL(k_compact, dim3((uint32_t)((R + 3) / 4)), dim3(kGB), 0, s,
                           (const int32_t *)B->receivers.ptr, R, (const int64_t *)cand_off.ptr,
                           (const int64_t *)B->row_ptr.ptr, (const int32_t *)st_src.ptr,
                           (const float *)st_w.ptr, (const float *)msc.ptr, (int32_t *)B->src.ptr,
                           (float *)B->w32.ptr, (__half *)B->w16.ptr, (float *)B->rscale.ptr);
    VH_HIP(hipGetLastError());
    if ((rc = upload_gather_rows(B, rp, s))) {
        cleanup();
        return rc;
    }
    VH_WAIT(c, s);
    cleanup();
    B->ntransfers = total;
    B->have_transfers = true;
    if (transfers) *transfers = total;
    if (candidates) *candidates = ncand;
    if (blocked) *blocked = (int64_t)hstats[0];
    return VH_OK;
}

int vh_transfers_read(vh_ctx *c, int64_t *row_ptr, int32_t *src, float *weight, int64_t capacity)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_loaded(c, true);
    if (rc) return rc;
    BounceData *B = c->bounce;
    if ((src || weight) && capacity < B->ntransfers) {
        set_error(std::to_string(B->ntransfers) + " transfers need as many slots, and " +
                  std::to_string(capacity) + " were given");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    if (row_ptr)
        VH_HIP(hipMemcpy(row_ptr, B->row_ptr.ptr, (size_t)(B->npatch + 1) * 8,
                         hipMemcpyDeviceToHost));
    if (src && B->ntransfers)
        VH_HIP(hipMemcpy(src, B->src.ptr, (size_t)B->ntransfers * 4, hipMemcpyDeviceToHost));
    if (weight && B->ntransfers)
        VH_HIP(hipMemcpy(weight, B->w32.ptr, (size_t)B->ntransfers * 4, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_transfers_export(vh_ctx *c, int32_t *counts, void *records, int64_t capacity)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_loaded(c, true);
    if (rc) return rc;
    BounceData *B = c->bounce;
    const int64_t T = B->ntransfers;
    if (records && capacity < T) {
        set_error(std::to_string(T) + " transfers need as many records, and " + std::to_string(capacity) +
                  " were given");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (counts) {
        std::vector<int64_t> rp((size_t)B->npatch + 1);
        VH_HIP(hipMemcpyAsync(rp.data(), B->row_ptr.ptr, rp.size() * 8, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
        for (int32_t p = 0; p < B->npatch; p++) {
            const int64_t n = rp[p + 1] - rp[p];
            if (n > INT32_MAX) {
                set_error("a patch's transfer list is longer than an int");
                return VH_EINVAL;
            }
            counts[p] = (int32_t)n;
        }
    }
    if (records && T > 0) {
        LocalRecords out;
        if ((rc = out.buf.ensure((size_t)T * 8, "transfer records"))) return rc;
        hipLaunchKernelGGL(k_records, dim3((uint32_t)((T + 255) / 256)), dim3(256), 0, s,
                           (const int32_t *)B->src.ptr, (const float *)B->w32.ptr, T, (uint2 *)out.buf.ptr);
        VH_HIP(hipGetLastError());
        VH_HIP(hipMemcpyAsync(records, out.buf.ptr, (size_t)T * 8, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
    }
    return VH_OK;
}

int vh_transfers_load(vh_ctx *c, const int64_t *row_ptr, const int32_t *src, const float *weight)
{
    clear_error();
    if (!c || !row_ptr) {
        set_error("NULL context or row offsets");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_loaded(c, false);
    if (rc) return rc;
    BounceData *B = c->bounce;
    int32_t P = B->npatch;
    if (row_ptr[0] != 0) {
        set_error("row_ptr[0] must be 0");
        return VH_EINVAL;
    }
    for (int32_t p = 0; p < P; p++)
        if (row_ptr[p + 1] < row_ptr[p] || (row_ptr[p + 1] > row_ptr[p] && B->h_child1[p] != -1)) {
            set_error("row_ptr is not monotone, or a parent patch has transfers");
            return VH_EINVAL;
        }
    int64_t T = row_ptr[P];
    if (T > 0 && (!src || !weight)) {
        set_error("NULL sources or weights");
        return VH_EINVAL;
    }
    for (int64_t k = 0; k < T; k++)
        if (src[k] < 0 || src[k] >= P) {
            set_error("transfer " + std::to_string(k) + " names patch " + std::to_string(src[k]) +
                      ", outside 0.." + std::to_string(P - 1));
            return VH_EINVAL;
        }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    std::vector<int64_t> rp(row_ptr, row_ptr + P + 1);
    if ((rc = upload(B->row_ptr, row_ptr, (size_t)P + 1, "row offsets", s)) ||
        (rc = alloc_transfers(B, T)) || (rc = upload(B->src, src, (size_t)T, "sources", s)) ||
        (rc = upload(B->w32, weight, (size_t)T, "weights", s)) ||
        (rc = upload_gather_rows(B, rp, s)))
        return rc;
    if (B->nrecv > 0)
        hipLaunchKernelGGL(k_half_rows, dim3((uint32_t)((B->nrecv + 3) / 4)), dim3(kGB), 0, s,
                           (const int32_t *)B->gather_rows.ptr, B->nrecv,
                           (const int64_t *)B->row_ptr.ptr, (const float *)B->w32.ptr,
                           (__half *)B->w16.ptr, (float *)B->rscale.ptr);
    VH_HIP(hipGetLastError());
    VH_WAIT(c, s);
    B->ntransfers = T;
    B->have_transfers = true;
    return VH_OK;
}

int vh_bounce(vh_ctx *c, const vh_bounce_opts *o, const float *emit0, const float *start,
              float *total_out, double *residuals, double *energies, int32_t *iters)
{
    clear_error();
    if (!c || !o) {
        set_error("NULL context or options");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_loaded(c, true);
    if (rc) return rc;
    if (!emit0 && !c->bounce->have_direct) {
        set_error("NULL emission and no vh_patch_lights result to start from");
        return VH_ESTATE;
    }
    if (o->max_iter < 1 || (o->weights != VH_WEIGHTS_F32 && o->weights != VH_WEIGHTS_F16) ||
        (o->mode != VH_BOUNCE_PROGRESSIVE && o->mode != VH_BOUNCE_FROM_START &&
         o->mode != VH_BOUNCE_FROM_PREDICTION) || std::isnan(o->tol)) {
        set_error("bad bounce options");
        return VH_EINVAL;
    }
    if (o->mode == VH_BOUNCE_FROM_PREDICTION && !c->bounce->have_pred) {
        set_error("no predicted start: call vh_mlp_predict first");
        return VH_ESTATE;
    }
    // the kernels know two forms: progressive (0) and from a start (1)
    const int kmode = o->mode == VH_BOUNCE_PROGRESSIVE ? 0 : 1;
    VH_HIP(hipSetDevice(c->device));
    BounceData *B = c->bounce;
    hipStream_t s = c->stream;
    int32_t P = B->npatch;
    std::vector<float> buf((size_t)P * 4, 0.0f);
    if (emit0) {
        for (int32_t p = 0; p < P; p++)
            for (int k = 0; k < 3; k++) buf[4 * p + k] = emit0[3 * p + k];
        VH_HIP(hipMemcpyAsync(B->e0.ptr, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, s));
    } else {
        VH_HIP(hipMemcpyAsync(B->e0.ptr, B->direct.ptr, buf.size() * 4, hipMemcpyDeviceToDevice, s));
    }
    B->have_e0 = true;
    B->have_total = false;
    float4 *startd = nullptr;
    if (o->mode == VH_BOUNCE_FROM_PREDICTION) {
        VH_HIP(hipMemcpyAsync(B->total.ptr, B->pred.ptr, buf.size() * 4, hipMemcpyDeviceToDevice, s));
        startd = (float4 *)B->total.ptr;
    }
    if (o->mode == VH_BOUNCE_FROM_START && start) {
        for (int32_t p = 0; p < P; p++)
            for (int k = 0; k < 3; k++) buf[4 * p + k] = start[3 * p + k];
        // the gathered-light buffer holds the start until k_init copies it
        VH_HIP(hipMemcpyAsync(B->add.ptr, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, s));
        VH_HIP(hipMemcpyAsync(B->total.ptr, B->add.ptr, buf.size() * 4, hipMemcpyDeviceToDevice, s));
        startd = (float4 *)B->total.ptr;
    }
    const float4 *e0 = (const float4 *)B->e0.ptr;
    float4 *emit = (float4 *)B->emit.ptr, *shoot = (float4 *)B->shoot.ptr;
    float4 *add = (float4 *)B->add.ptr, *total = (float4 *)B->total.ptr;
    const float4 *refl = (const float4 *)B->refl.ptr;
    const int2 *children = (const int2 *)B->children.ptr;
    const float4 *oa = (const float4 *)B->origin_area.ptr;
    const uint8_t *sky = (const uint8_t *)B->sky.ptr;
    // k_init reads the start from `total` and writes it back, so alias is safe per element
    hipLaunchKernelGGL(k_init, dim3((uint32_t)((P + kCB - 1) / kCB)), dim3(kCB), 0, s, P, kmode,
                       e0, (const float4 *)startd, children, sky, refl, emit, shoot, total, add);
    VH_HIP(hipGetLastError());
    const int32_t *levels = (const int32_t *)B->levels.ptr;
    if (kmode == 1)
        for (int32_t lv = 0; lv < B->height; lv++) {
            int32_t a = B->level_ptr[lv], n = B->level_ptr[lv + 1] - a;
            if (n > 0)
                hipLaunchKernelGGL(k_pullup<true>, dim3((uint32_t)((n + kCB - 1) / kCB)),
                                   dim3(kCB), 0, s, levels + a, n, children, oa, emit, shoot, refl);
        }
    VH_HIP(hipGetLastError());

    const uint8_t *bumped = B->has_bump ? (const uint8_t *)B->bumped.ptr : nullptr;
    float4 *add_b = (float4 *)B->add_b.ptr, *total_b = (float4 *)B->total_b.ptr;
    if (bumped) {
        const int64_t nb = (int64_t)P * 3;
        hipLaunchKernelGGL(k_zero4, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, s, add_b, nb);
        hipLaunchKernelGGL(k_zero4, dim3((uint32_t)((nb + 255) / 256)), dim3(256), 0, s, total_b, nb);
    }
    const float4 *npd = (const float4 *)B->normal_pd.ptr, *bnorm = (const float4 *)B->bnorm.ptr;
    int32_t nblk = (B->nleaves + kCB - 1) / kCB;
    bool f16 = o->weights == VH_WEIGHTS_F16;
    auto iteration = [&](bool timed) {
        if (B->nrecv > 0) {
            KernelTimer timer(c, VH_KERNEL_GATHER, s);
            if (!timed) timer.a = timer.b = nullptr;
            uint32_t g = (uint32_t)((B->nrecv + 3) / 4);
            if (f16)
                hipLaunchKernelGGL(k_gather<true>, dim3(g), dim3(kGB), 0, s,
                                   (const int32_t *)B->gather_rows.ptr, B->nrecv,
                                   (const int64_t *)B->row_ptr.ptr, (const int32_t *)B->src.ptr,
                                   (const float *)B->w32.ptr, (const __half *)B->w16.ptr,
                                   (const float *)B->rscale.ptr, (const float4 *)shoot, add, bumped,
                                   oa, npd, bnorm, add_b);
            else
                hipLaunchKernelGGL(k_gather<false>, dim3(g), dim3(kGB), 0, s,
                                   (const int32_t *)B->gather_rows.ptr, B->nrecv,
                                   (const int64_t *)B->row_ptr.ptr, (const int32_t *)B->src.ptr,
                                   (const float *)B->w32.ptr, (const __half *)B->w16.ptr,
                                   (const float *)B->rscale.ptr, (const float4 *)shoot, add, bumped,
                                   oa, npd, bnorm, add_b);
            if (timed) timer.done();
        }
        if (nblk > 0)
            hipLaunchKernelGGL(k_collect, dim3((uint32_t)nblk), dim3(kCB), 0, s,
                               (const int32_t *)B->leaves.ptr, B->nleaves, sky, kmode,
                               (const float4 *)add, total, emit, shoot, e0, refl,
                               (double *)B->partials.ptr, bumped, (const float4 *)add_b, total_b);
        for (int32_t lv = 0; lv < B->height; lv++) {
            int32_t a = B->level_ptr[lv], n = B->level_ptr[lv + 1] - a;
            if (n > 0)
                hipLaunchKernelGGL(k_pullup<true>, dim3((uint32_t)((n + kCB - 1) / kCB)),
                                   dim3(kCB), 0, s, levels + a, n, children, oa, emit, shoot, refl);
        }
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, (const double *)B->partials.ptr,
                           nblk, (double *)B->result.ptr);
        hipMemcpyAsync(B->h_result, B->result.ptr, 5 * 8, hipMemcpyDeviceToHost, s);
    };

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graphed = o->use_graph != 0 && !c->timing;
    if (graphed) {
        VH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        iteration(false);
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (e != hipSuccess) {
            set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
            return VH_EDEVICE;
        }
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(graph);
            set_error(std::string("graph instantiation failed: ") + hipGetErrorString(e));
            return VH_EDEVICE;
        }
    }
    int32_t it = 0;
    hipError_t failure = hipSuccess;
    int wrc = VH_OK;
    for (;;) {
        if (graphed)
            failure = hipGraphLaunch(exec, s);
        else {
            iteration(true);
            failure = hipGetLastError();
        }
        if (failure != hipSuccess) break;
        // bounded by the context's deadline and cancel flag; on either, the
        // iteration in flight (well under a millisecond) ends before its
        // graph is destroyed
        if ((wrc = wait_stream(c, s)) != VH_OK) {
            hipStreamSynchronize(s);
            break;
        }
        const double *r = B->h_result;
        double res = r[1] > 0 ? r[0] / r[1] : 0.0;
        if (residuals) residuals[it] = res;
        if (energies) {
            energies[3 * it] = r[2];
            energies[3 * it + 1] = r[3];
            energies[3 * it + 2] = r[4];
        }
        it++;
        bool stop = it >= o->max_iter;
        if (o->tol > 0)
            stop = stop || res <= o->tol;
        else
            stop = stop || (r[2] < 1.0 && r[3] < 1.0 && r[4] < 1.0);
        if (stop) break;
    }
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (failure != hipSuccess) {
        set_error(std::string("bounce iteration failed: ") + hipGetErrorString(failure));
        return VH_EDEVICE;
    }
    if (wrc != VH_OK) return wrc;
    // the parents' totals, pulled up once at the end as CollectLight leaves them
    for (int32_t lv = 0; lv < B->height; lv++) {
        int32_t a = B->level_ptr[lv], n = B->level_ptr[lv + 1] - a;
        if (n > 0)
            hipLaunchKernelGGL(k_pullup<false>, dim3((uint32_t)((n + kCB - 1) / kCB)), dim3(kCB),
                               0, s, levels + a, n, children, oa, total, nullptr, refl);
        if (n > 0 && bumped)
            hipLaunchKernelGGL(k_pullup_bump, dim3((uint32_t)((n + kCB - 1) / kCB)), dim3(kCB), 0, s,
                               levels + a, n, children, oa, bumped, total_b);
    }
    VH_HIP(hipGetLastError());
    if (total_out) {
        VH_HIP(hipMemcpyAsync(buf.data(), B->total.ptr, buf.size() * 4, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
        for (int32_t p = 0; p < P; p++)
            for (int k = 0; k < 3; k++) total_out[3 * p + k] = buf[4 * p + k];
    }
    VH_WAIT(c, s);
    if (iters) *iters = it;
    B->have_total = true;
    return VH_OK;
}

int vh_bounce_bump_totals(vh_ctx *c, float *out)
{
    clear_error();
    if (!c || !out) {
        set_error("NULL context or output");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    BounceData *B = c->bounce;
    if (!B || !B->have_total) {
        set_error("no bounce result: call vh_bounce first");
        return VH_ESTATE;
    }
    const int32_t P = B->npatch;
    if (!B->has_bump) {
        std::memset(out, 0, (size_t)P * 9 * sizeof(float));
        return VH_OK;
    }
    VH_HIP(hipSetDevice(c->device));
    std::vector<float4> buf((size_t)P * 3);
    VH_HIP(hipMemcpyAsync(buf.data(), B->total_b.ptr, buf.size() * sizeof(float4), hipMemcpyDeviceToHost,
                          c->stream));
    VH_WAIT(c, c->stream);
    for (size_t i = 0; i < buf.size(); i++) {
        out[3 * i] = buf[i].x;
        out[3 * i + 1] = buf[i].y;
        out[3 * i + 2] = buf[i].z;
    }
    return VH_OK;
}

int vh_patch_lights(vh_ctx *c, int32_t which, const float *sample_area, float *direct)
{
    clear_error();
    if (!c || (which != 0 && which != 1)) {
        set_error("NULL context or bad lighting pass");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_loaded(c, false);
    if (rc) return rc;
    GatherView gv;
    if (!gather_view(c, &gv) || !gv.have_results) {
        set_error("no gather results: call vh_gather_load and vh_gather_run first");
        return VH_ESTATE;
    }
    BounceData *B = c->bounce;
    if (gv.nfaces != B->nface) {
        set_error("the gather has " + std::to_string(gv.nfaces) + " faces and the patch system " +
                  std::to_string(B->nface) + "; they must be the same faces");
        return VH_EINVAL;
    }
    if (gv.nsamples > 0 && !sample_area) {
        set_error("NULL sample areas");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    int32_t P = B->npatch;
    if ((rc = upload(B->sample_area, sample_area, (size_t)gv.nsamples, "sample areas", s)) ||
        (rc = B->direct.ensure((size_t)P * 16 + 16, "patch direct light")))
        return rc;
    const float4 *oa = (const float4 *)B->origin_area.ptr;
    const int2 *children = (const int2 *)B->children.ptr;
    const int32_t *levels = (const int32_t *)B->levels.ptr;
    float4 *d = (float4 *)B->direct.ptr;
    if (B->nleaves > 0)
        hipLaunchKernelGGL(k_patch_lights, dim3((uint32_t)((B->nleaves + kCB - 1) / kCB)), dim3(kCB), 0,
                           s, (const int32_t *)B->leaves.ptr, B->nleaves, (const int32_t *)B->face.ptr,
                           (const uint8_t *)B->sky.ptr, (const float *)B->winding.ptr, gv.face_ptr,
                           gv.pos, (const float *)B->sample_area.ptr, which ? gv.hdr : gv.ldr, d);
    for (int32_t lv = 0; lv < B->height; lv++) {
        int32_t a = B->level_ptr[lv], n = B->level_ptr[lv + 1] - a;
        if (n > 0)
            hipLaunchKernelGGL(k_pullup<false>, dim3((uint32_t)((n + kCB - 1) / kCB)), dim3(kCB), 0, s,
                               levels + a, n, children, oa, d, nullptr, (const float4 *)B->refl.ptr);
    }
    VH_HIP(hipGetLastError());
    if (direct) {
        std::vector<float> buf((size_t)P * 4);
        VH_HIP(hipMemcpyAsync(buf.data(), d, buf.size() * 4, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
        for (int32_t p = 0; p < P; p++)
            for (int k = 0; k < 3; k++) direct[3 * p + k] = buf[4 * p + k];
    }
    VH_WAIT(c, s);
    B->have_direct = true;
    return VH_OK;
}

}  // extern "C"
