// The final light on gfx950: the luxel radial filter and the lighting lump.
//
// FinalLightFace.Run (SourceSharp.MapTools/Rad/Final/FinalLightFace.cs:93-299)
// splats each face's light samples, and its neighbours', onto the face's
// luxel grid (LuxelRadial.AddDirect), splats the bounced light of its own and
// its neighbours' leaf patches onto a second grid (AddBounced), then reads
// every luxel back normalised (Sample), clamps it to minlight and writes it
// as ColorRGBExp32, with the median of the face's flat luxels before it.
//
// The reference accumulates a luxel over the samples in a fixed order (own
// samples, then each neighbour's, then in sample order), one float add at a
// time.  Here a luxel belongs to one lane, so its sum is taken in exactly
// that order:
//
//   k_radial<BOUNCED>: one wave per 64 grid cells of one face.  The wave
//     walks the face's items (samples, or patches) 64 at a time: lane l
//     projects item l into this face's luxel space (the item's point, its
//     bounds carried corner by corner for a neighbour's sample, a patch's
//     winding bounds) and its integer window; a ballot keeps the items whose
//     window meets the wave's cells (most of a neighbour's samples miss a
//     small face), and the wave takes those in order, each broadcast from its
//     lane: every lane whose cell is in the item's window adds the item's
//     weighted light.  The cells go to HBM: weight and light per normal,
//     float4 each.
//   k_final: one wave per lit face.  Each lane samples its luxels from both
//     grids, adds them, clamps, encodes and writes the lump words, and keeps
//     the flat value of every luxel whose flat sample had weight; the median
//     of each channel is then the value whose rank (values below it, values
//     equal to it) covers count / 2.
//
// Every float operation is the reference's, in its order and precision; the
// double promotions (EQUAL_EPSILON, 0.1, RADIALDIST, 1.0 / weight) are done
// in double.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "vh_internal.h"
#include "vh_rgbe.h"

#pragma clang fp contract(off)

namespace {

constexpr int kWave = 64;
constexpr float kOneOverSqrt3 = 0.57735025882720947f;  // OO_SQRT_3, LuxelRadial.cs:51
constexpr float kWeightEps = 0.00001f;                 // WEIGHT_EPS, :48
constexpr double kRadialDist = 1.42;                   // RADIALDIST, :45
constexpr double kEqualEpsilon = 0.001;                // LightConstants.EqualEpsilonDouble
constexpr int32_t kMaxSide = 1 << 15;                  // a grid side; w * h stays below 2^30

// One face as the kernels read it (uniform per wave: scalar loads).
struct FaceDev {
    float o[3], ws[3], wt[3], ls[3], lt[3];
    int32_t min_s, min_t, w, h;
    float minlight;
    int32_t lit, nb;   // nb: 1, or 4 for a bumped face
    int32_t s0, s1;    // the face's gather samples
    int32_t pad;
    int64_t lux_base;  // first luxel in luxel_pos
};

// One light style of one lit face (FinalLightFace.Run's loop over styles).
struct UnitDev {
    int32_t face, slot, style, pad;
    int64_t cell_base;   // first cell of the unit's grids (nb * w * h float4)
    int64_t mval_base;   // first of its w * h median candidates
    int64_t lump_ofs;    // lump offset of its luxel 0
    int64_t median_ofs;  // lump offset of its median word
};

struct FinalArgs {
    const FaceDev *faces;
    const float *luxel_pos;     // 3 per luxel, or null: LuxelToWorld of the grid points
    const float4 *bounds;       // per gather sample
    const int32_t *nbr_ptr, *nbr;
    const int32_t *plist_ptr, *plist;  // leaf patches per face, in FacePatches order
    const float4 *porigin;
    const int32_t *wptr;
    const float *wpts;
    const float4 *plight;       // 4 per patch
    const float *spos;          // gather sample positions, 3 each
    const float4 *lv;           // gather light, [style slot][normal][sample]
    const uint8_t *fstyles;     // the gather's 4 style numbers per face
    int32_t nsamples;
    const UnitDev *units;
    const int2 *tiles;          // (unit, first cell)
    const int32_t *lit_units;
    float4 *cells, *pcells;
    float4 *mvals;
    uint32_t *lump;
    int32_t bounce, red_errors;
};

__device__ __forceinline__ float2 world_to_luxel(const FaceDev &f, float x, float y, float z)
{
    // FaceLightInfo.WorldToLuxel (FaceLightInfo.cs:264-270): the dot as
    // (x + y) + z, then the mins subtracted.
    float px = x - f.o[0], py = y - f.o[1], pz = z - f.o[2];
    float s = px * f.ws[0] + py * f.ws[1] + pz * f.ws[2];
    float t = px * f.wt[0] + py * f.wt[1] + pz * f.wt[2];
    return make_float2(s - (float)f.min_s, t - (float)f.min_t);
}

__device__ __forceinline__ float3 luxel_to_world(const FaceDev &f, float s, float t)
{
    // LuxelToWorld (FaceLightInfo.cs:285-291): origin, plus s, then plus t.
    s += (float)f.min_s;
    t += (float)f.min_t;
    float x = f.o[0] + f.ls[0] * s, y = f.o[1] + f.ls[1] * s, z = f.o[2] + f.ls[2] * s;
    return make_float3(x + f.lt[0] * t, y + f.lt[1] * t, z + f.lt[2] * t);
}

__device__ __forceinline__ float bcast(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ int bcast(int v, int j) { return __builtin_amdgcn_readlane(v, j); }

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ int32_t wave_inclusive_sum(int32_t v)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int32_t u = __shfl_up(v, d, kWave);
        if ((int)threadIdx.x >= d) v += u;
    }
    return v;
}

// One wave's cells of one face: the direct grid (BOUNCED false) or the patch
// grid.  NB is the face's normal count.
//
// The face's items form segments: its own samples (or leaf patches), then
// each neighbour's.  Lane q describes segment q (source face, first item,
// count); a wave prefix sum numbers the items, and the wave then takes them
// 64 at a time across segment boundaries: lane l finds its item's segment,
// projects the item and loads its light, and the items whose window meets
// the wave's cells are applied in order from registers, so no memory access
// waits inside the per-item loop.
// The slot of face f that holds light style `style`, or -1 (a neighbour's
// samples are matched by the style NUMBER, FinalLightFace.cs:340-360).
__device__ __forceinline__ int32_t slot_of_style(const FinalArgs &a, int32_t f, int32_t style)
{
    for (int k = 0; k < 4; k++)
        if ((int32_t)a.fstyles[(size_t)f * 4 + k] == style) return k;
    return -1;
}

template <bool BOUNCED, int NB>
__device__ void radial_tile(const FinalArgs &a, int32_t unit, int32_t c0)
{
    const UnitDev U = a.units[unit];
    const int32_t face = U.face;
    const FaceDev F = a.faces[face];
    const int lane = threadIdx.x;
    const int32_t w = F.w, ncell = F.w * F.h;
    const int32_t c = c0 + lane;
    const bool valid = c < ncell;
    const int32_t s = valid ? c % w : 0, t = valid ? c / w : 0;
    const float fs = (float)s, ft = (float)t;
    const int32_t t_lo = c0 / w, t_hi = (min(c0 + kWave, ncell) - 1) / w;

    float W = 0.0f;
    float L[NB][3];
#pragma unroll
    for (int b = 0; b < NB; b++) L[b][0] = L[b][1] = L[b][2] = 0.0f;

    const int32_t n0 = a.nbr_ptr[face], nseg = a.nbr_ptr[face + 1] - n0 + 1;
    for (int32_t q0 = 0; q0 < nseg; q0 += kWave) {
        const int32_t q = q0 + lane;
        int32_t src = face, first = 0, cnt = 0, sbump = 0, sslot = U.slot;
        if (q < nseg) {
            src = q == 0 ? face : a.nbr[n0 + q - 1];
            if (!BOUNCED) {
                // a neighbour without the style or without samples adds nothing
                const int32_t lit = a.faces[src].lit, nbs = a.faces[src].nb;
                if (q != 0) sslot = slot_of_style(a, src, U.style);
                first = a.faces[src].s0;
                cnt = (q == 0 || (lit && sslot >= 0)) ? a.faces[src].s1 - first : 0;
                sbump = nbs == 4;
            } else {
                first = a.plist_ptr[src];
                cnt = a.plist_ptr[src + 1] - first;
                sbump = q == 0 ? NB == 4 : a.faces[src].nb == 4;
            }
        }
        const int32_t incl = wave_inclusive_sum(cnt);
        const int32_t excl = incl - cnt;
        const int32_t total = bcast(incl, kWave - 1);
        const int32_t nq = min(kWave, nseg - q0);
        for (int32_t base = 0; base < total; base += kWave) {
            const int32_t g = base + lane;
            int32_t seg = 0;
            for (int k = 0; k < nq; k++) seg += bcast(incl, k) <= g;
            seg = min(seg, kWave - 1);
            const int32_t isrc = __shfl(src, seg, kWave);
            const int32_t ifirst = __shfl(first, seg, kWave);
            const int32_t iexcl = __shfl(excl, seg, kWave);
            const int32_t ibump = __shfl(sbump, seg, kWave);
            const int32_t islot = __shfl(sslot, seg, kWave);
            const bool mine = g < total;
            float cs = 0, ct = 0, pa = 0, pb = 0, pc = 0, pd = 0;
            float lt[NB][3];
#pragma unroll
            for (int b = 0; b < NB; b++) lt[b][0] = lt[b][1] = lt[b][2] = 0.0f;
            int32_t smin = 0, smax = 0, tmin = 0, tmax = 0;
            if (mine) {
                const int32_t item0 = ifirst + (g - iexcl);
                if (!BOUNCED) {
                    const int32_t item = item0;
                    const float *p = a.spos + (size_t)item * 3;
                    float2 cc = world_to_luxel(F, p[0], p[1], p[2]);
                    cs = cc.x;
                    ct = cc.y;
                    float4 bd = a.bounds[item];
                    if (q0 + seg == 0) {
                        pa = bd.x; pb = bd.y; pc = bd.z; pd = bd.w;
                    } else {
                        // the neighbour's bounds, corner by corner into this grid
                        const FaceDev S = a.faces[isrc];
                        float3 w0 = luxel_to_world(S, bd.x, bd.y);
                        float2 mn = world_to_luxel(F, w0.x, w0.y, w0.z);
                        float3 w1 = luxel_to_world(S, bd.z, bd.w);
                        float2 mx = world_to_luxel(F, w1.x, w1.y, w1.z);
                        pa = mn.x; pb = mn.y; pc = mx.x; pd = mx.y;
                    }
                    // AddDirect's window (LuxelRadial.cs:153-161)
                    smin = (int32_t)pa;
                    tmin = (int32_t)pb;
                    smax = (int32_t)(pc + 0.9999f) + 1;
                    tmax = (int32_t)(pd + 0.9999f) + 1;
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        if (b == 0 || ibump) {
                            float4 l = a.lv[((size_t)islot * 4 + b) * a.nsamples + item];
                            lt[b][0] = l.x; lt[b][1] = l.y; lt[b][2] = l.z;
                        }
                    }
                } else {
                    const int32_t item = a.plist[item0];
                    float4 o = a.porigin[item];
                    float2 cc = world_to_luxel(F, o.x, o.y, o.z);
                    cs = cc.x;
                    ct = cc.y;
                    // PatchLightmapCoordRange (FinalLightFace.cs:519-536)
                    float mns = 1e30f, mnt = 1e30f, mxs = -1e30f, mxt = -1e30f;
                    for (int32_t k = a.wptr[item]; k < a.wptr[item + 1]; k++) {
                        const float *p = a.wpts + (size_t)k * 3;
                        float2 st = world_to_luxel(F, p[0], p[1], p[2]);
                        mns = mns < st.x ? mns : st.x;
                        mxs = mxs > st.x ? mxs : st.x;
                        mnt = mnt < st.y ? mnt : st.y;
                        mxt = mxt > st.y ? mxt : st.y;
                    }
                    // AddBounced (LuxelRadial.cs:242-253): extents at least one
                    // luxel, the window RADIALDIST of them either side, in double
                    float ds = mxs - mns, dt = mxt - mnt;
                    ds = 1.0 > (double)ds ? 1.0f : ds;
                    dt = 1.0 > (double)dt ? 1.0f : dt;
                    pa = ds;
                    pb = dt;
                    smin = (int32_t)((double)cs - (double)ds * kRadialDist);
                    tmin = (int32_t)((double)ct - (double)dt * kRadialDist);
                    smax = (int32_t)((double)cs + (double)ds * kRadialDist + 1.0);
                    tmax = (int32_t)((double)ct + (double)dt * kRadialDist + 1.0);
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        if (b == 0 || ibump) {
                            float4 l = a.plight[(size_t)item * 4 + b];
                            lt[b][0] = l.x; lt[b][1] = l.y; lt[b][2] = l.z;
                        }
                    }
                }
                smin = clampi(smin, 0, 1 << 30);
                tmin = clampi(tmin, 0, 1 << 30);
                smax = min(smax, F.w);
                tmax = min(tmax, F.h);
            }
            // only the items whose window meets this wave's cells, in order
            const bool live = mine && smin < smax && tmin < tmax && tmax > t_lo && tmin <= t_hi;
            for (uint64_t todo = __ballot(live); todo; todo &= todo - 1) {
                const int j = __builtin_ctzll(todo);
                const int32_t jtmin = bcast(tmin, j), jtmax = bcast(tmax, j);
                const int32_t jsmin = bcast(smin, j), jsmax = bcast(smax, j);
                const float jcs = bcast(cs, j), jct = bcast(ct, j);
                const float ja = bcast(pa, j), jb = bcast(pb, j);
                const bool in = valid && s >= jsmin && s < jsmax && t >= jtmin && t < jtmax;
                float r = 0.0f;
                bool add = in;
                if (!BOUNCED) {
                    // AddDirect (LuxelRadial.cs:163-205): the overlap of the
                    // sample's box with the luxel's two-luxel window, over the
                    // Chebyshev distance floored at 0.1
                    const float jc = bcast(pc, j), jd = bcast(pd, j);
                    if (in) {
                        float x = ja - fs;
                        float s0 = x > -1.0f ? x : -1.0f;
                        x = jb - ft;
                        float t0 = x > -1.0f ? x : -1.0f;
                        x = jc - fs;
                        float s1 = x < 1.0f ? x : 1.0f;
                        x = jd - ft;
                        float t1 = x < 1.0f ? x : 1.0f;
                        float area = (s1 - s0) * (t1 - t0);
                        add = (double)area > kEqualEpsilon;
                        if (add) {
                            float ds = fabsf(jcs - fs), dt = fabsf(jct - ft);
                            r = ds > dt ? ds : dt;
                            r = (double)r < 0.1 ? (float)((double)area / 0.1) : area / r;
                        }
                    }
                } else if (in) {
                    // AddBounced (LuxelRadial.cs:255-291): 2 - d^2 in patch units
                    float ds = (jcs - fs) / ja, dt = (jct - ft) / jb;
                    r = 2.0f - (ds * ds + dt * dt);
                    add = r > 0.0f;
                }
                const bool jbump = bcast((int32_t)ibump, j) != 0;
                float l0[3] = {bcast(lt[0][0], j), bcast(lt[0][1], j), bcast(lt[0][2], j)};
                if (NB == 4 && jbump) {
                    float lb[NB][3];
#pragma unroll
                    for (int b = 1; b < NB; b++)
                        for (int k = 0; k < 3; k++) lb[b][k] = bcast(lt[b][k], j);
                    if (add) {
#pragma unroll
                        for (int k = 0; k < 3; k++) L[0][k] = L[0][k] + r * l0[k];
#pragma unroll
                        for (int b = 1; b < NB; b++)
                            for (int k = 0; k < 3; k++) L[b][k] = L[b][k] + r * lb[b][k];
                    }
                } else if (add) {
#pragma unroll
                    for (int k = 0; k < 3; k++) L[0][k] = L[0][k] + r * l0[k];
                    if (NB == 4) {
                        float bw = r * kOneOverSqrt3;
#pragma unroll
                        for (int b = 1; b < NB; b++)
                            for (int k = 0; k < 3; k++) L[b][k] = L[b][k] + bw * l0[k];
                    }
                }
                if (add) W = W + r;
            }
        }
    }
    if (valid) {
        float4 *out = (BOUNCED ? a.pcells : a.cells) + U.cell_base;
#pragma unroll
        for (int b = 0; b < NB; b++)
            out[(size_t)b * ncell + c] = make_float4(L[b][0], L[b][1], L[b][2], W);
    }
}

template <bool BOUNCED>
__global__ __launch_bounds__(kWave) void k_radial(FinalArgs a)
{
    const int2 tile = a.tiles[blockIdx.x];
    if (a.faces[a.units[tile.x].face].nb == 4)
        radial_tile<BOUNCED, 4>(a, tile.x, tile.y);
    else
        radial_tile<BOUNCED, 1>(a, tile.x, tile.y);
}

// LuxelRadial.Sample (LuxelRadial.cs:314-359) of one grid at cell (u, v),
// off_grid already decided; returns baseSampleOk.
template <int NB>
__device__ __forceinline__ bool sample_grid(const float4 *cells, int32_t ncell, bool off_grid,
                                            int32_t i, bool red_errors, float (&out)[NB][3])
{
    if (off_grid) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            out[b][0] = 2550.0f;
            out[b][1] = 0.0f;
            out[b][2] = 0.0f;
        }
        return false;
    }
    float wgt = cells[i].w;
    if (wgt > kWeightEps) {
        float sc = (float)(1.0 / (double)wgt);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            float4 l = cells[(size_t)b * ncell + i];
            out[b][0] = l.x * sc;
            out[b][1] = l.y * sc;
            out[b][2] = l.z * sc;
        }
        return true;
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
        out[b][0] = red_errors ? 2550.0f : 0.0f;
        out[b][1] = 0.0f;
        out[b][2] = 0.0f;
    }
    return false;
}

template <int NB>
__device__ void final_face(const FinalArgs &a, int32_t unit)
{
    const UnitDev U = a.units[unit];
    const FaceDev F = a.faces[U.face];
    const int lane = threadIdx.x;
    const int32_t w = F.w, h = F.h, n = F.w * F.h;
    const float4 *cells = a.cells + U.cell_base;
    const float4 *pcells = a.pcells + U.cell_base;
    uint32_t *out = a.lump + (U.lump_ofs >> 2);
    float4 *mv = a.mvals + U.mval_base;
    // bounced light goes into style 0 only (FinalLightFace.cs:182)
    const bool bounce = a.bounce && U.slot == 0;
    const bool no_samples = F.s1 == F.s0;
    const float ml = F.minlight;
    int32_t count = 0;
    for (int32_t j = lane; j - lane < n; j += kWave) {
        bool ok = false;
        if (j < n) {
            float3 p;
            if (a.luxel_pos) {
                const float *q = a.luxel_pos + (size_t)(F.lux_base + j) * 3;
                p = make_float3(q[0], q[1], q[2]);
            } else {
                // BuildFaceLuxels (Rad/Light/FaceSampleBuilder.cs:242-259)
                p = luxel_to_world(F, (float)(j % w), (float)(j / w));
            }
            float2 cc = world_to_luxel(F, p.x, p.y, p.z);
            int32_t u = (int32_t)(cc.x + 0.5f), v = (int32_t)(cc.y + 0.5f);
            bool off = u < 0 || u >= w || v < 0 || v >= h;
            int32_t i = off ? 0 : u + v * w;
            float lb[NB][3];
            ok = sample_grid<NB>(cells, n, off, i, a.red_errors, lb);
            if (bounce) {
                float pv[NB][3];
                sample_grid<NB>(pcells, n, off, i, a.red_errors, pv);
#pragma unroll
                for (int b = 0; b < NB; b++)
                    for (int k = 0; k < 3; k++) lb[b][k] = lb[b][k] + pv[b][k];
            }
            if (no_samples) {
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    lb[b][0] = 255.0f;
                    lb[b][1] = 0.0f;
                    lb[b][2] = 0.0f;
                }
                ok = false;
            }
#pragma unroll
            for (int b = 0; b < NB; b++) {
                float r = lb[b][0] > ml ? lb[b][0] : ml;
                float g = lb[b][1] > ml ? lb[b][1] : ml;
                float bl = lb[b][2] > ml ? lb[b][2] : ml;
                out[(size_t)b * n + j] = vh::encode_one(r, g, bl);
                if (b == 0) mv[j] = ok ? make_float4(r, g, bl, 1.0f) : make_float4(0, 0, 0, 0);
            }
        }
        count += __popcll(__ballot(ok));
    }
    __syncthreads();
    // FinalLightFace.Median (FinalLightFace.cs:466-472): element count / 2 of
    // the ascending order, per channel
    __shared__ float med[3];
    if (lane < 3) med[lane] = 0.0f;
    __syncthreads();
    if (count > 0) {
        const int32_t m = count >> 1;
        for (int32_t j = lane; j - lane < n; j += kWave) {
            float4 v = j < n ? mv[j] : make_float4(0, 0, 0, 0);
            int32_t lt[3] = {0, 0, 0}, eq[3] = {0, 0, 0};
            for (int32_t i = 0; i < n; i++) {
                float4 q = mv[i];
                int32_t use = q.w != 0.0f;
                lt[0] += use & (q.x < v.x);
                eq[0] += use & (q.x == v.x);
                lt[1] += use & (q.y < v.y);
                eq[1] += use & (q.y == v.y);
                lt[2] += use & (q.z < v.z);
                eq[2] += use & (q.z == v.z);
            }
            if (j < n && v.w != 0.0f) {
                if (lt[0] <= m && m < lt[0] + eq[0]) med[0] = v.x;
                if (lt[1] <= m && m < lt[1] + eq[1]) med[1] = v.y;
                if (lt[2] <= m && m < lt[2] + eq[2]) med[2] = v.z;
            }
        }
    }
    __syncthreads();
    // stored before lightofs, in reverse style order (FinalLightFace.cs:270-279)
    if (lane == 0) a.lump[U.median_ofs >> 2] = vh::encode_one(med[0], med[1], med[2]);
}

__global__ __launch_bounds__(kWave) void k_final(FinalArgs a)
{
    const int32_t unit = a.lit_units[blockIdx.x];
    if (a.faces[a.units[unit].face].nb == 4)
        final_face<4>(a, unit);
    else
        final_face<1>(a, unit);
}

// The bounce's light as TotalLight: slot 0 the flat light, slots 1-3 the bump
// directions' (zero when the bounce had no bumped patches; then only flat
// patches are read, which the host checks).
__global__ void k_plight(const float4 *__restrict__ total, const float4 *__restrict__ total_b,
                         int32_t npatch, float4 *__restrict__ plight)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npatch * 4) return;
    int32_t p = i >> 2, b = i & 3;
    float4 t = b == 0 ? total[p] : (total_b ? total_b[(size_t)p * 3 + b - 1] : make_float4(0, 0, 0, 0));
    plight[i] = make_float4(t.x, t.y, t.z, 0.0f);
}

}  // namespace

struct FinalData {
    int32_t nfaces = 0, npatch = 0, nsamples = 0, nlit = 0;
    int64_t nluxels = 0, ncells = 0, lump_bytes = 0, ntiles = 0, nbtiles = 0;
    std::vector<int32_t> h_face_ptr, h_normals, h_offsets;
    std::vector<FaceDev> h_faces;
    uint64_t style_generation = ~0ull;  // the gather styles the layout was made for
    bool bumped_owner = false;  // some bumped face owns a leaf patch
    bool have_luxels = false;   // luxel points given (else made from the frames)
    bool have_lump = false;     // a vh_final_light has filled the lump
    vh::DevBuf faces, luxel_pos, bounds, nbr_ptr, nbr, plist_ptr, plist, porigin, wptr, wpts;
    vh::DevBuf plight, units, tiles, lit_faces, cells, pcells, mvals, lump;
};

namespace vh {
void final_release(vh_ctx *c)
{
    FinalData *D = c->final_;
    if (!D) return;
    for (DevBuf *b : {&D->faces, &D->luxel_pos, &D->bounds, &D->nbr_ptr, &D->nbr, &D->plist_ptr,
                      &D->plist, &D->porigin, &D->wptr, &D->wpts, &D->plight, &D->units, &D->tiles,
                      &D->lit_faces, &D->cells, &D->pcells, &D->mvals, &D->lump})
        b->release();
    delete D;
    c->final_ = nullptr;
}

bool final_lump(vh_ctx *c, FinalLump *v)
{
    FinalData *D = c->final_;
    if (!D || !D->have_lump) return false;
    v->lump = D->lump.ptr;
    v->bytes = D->lump_bytes;
    return true;
}
}  // namespace vh

using namespace vh;

namespace {
template <typename T>
int put(DevBuf &b, const T *host, size_t count, const char *what, hipStream_t s)
{
    int rc = b.ensure(count * sizeof(T) + 16, what);
    if (rc) return rc;
    if (count) VH_HIP(hipMemcpyAsync(b.ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    return VH_OK;
}

// LightmapOffsets.Compute (Rad/Light/LightmapOffsets.cs:60-123) over the
// gather's face styles: a lit face reserves a median word per style, then 4
// bytes per luxel per normal per style; its light offset is just past the
// median words.  The (face, style) units are listed style slot by style
// slot, so the slot-0 units -- the only ones the bounce reaches -- and their
// tiles come first.
int final_layout(FinalData *D, const GatherView &gv, hipStream_t st)
{
    const int32_t F = D->nfaces;
    const std::vector<uint8_t> &sty = *gv.h_face_styles;
    std::vector<int32_t> offsets(F, -1), nst(F, 0);
    int64_t bytes = 0;
    for (int32_t f = 0; f < F; f++) {
        const FaceDev &o = D->h_faces[f];
        int k = 0;
        while (k < 4 && sty[(size_t)f * 4 + k] != 255) k++;
        if (!o.lit || k == 0) continue;
        nst[f] = k;
        bytes += 4 * (int64_t)k;
        offsets[f] = (int32_t)bytes;
        bytes += 4 * (int64_t)o.nb * o.w * o.h * k;
        if (bytes > INT32_MAX) {
            set_error("the lump would pass 2 GiB");
            return VH_EINVAL;
        }
    }
    std::vector<UnitDev> units;
    std::vector<int2> tiles;
    std::vector<int32_t> lit;
    int64_t cells = 0, mv = 0, nbtiles = 0;
    for (int slot = 0; slot < 4; slot++) {
        for (int32_t f = 0; f < F; f++) {
            if (slot >= nst[f]) continue;
            const FaceDev &o = D->h_faces[f];
            const int64_t n = (int64_t)o.w * o.h;
            UnitDev u;
            std::memset(&u, 0, sizeof u);
            u.face = f;
            u.slot = slot;
            u.style = sty[(size_t)f * 4 + slot];
            u.cell_base = cells;
            u.mval_base = mv;
            u.lump_ofs = offsets[f] + 4 * (int64_t)slot * o.nb * n;
            u.median_ofs = offsets[f] - 4 * (int64_t)(slot + 1);
            const int32_t id = (int32_t)units.size();
            units.push_back(u);
            lit.push_back(id);
            for (int64_t t0 = 0; t0 < n; t0 += kWave) tiles.push_back(make_int2(id, (int32_t)t0));
            cells += (int64_t)o.nb * n;
            mv += n;
        }
        if (slot == 0) nbtiles = (int64_t)tiles.size();
    }
    D->h_offsets = offsets;
    D->lump_bytes = bytes;
    D->ncells = cells;
    D->nlit = (int32_t)lit.size();
    D->ntiles = (int64_t)tiles.size();
    D->nbtiles = nbtiles;
    int rc;
    if ((rc = put(D->units, units.data(), units.size(), "final units", st)) ||
        (rc = put(D->tiles, tiles.data(), tiles.size(), "tiles", st)) ||
        (rc = put(D->lit_faces, lit.data(), lit.size(), "lit units", st)) ||
        (rc = D->cells.ensure((size_t)std::max<int64_t>(cells, 1) * sizeof(float4), "cells")) ||
        (rc = D->pcells.ensure((size_t)std::max<int64_t>(cells, 1) * sizeof(float4), "patch cells")) ||
        (rc = D->mvals.ensure((size_t)std::max<int64_t>(mv, 1) * sizeof(float4), "median values")) ||
        (rc = D->lump.ensure((size_t)std::max<int64_t>(bytes, 4), "lump")))
        return rc;
    VH_HIP(hipStreamSynchronize(st));  // the host vectors above must outlive the copies
    D->style_generation = gv.style_generation;
    return VH_OK;
}
}  // namespace

extern "C" {

int vh_final_load(vh_ctx *c, const vh_final_desc *d)
{
    clear_error();
    if (!c || !d) {
        set_error("NULL context or descriptor");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    GatherView gv;
    if (!gather_view(c, &gv)) {
        set_error("no gather loaded: call vh_gather_load first");
        return VH_ESTATE;
    }
    const int32_t F = d->nfaces, P = d->npatch;
    if (F != gv.nfaces || P < 0 || (F > 0 && (!d->faces || !d->nbr_ptr)) ||
        (gv.nsamples > 0 && !d->sample_bounds) ||
        (P > 0 && (!d->patch_origin || !d->patch_face || !d->patch_child1 || !d->winding_ptr))) {
        set_error("incomplete final-light descriptor (nfaces must be the gather's)");
        return VH_EINVAL;
    }
    const std::vector<int32_t> &fp = *gv.h_face_ptr, &nc = *gv.h_normals;
    std::vector<FaceDev> fd(F > 0 ? F : 1);
    int64_t lux = 0;
    for (int32_t f = 0; f < F; f++) {
        const vh_final_face &x = d->faces[f];
        if (x.width < 0 || x.height < 0 || x.width > kMaxSide || x.height > kMaxSide ||
            (int64_t)x.width * x.height >= (1 << 30) || (x.lit != 0 && x.lit != 1)) {
            set_error("face " + std::to_string(f) + ": bad grid size or lit flag");
            return VH_EINVAL;
        }
        FaceDev &o = fd[f];
        std::memset(&o, 0, sizeof o);
        for (int k = 0; k < 3; k++) {
            o.o[k] = x.origin[k];
            o.ws[k] = x.w2l_s[k];
            o.wt[k] = x.w2l_t[k];
            o.ls[k] = x.l2w_s[k];
            o.lt[k] = x.l2w_t[k];
        }
        o.min_s = x.min_s;
        o.min_t = x.min_t;
        o.w = x.width;
        o.h = x.height;
        o.minlight = x.minlight;
        o.lit = x.lit;
        o.nb = nc[f] == 4 ? 4 : 1;
        o.s0 = fp[f];
        o.s1 = fp[f + 1];
        o.lux_base = lux;
        lux += (int64_t)x.width * x.height;
    }
    if (F > 0 && d->nbr_ptr[0] != 0) {
        set_error("nbr_ptr must start at 0");
        return VH_EINVAL;
    }
    for (int32_t f = 0; f < F; f++) {
        if (d->nbr_ptr[f + 1] < d->nbr_ptr[f]) {
            set_error("nbr_ptr must not decrease");
            return VH_EINVAL;
        }
    }
    const int32_t nn = F > 0 ? d->nbr_ptr[F] : 0;
    if (nn > 0 && !d->nbr) {
        set_error("nbr is required");
        return VH_EINVAL;
    }
    for (int32_t k = 0; k < nn; k++) {
        if (d->nbr[k] < 0 || d->nbr[k] >= F) {
            set_error("neighbour " + std::to_string(d->nbr[k]) + " out of range");
            return VH_EINVAL;
        }
    }
    // each face's leaf patches, highest index first
    std::vector<int32_t> pptr(F + 1, 0), plist;
    bool bumped_owner = false;
    if (P > 0) {
        if (d->winding_ptr[0] != 0) {
            set_error("winding_ptr must start at 0");
            return VH_EINVAL;
        }
        for (int32_t p = 0; p < P; p++) {
            if (d->patch_face[p] < 0 || d->patch_face[p] >= F ||
                d->winding_ptr[p + 1] < d->winding_ptr[p]) {
                set_error("patch " + std::to_string(p) + ": bad face or winding range");
                return VH_EINVAL;
            }
            if (d->patch_child1[p] < 0) pptr[d->patch_face[p] + 1]++;
        }
        if (d->winding_ptr[P] > 0 && !d->winding) {
            set_error("winding is required");
            return VH_EINVAL;
        }
        for (int32_t f = 0; f < F; f++) {
            if (pptr[f + 1] > 0 && fd[f].nb == 4) bumped_owner = true;
            pptr[f + 1] += pptr[f];
        }
        plist.resize(pptr[F] > 0 ? pptr[F] : 1);
        std::vector<int32_t> fill(pptr.begin(), pptr.end() - 1);
        for (int32_t p = P - 1; p >= 0; p--)
            if (d->patch_child1[p] < 0) plist[fill[d->patch_face[p]]++] = p;
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    if (!c->final_) c->final_ = new FinalData();
    FinalData *D = c->final_;
    D->nfaces = F;
    D->npatch = P;
    D->nsamples = gv.nsamples;
    D->nluxels = lux;
    D->h_face_ptr = fp;
    D->h_normals = nc;
    D->h_faces.assign(fd.begin(), fd.begin() + F);
    D->bumped_owner = bumped_owner;
    D->have_luxels = d->luxel_pos != nullptr;
    D->have_lump = false;
    std::vector<float4> porg((size_t)std::max(P, 1));
    for (int32_t p = 0; p < P; p++)
        porg[p] = make_float4(d->patch_origin[3 * p], d->patch_origin[3 * p + 1],
                              d->patch_origin[3 * p + 2], 0.0f);
    const int32_t nw = P > 0 ? d->winding_ptr[P] : 0;
    const int32_t zero[2] = {0, 0};
    int rc;
    if ((rc = put(D->faces, fd.data(), (size_t)std::max(F, 1), "final faces", st)) ||
        (rc = put(D->luxel_pos, d->luxel_pos, d->luxel_pos ? (size_t)lux * 3 : 0, "luxel positions", st)) ||
        (rc = put(D->bounds, d->sample_bounds, (size_t)gv.nsamples * 4, "sample bounds", st)) ||
        (rc = put(D->nbr_ptr, F > 0 ? d->nbr_ptr : zero, (size_t)F + 1, "neighbour offsets", st)) ||
        (rc = put(D->nbr, d->nbr, (size_t)nn, "neighbours", st)) ||
        (rc = put(D->plist_ptr, pptr.data(), pptr.size(), "patch list offsets", st)) ||
        (rc = put(D->plist, plist.data(), (size_t)pptr[F], "patch lists", st)) ||
        (rc = put(D->porigin, porg.data(), (size_t)P, "patch origins", st)) ||
        (rc = put(D->wptr, P > 0 ? d->winding_ptr : zero, (size_t)P + 1, "winding offsets", st)) ||
        (rc = put(D->wpts, d->winding, (size_t)nw * 3, "windings", st)) ||
        (rc = D->plight.ensure((size_t)std::max(P, 1) * 4 * sizeof(float4), "patch lights")))
        return rc;
    // the layout follows the gather's light styles (style 0 alone until a run)
    return final_layout(D, gv, st);
}

int vh_final_layout(vh_ctx *c, int32_t *offsets, int64_t *bytes)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->final_) {
        set_error("no final-light inputs: call vh_final_load first");
        return VH_ESTATE;
    }
    GatherView gv;
    if (gather_view(c, &gv) && gv.nfaces == c->final_->nfaces &&
        gv.style_generation != c->final_->style_generation) {
        VH_HIP(hipSetDevice(c->device));
        int rc = final_layout(c->final_, gv, c->stream);
        if (rc) return rc;
    }
    if (offsets && !c->final_->h_offsets.empty())
        std::memcpy(offsets, c->final_->h_offsets.data(), c->final_->h_offsets.size() * 4);
    if (bytes) *bytes = c->final_->lump_bytes;
    return VH_OK;
}

int vh_final_light(vh_ctx *c, const vh_final_opts *o, const float *patch_light, uint8_t *lump,
                   int64_t capacity)
{
    clear_error();
    if (!c || !o || (o->which != 0 && o->which != 1) || o->bounce < VH_FINAL_NO_BOUNCE ||
        o->bounce > VH_FINAL_BOUNCE_HOST || !lump) {
        set_error("bad final-light arguments");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    FinalData *D = c->final_;
    GatherView gv;
    if (!D || !gather_view(c, &gv)) {
        set_error("no final-light inputs: call vh_gather_load and vh_final_load first");
        return VH_ESTATE;
    }
    if (!gv.have_results) {
        set_error("no gather results: call vh_gather_run first");
        return VH_ESTATE;
    }
    if (*gv.h_face_ptr != D->h_face_ptr || *gv.h_normals != D->h_normals) {
        set_error("the gather was reloaded with other faces: call vh_final_load again");
        return VH_ESTATE;
    }
    if (gv.style_generation != D->style_generation) {
        VH_HIP(hipSetDevice(c->device));
        int rc = final_layout(D, gv, c->stream);
        if (rc) return rc;
    }
    if (capacity < D->lump_bytes) {
        set_error(std::to_string(D->lump_bytes) + " lump bytes need as much room, and " +
                  std::to_string(capacity) + " were given");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int32_t P = D->npatch;
    if (o->bounce == VH_FINAL_BOUNCE_DEVICE) {
        BounceView bv;
        if (!bounce_view(c, &bv) || !bv.total) {
            set_error("no bounce result on the device: call vh_bounce first");
            return VH_ESTATE;
        }
        if (bv.npatch != P) {
            set_error("the bounce's patch system has " + std::to_string(bv.npatch) +
                      " patches and the final light's " + std::to_string(P));
            return VH_EINVAL;
        }
        if (D->bumped_owner && !bv.total_b) {
            set_error("the bounce had no bumped patches and a bumped face owns leaf patches: load "
                      "the patches with their bump flags, or pass TotalLight per normal with "
                      "VH_FINAL_BOUNCE_HOST");
            return VH_ENOTSUP;
        }
        if (P > 0)
            hipLaunchKernelGGL(k_plight, dim3((uint32_t)((P * 4 + 255) / 256)), dim3(256), 0, st,
                               bv.total, bv.total_b, P, (float4 *)D->plight.ptr);
    } else if (o->bounce == VH_FINAL_BOUNCE_HOST) {
        if (P > 0 && !patch_light) {
            set_error("patch_light is required with VH_FINAL_BOUNCE_HOST");
            return VH_EINVAL;
        }
        std::vector<float4> pl((size_t)P * 4);
        for (int64_t i = 0; i < (int64_t)P * 4; i++)
            pl[i] = make_float4(patch_light[3 * i], patch_light[3 * i + 1], patch_light[3 * i + 2], 0);
        if (P > 0)
            VH_HIP(hipMemcpyAsync(D->plight.ptr, pl.data(), pl.size() * sizeof(float4),
                                  hipMemcpyHostToDevice, st));
        VH_WAIT(c, st);
    }
    FinalArgs a;
    a.faces = (const FaceDev *)D->faces.ptr;
    a.luxel_pos = D->have_luxels ? (const float *)D->luxel_pos.ptr : nullptr;
    a.bounds = (const float4 *)D->bounds.ptr;
    a.nbr_ptr = (const int32_t *)D->nbr_ptr.ptr;
    a.nbr = (const int32_t *)D->nbr.ptr;
    a.plist_ptr = (const int32_t *)D->plist_ptr.ptr;
    a.plist = (const int32_t *)D->plist.ptr;
    a.porigin = (const float4 *)D->porigin.ptr;
    a.wptr = (const int32_t *)D->wptr.ptr;
    a.wpts = (const float *)D->wpts.ptr;
    a.plight = (const float4 *)D->plight.ptr;
    a.spos = gv.pos;
    a.lv = o->which ? gv.hdr : gv.ldr;
    a.fstyles = gv.face_styles;
    a.nsamples = gv.nsamples;
    a.units = (const UnitDev *)D->units.ptr;
    a.tiles = (const int2 *)D->tiles.ptr;
    a.lit_units = (const int32_t *)D->lit_faces.ptr;
    a.cells = (float4 *)D->cells.ptr;
    a.pcells = (float4 *)D->pcells.ptr;
    a.mvals = (float4 *)D->mvals.ptr;
    a.lump = (uint32_t *)D->lump.ptr;
    a.bounce = o->bounce != VH_FINAL_NO_BOUNCE;
    a.red_errors = o->red_errors != 0;
    {
        KernelTimer timer(c, VH_KERNEL_FINAL, st);
        if (D->ntiles > 0) {
            hipLaunchKernelGGL(k_radial<false>, dim3((uint32_t)D->ntiles), dim3(kWave), 0, st, a);
            if (a.bounce && D->nbtiles > 0)  // style 0's units only: the first tiles
                hipLaunchKernelGGL(k_radial<true>, dim3((uint32_t)D->nbtiles), dim3(kWave), 0, st, a);
        }
        if (D->nlit > 0)
            hipLaunchKernelGGL(k_final, dim3((uint32_t)D->nlit), dim3(kWave), 0, st, a);
        timer.done();
    }
    VH_HIP(hipGetLastError());
    VH_WAIT(c, st);  // bounded: the lump reaches the caller only once it is made
    hipError_t e = hipSuccess;
    if (D->lump_bytes > 0)
        e = hipMemcpyAsync(lump, D->lump.ptr, (size_t)D->lump_bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        set_error(std::string("final light failed: ") + hipGetErrorString(e));
        return VH_EDEVICE;
    }
    D->have_lump = true;
    return VH_OK;
}

}  // extern "C"
