// The direct-light gather and the RGBE encode on gfx950.
//
// One lane per sample.  Lanes are grouped four by four exactly as the
// reference's SampleGroup (FaceLightJob.EmitGroup, Rad/Light/FaceLightJob.cs:329-398):
// a group is four consecutive samples of one face, a short last group repeats
// its last sample in the padding lanes.  Decisions the reference makes per
// group (the sun's "all four dots are zero", ResolveGroup's "every fxdot is
// zero, skip the light" that also gates the direct-sun amount of every lane)
// are 4-lane shuffles.  Per light, in list order, each lane computes
// DirectLightGatherer's dot and falloff, traces its own visibility or sky rays
// inline through the BVH (vh_traverse.h) and adds (dot * mask) * falloff *
// intensity to its LightingValue per normal (LDR and HDR together: the two
// passes differ only in intensity).
//
// Scheduling: a wave takes a chunk of 16 groups (64 samples).  The persistent
// form launches a fixed number of waves that pull chunks from an atomic
// ticket; the grid form launches one wave per chunk; the per-face form is one
// launch per face, the unbatched baseline.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vh_detmath.h"
#include "vh_rgbe.h"
#include "vh_traverse.h"

#pragma clang fp contract(off)

struct GatherData {
    int32_t nfaces = 0, nsamples = 0, nlights = 0, pvs_bytes = 0;
    // light styles: the lights sorted by style group (group 0 is style 0,
    // then the other styles in ascending number), each group in light order;
    // per face up to four slots (slot 0 style 0, the rest by first touch)
    int32_t ngroups_style = 1, nslots = 1;
    std::vector<int32_t> h_group_style;  // style number of each group
    std::vector<uint8_t> h_face_styles;  // 4 per face, after a run
    int64_t dropped_styles = 0;          // (face, style) pairs past the fourth slot, last run
    uint64_t style_generation = 0;       // changes whenever h_face_styles may have
    vh::DevBuf light_index, group_ptr, first_touch, slot_group, face_styles, sample_face, stage_ldr,
        stage_hdr, dropped;
    int32_t n_sky_ambient = 0;  // lights of type sky ambient (the rest light the direct points)
    int32_t sun_samples = 1, sky_count = 0, ao_count = 0;
    float sun_extent = 0.0f, ao_radius = 0.0f;
    int64_t ngroups = 0, nchunks = 0;
    std::vector<int32_t> h_face_ptr, h_normals;
    std::vector<int64_t> h_face_chunk_ptr;  // chunks of face f: [ptr[f], ptr[f+1])
    vh::DevBuf face_ptr, face_nc, face_normals, pos, cluster, lights, pvs, jitter, sky, ao_dirs;
    vh::DevBuf chunk_face, chunk_first;  // per chunk of 16 groups within one face
    vh::DevBuf out_ldr, out_hdr, out_ao, ticket;
    // supersampling's working set, kept between calls (grown, never shrunk)
    vh::DevBuf ss_frames, ss_lux_ptr, ss_st_in, ss_luxmap, ss_st, ss_inten, ss_processed, ss_flags, ss_counts,
        ss_off, ss_dptr, ss_aptr, ss_dcf, ss_dcs, ss_acf, ss_acs, ss_dpos, ss_dcl, ss_apos, ss_acl, ss_sel_sample,
        ss_sel_face, ss_dl, ss_dh, ss_al, ss_ah, ss_sface;
    bool have_results = false;
};

namespace vh {

void gather_release(vh_ctx *c)
{
    GatherData *G = c->gather;
    if (!G) return;
    for (DevBuf *b : {&G->light_index, &G->group_ptr, &G->first_touch, &G->slot_group, &G->face_styles,
                      &G->sample_face, &G->stage_ldr, &G->stage_hdr, &G->dropped})
        b->release();
    for (DevBuf *b : {&G->face_ptr, &G->face_nc, &G->face_normals, &G->pos, &G->cluster,
                      &G->lights, &G->pvs, &G->jitter, &G->sky, &G->ao_dirs, &G->chunk_face,
                      &G->chunk_first, &G->out_ldr, &G->out_hdr, &G->out_ao, &G->ticket, &G->ss_frames,
                      &G->ss_lux_ptr, &G->ss_st_in, &G->ss_luxmap, &G->ss_st, &G->ss_inten, &G->ss_processed,
                      &G->ss_flags, &G->ss_counts, &G->ss_off, &G->ss_dptr, &G->ss_aptr, &G->ss_dcf, &G->ss_dcs,
                      &G->ss_acf, &G->ss_acs, &G->ss_dpos, &G->ss_dcl, &G->ss_apos, &G->ss_acl,
                      &G->ss_sel_sample, &G->ss_sel_face, &G->ss_dl, &G->ss_dh, &G->ss_al, &G->ss_ah,
                      &G->ss_sface})
        b->release();
    delete G;
    c->gather = nullptr;
}

bool gather_view(vh_ctx *c, GatherView *v)
{
    GatherData *G = c->gather;
    if (!G) return false;
    v->nfaces = G->nfaces;
    v->nsamples = G->nsamples;
    v->face_ptr = (const int32_t *)G->face_ptr.ptr;
    v->face_nc = (const int32_t *)G->face_nc.ptr;
    v->h_face_ptr = &G->h_face_ptr;
    v->h_normals = &G->h_normals;
    v->pos = (const float *)G->pos.ptr;
    v->ldr = (const float4 *)G->out_ldr.ptr;
    v->hdr = (const float4 *)G->out_hdr.ptr;
    v->have_results = G->have_results;
    v->nslots = G->nslots;
    v->face_styles = (const uint8_t *)G->face_styles.ptr;
    v->h_face_styles = &G->h_face_styles;
    v->style_generation = G->style_generation;
    return true;
}

namespace {

constexpr int kWave = 64;
constexpr float kMaxTrace = (float)(1.732050807569 * (2 * 16384));  // LightConstants.MaxTraceLength
constexpr float kDistEpsilon = 0.03125f;                           // LightConstants.DistEpsilon
constexpr float kEqualEpsilon = 0.001f;                            // LightConstants.EqualEpsilon

struct GatherArgs {
    const int32_t *face_ptr;
    const int32_t *face_nc;
    const float *face_normals;  // 12 per face
    const float *pos;
    const int32_t *cluster;
    const vh_light *lights;      // sorted by style group
    const int32_t *light_index;  // each sorted light's index in the caller's list
    const int32_t *group_ptr;    // style group g: sorted lights [group_ptr[g], group_ptr[g + 1])
    int32_t ngroups;
    const int32_t *face_group;   // the resample: per face the one group it lights (-1 none), stride 4
    unsigned long long *first_touch;  // per (face, group): first (group in face, light) to touch it
    int32_t nlights;
    const uint8_t *pvs;
    int32_t pvs_bytes;
    int32_t nsamples;
    float sun_extent;
    int32_t sun_samples;
    const float *jitter;
    int32_t sky_count;
    const float *sky;
    int32_t ao_count;
    float ao_radius;
    const float *ao_dirs;
    const int32_t *chunk_face;
    const int32_t *chunk_first;
    int64_t nchunks;
    float4 *out_ldr;  // per style group, 4 normals x nsamples (the resample: one group)
    float4 *out_hdr;
    float *out_ao;
    const float4 *nodes;
    const float4 *tris;
    unsigned long long *rays;  // rays traced, summed per wave
    unsigned long long *counters;  // COUNT launches: node visits, triangle tests, node and triangle slots
    int32_t light_filter;      // 0 every light; 1 sky ambient only; 2 all but sky ambient
    int32_t no_skip;           // ResolveResample: every light's record is added
    // the 3D skybox (vh_gather_skybox); sb_ncam 0 when there is none
    int32_t sb_ncam, sb_nnodes, sb_nleaves;
    const float4 *sb_cam;        // origin, world_to_sky
    const int4 *sb_nodes;        // children, plane type
    const float4 *sb_planes;     // the node's plane: normal, dist
    const uint8_t *sb_recurse;   // per leaf
};

__device__ __forceinline__ float maxps(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float minps(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx) + (ay * by) + (az * bz);
}

// any lane of this lane's 4-lane group
__device__ __forceinline__ bool group_any(bool v)
{
    int x = v ? 1 : 0;
    x |= __shfl_xor(x, 1, 64);
    x |= __shfl_xor(x, 2, 64);
    return x != 0;
}
__device__ __forceinline__ bool group_all(bool v) { return !group_any(!v); }

__device__ __forceinline__ bool pvs_check(const GatherArgs &a, int32_t light, int32_t cl)
{
    if (cl < 0 || a.pvs_bytes == 0 || cl >= a.pvs_bytes * 8) return true;
    return (a.pvs[(size_t)light * a.pvs_bytes + (cl >> 3)] >> (cl & 7)) & 1;
}

// LightRayLog.MakeRay (correct mode): the whole segment, reach 1.
__device__ __forceinline__ Ray segment(float px, float py, float pz, float ex, float ey, float ez)
{
    Ray r;
    r.ox = px;
    r.oy = py;
    r.oz = pz;
    r.dx = ex - px;
    r.dy = ey - py;
    r.dz = ez - pz;
    r.reach = 1.0f;
    return r;
}

// The lanes that call these together trace together (packet_any and
// packet_closest over the wave's stack `wstk`): the rays of one light, one sun
// or sky direction, or one AO direction of one face leave from the samples of
// one face, so they travel as a packet.
// Traversal work of one lane (counted only in the COUNT instantiation):
// node visits, triangle tests, and the wave's node and triangle slots.
struct Work {
    uint32_t visit = 0, test = 0, slots[2] = {0, 0};
};

template <bool COUNT>
__device__ __forceinline__ bool blocked(const GatherArgs &a, int *wstk, const Ray &r, int &nr, Work &wk)
{
    const bool fin = ray_finite(r);
    nr += fin ? 1 : 0;
    return packet_any<COUNT>(a.nodes, a.tris, wstk, r, 0.0f, fin, &wk.visit, &wk.test, nullptr,
                             COUNT ? wk.slots : nullptr);
}

// LightRayLog.Occlusion: the first hit short of the end, unless it is sky.
template <bool COUNT>
__device__ __forceinline__ float sky_occlusion(const GatherArgs &a, int *wstk, const Ray &r, int &nr,
                                               Work &wk)
{
    const bool fin = ray_finite(r);
    nr += fin ? 1 : 0;
    float t;
    int32_t id;
    const int32_t h = packet_closest<kClosestShort, -1, COUNT>(a.nodes, a.tris, wstk, r, 0.0f, fin, t, id,
                                                               &wk.visit, &wk.test, wk.slots);
    return (fin && h >= 0 && (id & VH_TRACE_ID_SKY) == 0) ? 1.0f : 0.0f;
}

// Whether a sky test starting at p continues into the skyboxes:
// CompiledBspTree.LeafFromPoint (Rad/Light/CompiledBspTree.cs) from node 0 --
// an axial plane (type < 3) compares its one coordinate, any other the dot
// product, and a point on the plane goes to the front child; a child past the
// last node is leaf 0 -- then the leaf's flag (DirectLightGatherer.LaneRecurses,
// DirectLightGatherer.cs:1611-1621: an area of the map with no camera in it).
// The host has checked that no descent meets a node twice, so it ends within
// sb_nnodes steps.
__device__ __forceinline__ bool skybox_recurses(const GatherArgs &a, float px, float py, float pz)
{
    int32_t node = 0;
    for (int32_t step = 0; node >= 0 && step <= a.sb_nnodes; step++) {
        if (node >= a.sb_nnodes) {
            node = -1;  // leaf 0
            break;
        }
        const int4 nd = a.sb_nodes[node];
        const float4 pl = a.sb_planes[node];
        float dist;
        if (nd.z < 3)
            dist = (nd.z == 0 ? px : (nd.z == 1 ? py : pz)) - pl.w;
        else
            dist = dot3(pl.x, pl.y, pl.z, px, py, pz) - pl.w;
        node = dist < 0.0f ? nd.y : nd.x;
    }
    const int32_t leaf = -1 - node;
    return node < 0 && leaf < a.sb_nleaves && a.sb_recurse[leaf] != 0;
}

// TestLineDoesHitSky (DirectLightGatherer.cs:510-714) for this lane's segment
// p -> e: the occlusion, before its clamp to [0, 1].  `want` lanes trace their
// first ray (LightRayLog.Occlusion: a hit short of the end that is not sky);
// with a skybox, a lane whose start recurses (`rec`) in a group whose first
// rays are not all blocked casts one more ray per camera from origin +
// start * world_to_sky along its direction normalised (Vec3.Normalise: three
// divides by the length, a zero vector for a zero length), MaxTraceLength
// long, and adds its answer as the reference does: occlusion + 1 - fraction.
// All four lanes of a group call this together (the group's "all blocked"
// is a shuffle); padding lanes pass want = false, and the group's other lanes
// then decide it, as the padding copies the last real lane.
template <bool COUNT>
__device__ __forceinline__ float sky_test(const GatherArgs &a, int *wstk, bool want, bool rec, float px,
                                          float py, float pz, float ex, float ey, float ez, int &nr, Work &wk)
{
    float occ = 0.0f;
    if (want) occ = sky_occlusion<COUNT>(a, wstk, segment(px, py, pz, ex, ey, ez), nr, wk);
    if (a.sb_ncam == 0) return occ;
    const bool fully = group_all(!want || occ >= 1.0f);
    const bool go = want && rec && !fully;
    if (__builtin_amdgcn_ballot_w64(go) == 0) return occ;
    const float vx = ex - px, vy = ey - py, vz = ez - pz;
    const float len = sqrtf((vx * vx) + (vy * vy) + (vz * vz));
    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
    if (len != 0.0f) {
        ux = vx / len;
        uy = vy / len;
        uz = vz / len;
    }
    for (int32_t k = 0; k < a.sb_ncam; k++) {
        const float4 cam = a.sb_cam[k];
        const float sx = cam.x + (px * cam.w), sy = cam.y + (py * cam.w), sz = cam.z + (pz * cam.w);
        const float tx = sx + (ux * kMaxTrace), ty = sy + (uy * kMaxTrace), tz = sz + (uz * kMaxTrace);
        float inner = 0.0f;
        if (go) inner = sky_occlusion<COUNT>(a, wstk, segment(sx, sy, sz, tx, ty, tz), nr, wk);
        if (go) {
            const float fraction = 1.0f - minps(maxps(inner, 0.0f), 1.0f);
            occ = occ + 1.0f - fraction;
        }
    }
    return occ;
}

__device__ __forceinline__ float inverse_quadratic(const vh_light &l, float d)
{
    float f = d * d;
    f *= l.quadratic_attn;
    f += l.linear_attn * d;
    f += l.constant_attn;
    return 1.0f / f;
}

// One wave lights one chunk: 16 groups of 4 samples of one face.  NN is the
// face's normal count (1, or 4 when bumped), so every per-normal array is in
// registers.
template <int NN, bool COUNT>
__device__ void light_chunk_n(const GatherArgs &a, int64_t chunk, int *stk)
{
    Work wk;
    int lane = threadIdx.x & 63;
    int sub = lane & 3;
    int32_t f = a.chunk_face[chunk];
    int32_t gfirst = a.chunk_first[chunk] + (lane >> 2) * 4;  // this lane's group
    int32_t fend = a.face_ptr[f + 1];
    bool group_live = gfirst < fend;
    int count = group_live ? min(4, fend - gfirst) : 1;
    bool real = group_live && sub < count;
    int32_t s = group_live ? gfirst + min(sub, count - 1) : a.face_ptr[f];
    constexpr int normals = NN;
    const float *fnp = a.face_normals + (size_t)f * 12;
    float n0x = fnp[0], n0y = fnp[1], n0z = fnp[2];
    float bn[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) bn[k][c] = fnp[3 + 3 * k + c];
    // ComputeIlluminationPointAndNormals: one unit off the face
    float px = a.pos[3 * (size_t)s] + n0x;
    float py = a.pos[3 * (size_t)s + 1] + n0y;
    float pz = a.pos[3 * (size_t)s + 2] + n0z;
    int32_t cl = a.cluster[s];
    // whether this lane's sky tests continue into the skyboxes (its start is
    // its point: FaceLightJob passes no epsilon, FaceLightJob.cs:392,514)
    const bool rec = a.sb_ncam > 0 && skybox_recurses(a, px, py, pz);

    int nr = 0;
    // this group's index in its face: the first-touch keys order groups
    const int32_t grp = group_live ? (gfirst - a.face_ptr[f]) >> 2 : 0;
    // style groups in turn (one, for the resample); within a group the
    // lights keep their order, so each style slot sums as the reference does
    int32_t g_lo = 0, g_hi = a.ngroups;
    if (a.face_group) {
        g_lo = a.face_group[(size_t)f * 4];
        g_hi = g_lo + 1;
    }
    for (int32_t g = g_lo; g < g_hi; g++) {
    float4 acc_l[4], acc_h[4];
    for (int n = 0; n < 4; n++) {
        acc_l[n] = make_float4(0, 0, 0, 0);
        acc_h[n] = make_float4(0, 0, 0, 0);
    }
    unsigned long long touch = ~0ull;
    const int32_t k_lo = g >= 0 ? a.group_ptr[g] : 0, k_hi = g >= 0 ? a.group_ptr[g + 1] : 0;
    for (int32_t k = k_lo; k < k_hi; k++) {
        const int32_t li = a.light_index[k];
        const vh_light &L = a.lights[k];
        bool inmask = real && pvs_check(a, li, cl);
        if (!group_any(inmask)) continue;
        // EmitResample (FaceLightJob.cs:474-482): ambient points see only sky
        // ambient, direct points every other light
        if (a.light_filter != 0 &&
            ((a.light_filter == 1) != (L.type == VH_EMIT_SKYAMBIENT)))
            continue;
        float dot[4] = {0, 0, 0, 0};
        float falloff = 0.0f, sun = 0.0f;
        if (L.type == VH_EMIT_POINT || L.type == VH_EMIT_SURFACE || L.type == VH_EMIT_SPOTLIGHT) {
            // GatherStandardLane (DirectLightGatherer.cs:1169-1306)
            float sx = L.origin[0], sy = L.origin[1], sz = L.origin[2];
            float dx = sx - px, dy = sy - py, dz = sz - pz;
            float dist2 = (dx * dx) + (dy * dy) + (dz * dz);
            float rpc = 1.0f / sqrtf(dist2);
            dx = dx * rpc;
            dy = dy * rpc;
            dz = dz * rpc;
            float dist = sqrtf(dist2);
            float d0 = maxps(0.0f, dot3(dx, dy, dz, n0x, n0y, n0z));
            bool hard = L.end_fade > L.start_fade;
            bool ok = !(hard && !(dist <= L.end_fade));
            dist = maxps(dist, 1.0f);
            float eval = minps(dist, L.cap_dist);
            float fo = 0.0f;
            if (ok) {
                if (L.type == VH_EMIT_POINT) {
                    fo = inverse_quadratic(L, eval);
                } else if (L.type == VH_EMIT_SURFACE) {
                    float dot2 = maxps(0.0f, -dot3(dx, dy, dz, L.normal[0], L.normal[1], L.normal[2]));
                    if (d0 == 0.0f) ok = false;
                    fo = dot2 / dist2;
                    sx = sx + (L.normal[0] * kDistEpsilon);
                    sy = sy + (L.normal[1] * kDistEpsilon);
                    sz = sz + (L.normal[2] * kDistEpsilon);
                } else {
                    float dot2 = -dot3(dx, dy, dz, L.normal[0], L.normal[1], L.normal[2]);
                    if (!(dot2 > L.stopdot2)) {
                        ok = false;
                    } else {
                        fo = inverse_quadratic(L, eval) * dot2;
                        bool fringe = dot2 <= L.stopdot;
                        float mult = (dot2 - L.stopdot2) / (L.stopdot - L.stopdot2);
                        mult = minps(mult, 1.0f);
                        mult = maxps(mult, 0.0f);
                        // DetMathF.Pow, correctly rounded
                        if (L.exponent != 0.0f && L.exponent != 1.0f)
                            mult = detmath::pow_f32(mult, L.exponent);
                        mult = fringe ? mult : 1.0f;
                        fo = mult * fo;
                    }
                }
            }
            if (ok && hard) {
                float t = (dist - L.start_fade) / (L.end_fade - L.start_fade);
                t = minps(t, 1.0f);
                t = maxps(t, 0.0f);
                t = 1.0f - t;
                float mult = (6.0f * t) - 15.0f;
                mult = (mult * t) + 10.0f;
                mult = (t * t) * mult;
                mult = t * mult;
                fo = mult * fo;
            }
            if (ok) {
                dot[0] = d0;
                falloff = fo;
                _Pragma("unroll") for (int n = 1; n < normals; n++)
                    dot[n] = maxps(0.0f, dot3(bn[n - 1][0], bn[n - 1][1], bn[n - 1][2], dx, dy, dz));
                if (inmask && d0 != 0.0f && fo != 0.0f) {
                    if (blocked<COUNT>(a, stk, segment(px, py, pz, sx, sy, sz), nr, wk)) dot[0] = 0.0f * dot[0];
                }
            }
        } else if (L.type == VH_EMIT_SKYLIGHT) {
            // EmitSkyLight / ResolveSkyLight (DirectLightGatherer.cs:716-845)
            float dl = maxps(-dot3(n0x, n0y, n0z, L.normal[0], L.normal[1], L.normal[2]), 0.0f);
            if (!group_all(dl == 0.0f)) {
                float js = (float)(1.732050807569 * (2 * 16384) * (double)a.sun_extent);
                float total = 0.0f;
                for (int d = 0; d < a.sun_samples; d++) {
                    float ex = L.normal[0] * -kMaxTrace, ey = L.normal[1] * -kMaxTrace,
                          ez = L.normal[2] * -kMaxTrace;
                    if (d != 0) {
                        const float *jt = a.jitter + 3 * (size_t)(d - 1);
                        ex = ex + (jt[0] * js);
                        ey = ey + (jt[1] * js);
                        ez = ez + (jt[2] * js);
                    }
                    // EmitSkyLight traces every real lane (DirectLightGatherer.cs:780-790)
                    const float occ =
                        sky_test<COUNT>(a, stk, real, rec, px, py, pz, ex + px, ey + py, ez + pz, nr, wk);
                    total += 1.0f - minps(maxps(occ, 0.0f), 1.0f);
                }
                float see = total * (1.0f / (float)a.sun_samples);
                dot[0] = dl * see;
                falloff = 1.0f;
                sun = see * 10000.0f;
                _Pragma("unroll") for (int n = 1; n < normals; n++)
                    dot[n] = (-dot3(bn[n - 1][0], bn[n - 1][1], bn[n - 1][2], L.normal[0],
                                    L.normal[1], L.normal[2])) *
                             see;
            }
        } else if (L.type == VH_EMIT_SKYAMBIENT) {
            // AmbientSky (DirectLightGatherer.cs:847-1012)
            float sumdot = 0.0f;
            float amb[4] = {0, 0, 0, 0}, poss[4] = {0, 0, 0, 0};
            for (int32_t k = 0; k < a.sky_count; k++) {
                float ax = a.sky[3 * k], ay = a.sky[3 * k + 1], az = a.sky[3 * k + 2];
                float dd = -dot3(n0x, n0y, n0z, ax, ay, az);
                bool valid = dd > kEqualEpsilon;  // the same for every lane of a flat face
                if (!group_any(valid)) continue;
                float dts[4];
                dts[0] = valid ? dd : 0.0f;
                sumdot = dts[0] + sumdot;
                poss[0] = (valid ? 1.0f : 0.0f) + poss[0];
                bool tr = valid;
                _Pragma("unroll") for (int n = 1; n < normals; n++) {
                    float d2 = -dot3(bn[n - 1][0], bn[n - 1][1], bn[n - 1][2], ax, ay, az);
                    bool v2 = d2 > kEqualEpsilon;
                    dts[n] = v2 ? d2 : 0.0f;
                    poss[n] = (valid && v2 ? 1.0f : 0.0f) + poss[n];
                    tr = tr || v2;
                }
                // EmitAmbientRays / ResolveAmbientRays (DirectLightGatherer.cs:1014-1069):
                // with a skybox every real lane is traced through the sky test
                // (the recursion couples the group's lanes); without one only a
                // lane whose answer can count
                const bool want = a.sb_ncam > 0 ? real : (real && tr && inmask);
                const float occ = sky_test<COUNT>(a, stk, want, rec, px, py, pz, (ax * -kMaxTrace) + px,
                                                  (ay * -kMaxTrace) + py, (az * -kMaxTrace) + pz, nr, wk);
                const float frac = want ? 1.0f - minps(maxps(occ, 0.0f), 1.0f) : 1.0f;
                _Pragma("unroll") for (int n = 0; n < normals; n++) amb[n] = amb[n] + (frac * dts[n]);
            }
            falloff = 1.0f;
            _Pragma("unroll") for (int n = 0; n < normals; n++) {
                float factor = poss[n] / poss[0];
                dot[n] = amb[n] / (factor * sumdot);
            }
        } else {
            continue;
        }
        // Resolve: a light behind the face lights no bump direction
        float d0 = maxps(dot[0], 0.0f);
        dot[0] = d0;
        _Pragma("unroll") for (int n = 1; n < normals; n++) dot[n] = d0 > 0.0f ? maxps(dot[n], 0.0f) : 0.0f;
        // ResolveGroup: (dot * mask) * falloff; a group whose every fxdot is 0 skips the light
        float fx[4];
        bool nz = false;
        _Pragma("unroll") for (int n = 0; n < normals; n++) {
            float masked = inmask ? dot[n] : 0.0f;
            fx[n] = masked * falloff;
            nz = nz || fx[n] != 0.0f;
        }
        if (!a.no_skip && !group_any(nz)) continue;
        // FindOrAllocateStyle's order: the first (group, light) to touch it
        if (touch == ~0ull) touch = (unsigned long long)grp * (unsigned long long)a.nlights + (unsigned long long)li;
        if (real) {
            _Pragma("unroll") for (int n = 0; n < normals; n++) {
                acc_l[n].x = acc_l[n].x + (fx[n] * L.intensity[0]);
                acc_l[n].y = acc_l[n].y + (fx[n] * L.intensity[1]);
                acc_l[n].z = acc_l[n].z + (fx[n] * L.intensity[2]);
                acc_l[n].w += sun;
                acc_h[n].x = acc_h[n].x + (fx[n] * L.intensity_hdr[0]);
                acc_h[n].y = acc_h[n].y + (fx[n] * L.intensity_hdr[1]);
                acc_h[n].z = acc_h[n].z + (fx[n] * L.intensity_hdr[2]);
                acc_h[n].w += sun;
            }
        }
    }
    // every sample's value is written, zeros included (no clearing pass)
    if (real && g >= 0) {
        const size_t og = a.face_group ? 0 : (size_t)g;
        _Pragma("unroll") for (int n = 0; n < 4; n++) {
            a.out_ldr[(og * 4 + n) * a.nsamples + s] = acc_l[n];
            a.out_hdr[(og * 4 + n) * a.nsamples + s] = acc_h[n];
        }
    }
    if (a.first_touch && g >= 0) {
        // one atomic per wave and group: the wave's groups are one face's
        for (int o = 32; o > 0; o >>= 1) {
            unsigned long long v = __shfl_xor(touch, o, 64);
            touch = v < touch ? v : touch;
        }
        if (lane == 0 && touch != ~0ull) atomicMin(&a.first_touch[(size_t)f * a.ngroups + g], touch);
    }
    }
    if (real && a.face_group && a.face_group[(size_t)f * 4] < 0) {
        // a resample face without the style: its points gather nothing
        _Pragma("unroll") for (int n = 0; n < 4; n++) {
            a.out_ldr[(size_t)n * a.nsamples + s] = make_float4(0, 0, 0, 0);
            a.out_hdr[(size_t)n * a.nsamples + s] = make_float4(0, 0, 0, 0);
        }
    }
    if (real && a.ao_count > 0) {
        int unblocked = 0;
        const float *dirs = a.ao_dirs + (size_t)f * a.ao_count * 3;
        for (int32_t k = 0; k < a.ao_count; k++) {
            Ray r;
            r.ox = px;
            r.oy = py;
            r.oz = pz;
            r.dx = dirs[3 * k] * a.ao_radius;
            r.dy = dirs[3 * k + 1] * a.ao_radius;
            r.dz = dirs[3 * k + 2] * a.ao_radius;
            r.reach = 1.0f;
            if (!blocked<COUNT>(a, stk, r, nr, wk)) unblocked++;
        }
        a.out_ao[s] = (float)unblocked / (float)a.ao_count;
    }
    // every lane of the wave is here: one atomic per chunk
    for (int o = 32; o > 0; o >>= 1) nr += __shfl_xor(nr, o, 64);
    if (lane == 0) atomicAdd(a.rays, (unsigned long long)nr);
    if (COUNT) {
        unsigned long long v[4] = {wk.visit, wk.test, wk.slots[0], wk.slots[1]};
        for (int q = 0; q < 4; q++) {
            for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        }
        // slots are per wave (every lane holds the same count)
        if (lane == 0) {
            atomicAdd(&a.counters[0], v[0]);
            atomicAdd(&a.counters[1], v[1]);
            atomicAdd(&a.counters[2], v[2] / 64);
            atomicAdd(&a.counters[3], v[3] / 64);
        }
    }
}

template <bool COUNT>
__device__ __forceinline__ void light_chunk(const GatherArgs &a, int64_t chunk, int *stk)
{
    if (a.face_nc[a.chunk_face[chunk]] == 4)
        light_chunk_n<4, COUNT>(a, chunk, stk);
    else
        light_chunk_n<1, COUNT>(a, chunk, stk);
}

// PERSISTENT: each wave takes `grain` consecutive chunks per ticket until the
// queue drains (every wave reads the ticket past the end and leaves).
// Five waves per SIMD: the walk is latency-bound, and at the register count
// the accumulators want (143 VGPRs, three waves) the batching config ran 12 %
// slower than here at 96 VGPRs, where a few dozen values spill to scratch.
template <bool PERSISTENT, bool COUNT = false>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(5))) void k_facelight(
    GatherArgs a, int64_t first_chunk, unsigned long long *__restrict__ ticket, int32_t grain = 1)
{
    __shared__ int stk[3 * kMaxDepth];  // the wave's traversal stack
    if (!PERSISTENT) {
        light_chunk<COUNT>(a, first_chunk + blockIdx.x, stk);
        return;
    }
    for (;;) {
        unsigned long long c = 0;
        if (threadIdx.x == 0) c = atomicAdd(ticket, (unsigned long long)grain);
        c = __shfl(c, 0, 64);
        if ((int64_t)c >= a.nchunks) break;
        for (int32_t j = 0; j < grain && (int64_t)c + j < a.nchunks; j++)
            light_chunk<COUNT>(a, (int64_t)c + j, stk);
    }
}

// FaceLight.FindOrAllocateStyle over a whole face at once (FaceLight.cs:204-222,
// FaceLightJob.cs:185,435): slot 0 is style 0; the other styles that touched
// the face take slots 1-3 in the order of their first touch (group in the
// face, then light); a fifth is dropped.  A thread per face.
__global__ void k_assign_slots(const unsigned long long *__restrict__ first_touch, int32_t nfaces,
                               int32_t ngroups, const int32_t *__restrict__ group_style,
                               int32_t *__restrict__ slot_group, uint8_t *__restrict__ face_styles,
                               unsigned long long *__restrict__ dropped)
{
    const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfaces) return;
    const unsigned long long *ft = first_touch + (size_t)f * ngroups;
    int32_t sg[4] = {0, -1, -1, -1};
    unsigned long long last = 0;  // keys are distinct: each (group, light) touches one style
    int k = 1;
    for (; k < 4; k++) {
        unsigned long long best = ~0ull;
        int32_t bg = -1;
        for (int32_t g = 1; g < ngroups; g++) {
            const unsigned long long v = ft[g];
            if (v != ~0ull && (k == 1 || v > last) && v < best) {
                best = v;
                bg = g;
            }
        }
        if (bg < 0) break;
        sg[k] = bg;
        last = best;
    }
    int32_t more = 0;
    if (k == 4)
        for (int32_t g = 1; g < ngroups; g++) more += ft[g] != ~0ull && ft[g] > last;
    if (more) atomicAdd(dropped, (unsigned long long)more);
    for (int j = 0; j < 4; j++) {
        slot_group[(size_t)f * 4 + j] = sg[j];
        face_styles[(size_t)f * 4 + j] = sg[j] < 0 ? (uint8_t)255 : (uint8_t)group_style[sg[j]];
    }
}

// Per-group light into the faces' slots: thread per (slot, normal, sample).
__global__ void k_fill_slots(const float4 *__restrict__ stage_l, const float4 *__restrict__ stage_h,
                             const int32_t *__restrict__ slot_group, const int32_t *__restrict__ sample_face,
                             int32_t nsamples, int32_t nslots, float4 *__restrict__ out_l,
                             float4 *__restrict__ out_h)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)4 * nsamples;
    if (i >= per * nslots) return;
    const int32_t k = (int32_t)(i / per), n = (int32_t)((i / nsamples) % 4);
    const int32_t s = (int32_t)(i % nsamples);
    const int32_t g = slot_group[(size_t)sample_face[s] * 4 + k];
    float4 l = make_float4(0, 0, 0, 0), h = l;
    if (g >= 0) {
        l = stage_l[((size_t)g * 4 + n) * nsamples + s];
        h = stage_h[((size_t)g * 4 + n) * nsamples + s];
    }
    out_l[i] = l;
    out_h[i] = h;
}

__global__ void k_encode(const float *__restrict__ rgb, int64_t n, int32_t stride,
                         uint32_t *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *v = rgb + i * stride;
    out[i] = encode_one(v[0], v[1], v[2]);
}

// The lump of one gather: thread per (face sample, normal).
__global__ void k_encode_lump(const float4 *__restrict__ lv, int32_t nsamples,
                              const int32_t *__restrict__ face_of, const int32_t *__restrict__ face_ptr,
                              const int32_t *__restrict__ face_nc,
                              const int64_t *__restrict__ face_base, uint32_t *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nsamples * 4) return;
    int32_t s = (int32_t)(i % nsamples);
    int32_t b = (int32_t)(i / nsamples);
    int32_t f = face_of[s];
    if (b >= face_nc[f]) return;
    int32_t j = s - face_ptr[f];
    int32_t ns = face_ptr[f + 1] - face_ptr[f];
    float4 v = lv[(size_t)b * nsamples + s];
    out[face_base[f] / 4 + (int64_t)b * ns + j] = encode_one(v.x, v.y, v.z);
}

template <typename T>
int upload(DevBuf &b, const T *host, size_t count, const char *what, hipStream_t s)
{
    int rc = b.ensure(count * sizeof(T) + 16, what);
    if (rc) return rc;
    if (count) VH_HIP(hipMemcpyAsync(b.ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    return VH_OK;
}

GatherArgs gather_args(vh_ctx *c)
{
    GatherData *G = c->gather;
    GatherArgs a;
    a.face_ptr = (const int32_t *)G->face_ptr.ptr;
    a.face_nc = (const int32_t *)G->face_nc.ptr;
    a.face_normals = (const float *)G->face_normals.ptr;
    a.pos = (const float *)G->pos.ptr;
    a.cluster = (const int32_t *)G->cluster.ptr;
    a.lights = (const vh_light *)G->lights.ptr;
    a.light_index = (const int32_t *)G->light_index.ptr;
    a.group_ptr = (const int32_t *)G->group_ptr.ptr;
    a.ngroups = G->ngroups_style;
    a.face_group = nullptr;
    a.first_touch = G->ngroups_style > 1 ? (unsigned long long *)G->first_touch.ptr : nullptr;
    a.nlights = G->nlights;
    a.pvs = (const uint8_t *)G->pvs.ptr;
    a.pvs_bytes = G->pvs_bytes;
    a.nsamples = G->nsamples;
    a.sun_extent = G->sun_extent;
    a.sun_samples = G->sun_samples;
    a.jitter = (const float *)G->jitter.ptr;
    a.sky_count = G->sky_count;
    a.sky = (const float *)G->sky.ptr;
    a.ao_count = G->ao_count;
    a.ao_radius = G->ao_radius;
    a.ao_dirs = (const float *)G->ao_dirs.ptr;
    a.chunk_face = (const int32_t *)G->chunk_face.ptr;
    a.chunk_first = (const int32_t *)G->chunk_first.ptr;
    a.nchunks = G->nchunks;
    // one style group: straight into slot 0; else per group, then into slots
    a.out_ldr = (float4 *)(G->ngroups_style > 1 ? G->stage_ldr.ptr : G->out_ldr.ptr);
    a.out_hdr = (float4 *)(G->ngroups_style > 1 ? G->stage_hdr.ptr : G->out_hdr.ptr);
    a.out_ao = (float *)G->out_ao.ptr;
    a.nodes = (const float4 *)c->nodes.ptr;
    a.tris = (const float4 *)c->tris.ptr;
    a.rays = (unsigned long long *)G->ticket.ptr + 1;
    a.counters = (unsigned long long *)c->counters.ptr;
    a.light_filter = 0;
    a.no_skip = 0;
    a.sb_ncam = c->sb_ncam;
    a.sb_nnodes = c->sb_nnodes;
    a.sb_nleaves = c->sb_nleaves;
    a.sb_cam = (const float4 *)c->sb_cam.ptr;
    a.sb_nodes = (const int4 *)c->sb_nodes.ptr;
    a.sb_planes = (const float4 *)c->sb_planes.ptr;
    a.sb_recurse = (const uint8_t *)c->sb_recurse.ptr;
    return a;
}

// ---------------------------------------------------------------------------
// Supersampling (FaceLightJob's StyleSupersample passes, FaceLightJob.cs:698-1207)

constexpr float kSupersampleGradient = 0.0625f;  // LightConstants.SupersampleGradient

// ComputeLuxelIntensity: R + G + B to a linear perception space,
// DetMath.PowToSingle(I / 256.0, 1.0 / 2.2)
__device__ __forceinline__ float ss_intensity(float4 v)
{
    float i = (v.x + v.y) + v.z;
    return detmath::pow_to_single((double)i / 256.0, 1.0 / 2.2);
}

// One wave per face: its luxel -> sample map, each sample's luxel and its
// intensity per normal.
__global__ __launch_bounds__(kWave) void k_ss_init(const int32_t *__restrict__ face_ptr,
                                                   const int32_t *__restrict__ face_nc,
                                                   const vh_final_face *__restrict__ frames,
                                                   const int64_t *__restrict__ lux_ptr,
                                                   const int32_t *__restrict__ st_in,
                                                   int32_t *__restrict__ luxmap,
                                                   int32_t *__restrict__ st,
                                                   int32_t *__restrict__ sface,
                                                   const float4 *__restrict__ light,
                                                   float *__restrict__ inten, int32_t S)
{
    const int32_t f = blockIdx.x, lane = threadIdx.x;
    const int32_t a = face_ptr[f], b = face_ptr[f + 1], nc = face_nc[f];
    const int32_t w = frames[f].width, h = frames[f].height;
    const int64_t base = lux_ptr[f];
    for (int64_t j = lane; j < (int64_t)w * h; j += kWave) luxmap[base + j] = -1;
    __syncthreads();
    for (int32_t s = a + lane; s < b; s += kWave) {
        int32_t ss, tt;
        if (st_in) {
            ss = st_in[2 * (size_t)s];
            tt = st_in[2 * (size_t)s + 1];
        } else {
            ss = (s - a) % w;
            tt = (s - a) / w;
        }
        st[2 * (size_t)s] = ss;
        st[2 * (size_t)s + 1] = tt;
        sface[s] = f;
        luxmap[base + ss + (int64_t)tt * w] = s;
        for (int n = 0; n < nc; n++) inten[(size_t)n * S + s] = ss_intensity(light[(size_t)n * S + s]);
    }
}

// ComputeGradients + SelectForPass, a thread per sample: flags the selected
// samples, marks them processed and counts them per face (counts zeroed
// before the launch; a count does not depend on the order of the adds).
__global__ void k_ss_select(const int32_t *__restrict__ sface, const int32_t *__restrict__ face_nc,
                            const vh_final_face *__restrict__ frames, const int64_t *__restrict__ lux_ptr,
                            const int32_t *__restrict__ luxmap, const int32_t *__restrict__ st,
                            const float *__restrict__ inten, int32_t S, uint8_t *__restrict__ processed,
                            uint8_t *__restrict__ flags, int32_t *__restrict__ counts, int32_t garbage,
                            const int32_t *__restrict__ face_group)
{
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int32_t f = sface[s], nc = face_nc[f];
    // a face without this style slot has no StyleSupersample for it
    if (face_group[(size_t)f * 4] < 0) {
        flags[s] = 0;
        return;
    }
    const int32_t w = frames[f].width, h = frames[f].height;
    const int64_t base = lux_ptr[f];
    // the neighbours in the order Step visits them
    const int ds[8] = {-1, 0, 1, -1, 0, 1, -1, 1};
    const int dt[8] = {-1, -1, -1, 1, 1, 1, 0, 0};
    {
        bool sel = false;
        if (!processed[s]) {
            const int32_t ss = st[2 * (size_t)s], tt = st[2 * (size_t)s + 1];
            float g = 0.0f;
            for (int n = 0; n < nc; n++) {
                const float c = inten[(size_t)n * S + s];
                for (int k = 0; k < 8; k++) {
                    const int32_t x = ss + ds[k], y = tt + dt[k];
                    if (x < 0 || x >= w || y < 0 || y >= h) continue;
                    const int32_t j = luxmap[base + x + (int64_t)y * w];
                    if (j < 0) {
                        if (garbage) g = __builtin_inff();
                        continue;
                    }
                    const float v = fabsf(c - inten[(size_t)n * S + j]);
                    g = g > v ? g : v;
                }
            }
            sel = !(g < kSupersampleGradient);
            if (sel) processed[s] = 1;
        }
        flags[s] = sel ? 1 : 0;
        if (sel) atomicAdd(&counts[f], 1);
    }
}

// EmitSupersamplePoint's points, one wave per face: item i (the i-th selected
// sample in sample order) owns direct points 16 i .. 16 i + 15 (row-major,
// 4 x 4) and ambient points 4 i .. 4 i + 3.
__global__ __launch_bounds__(kWave) void k_ss_points(const int32_t *__restrict__ face_ptr,
                                                     const vh_final_face *__restrict__ frames,
                                                     const uint8_t *__restrict__ flags,
                                                     const int32_t *__restrict__ item_off,
                                                     const float *__restrict__ pos,
                                                     const int32_t *__restrict__ cluster,
                                                     float *__restrict__ dpos, int32_t *__restrict__ dcl,
                                                     float *__restrict__ apos, int32_t *__restrict__ acl,
                                                     int32_t *__restrict__ sel_sample,
                                                     int32_t *__restrict__ sel_face)
{
    const int32_t f = blockIdx.x, lane = threadIdx.x;
    int32_t run = item_off[f];
    if (item_off[f + 1] == run) return;
    const int32_t a = face_ptr[f], b = face_ptr[f + 1];
    const vh_final_face &F = frames[f];
    const float a4[4] = {-0.375f, -0.125f, 0.125f, 0.375f};  // csshift + c * cscale, 4 wide
    const float a2[2] = {-0.25f, 0.25f};                      // 2 wide
    for (int32_t first = a; first < b; first += kWave) {
        const int32_t s = first + lane;
        const bool sel = s < b && flags[s];
        const unsigned long long m = __ballot(sel);
        if (sel) {
            const int32_t item = run + __popcll(m & ((1ull << lane) - 1ull));
            sel_sample[item] = s;
            sel_face[item] = f;
            const int32_t cl = cluster[s];
            // WorldToLuxel of the sample
            const float px = pos[3 * (size_t)s] - F.origin[0], py = pos[3 * (size_t)s + 1] - F.origin[1],
                        pz = pos[3 * (size_t)s + 2] - F.origin[2];
            const float os = (((px * F.w2l_s[0]) + (py * F.w2l_s[1])) + (pz * F.w2l_s[2])) - (float)F.min_s;
            const float ot = (((px * F.w2l_t[0]) + (py * F.w2l_t[1])) + (pz * F.w2l_t[2])) - (float)F.min_t;
            for (int p = 0; p < 20; p++) {
                float cs, ct;
                if (p < 16) {
                    cs = os + a4[p >> 2];
                    ct = ot + a4[p & 3];
                } else {
                    cs = os + a2[(p - 16) >> 1];
                    ct = ot + a2[(p - 16) & 1];
                }
                // LuxelToWorld: the two multiply-adds in turn
                cs = cs + (float)F.min_s;
                ct = ct + (float)F.min_t;
                float q[3];
                for (int k = 0; k < 3; k++) {
                    q[k] = F.origin[k] + (F.l2w_s[k] * cs);
                    q[k] = q[k] + (F.l2w_t[k] * ct);
                }
                float *o = p < 16 ? dpos + 3 * ((size_t)16 * item + p) : apos + 3 * ((size_t)4 * item + (p - 16));
                o[0] = q[0];
                o[1] = q[1];
                o[2] = q[2];
                if (p < 16)
                    dcl[(size_t)16 * item + p] = cl;
                else
                    acl[(size_t)4 * item + (p - 16)] = cl;
            }
        }
        run += __popcll(m);
    }
}

// ResolveSupersampleItem: the points summed in order, direct / 16 + ambient / 4,
// written over the sample's light; its intensity recomputed.
// A kind with no light is not launched: its points would gather exactly +0,
// so its sum is taken as +0 (have_d / have_a 0).
__global__ void k_ss_resolve(int32_t total, const int32_t *__restrict__ sel_sample,
                             const int32_t *__restrict__ sel_face, const int32_t *__restrict__ face_nc,
                             const float4 *__restrict__ dout, const float4 *__restrict__ aout, int32_t have_d,
                             int32_t have_a, float4 *__restrict__ light, float *__restrict__ inten, int32_t S)
{
    const int32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= total) return;
    const int32_t s = sel_sample[item], nc = face_nc[sel_face[item]];
    const size_t nd = (size_t)16 * total, na = (size_t)4 * total;
    for (int n = 0; n < nc; n++) {
        float4 d = make_float4(0, 0, 0, 0), m = make_float4(0, 0, 0, 0);
        for (int p = 0; p < (have_d ? 16 : 0); p++) {
            const float4 x = dout[n * nd + (size_t)16 * item + p];
            d.x = d.x + x.x;
            d.y = d.y + x.y;
            d.z = d.z + x.z;
            d.w = d.w + x.w;
        }
        for (int p = 0; p < (have_a ? 4 : 0); p++) {
            const float4 x = aout[n * na + (size_t)4 * item + p];
            m.x = m.x + x.x;
            m.y = m.y + x.y;
            m.z = m.z + x.z;
            m.w = m.w + x.w;
        }
        float4 v;
        v.x = (0.0f + (0.0625f * d.x)) + (0.25f * m.x);
        v.y = (0.0f + (0.0625f * d.y)) + (0.25f * m.y);
        v.z = (0.0f + (0.0625f * d.z)) + (0.25f * m.z);
        v.w = (0.0f + (0.0625f * d.w)) + (0.25f * m.w);
        light[(size_t)n * S + s] = v;
        inten[(size_t)n * S + s] = ss_intensity(v);
    }
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

int vh_gather_load(vh_ctx *c, const vh_gather_desc *d)
{
    clear_error();
    if (!c || !d) {
        set_error("NULL context or descriptor");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int32_t F = d->nfaces, S = d->nsamples, NL = d->nlights;
    if (F < 0 || S < 0 || NL < 0 || d->pvs_bytes < 0 || (F > 0 && (!d->face_sample_ptr ||
        !d->face_normal_count || !d->face_normals)) || (S > 0 && (!d->sample_pos ||
        !d->sample_cluster)) || (NL > 0 && !d->lights) || (d->pvs_bytes > 0 && NL > 0 &&
        !d->light_pvs)) {
        set_error("incomplete gather descriptor");
        return VH_EINVAL;
    }
    if (F > 0 && (d->face_sample_ptr[0] != 0 || d->face_sample_ptr[F] != S)) {
        set_error("face_sample_ptr must run from 0 to nsamples");
        return VH_EINVAL;
    }
    for (int32_t f = 0; f < F; f++) {
        if (d->face_sample_ptr[f + 1] < d->face_sample_ptr[f]) {
            set_error("face_sample_ptr is not monotone");
            return VH_EINVAL;
        }
        if (d->face_normal_count[f] != 1 && d->face_normal_count[f] != 4) {
            set_error("a face has 1 normal, or 4 when bumped");
            return VH_EINVAL;
        }
    }
    for (int32_t l = 0; l < NL; l++) {
        const vh_light &L = d->lights[l];
        if (L.style < 0 || L.style > 254) {
            set_error("light " + std::to_string(l) + " has style " + std::to_string(L.style) +
                      "; a light style runs from 0 to 254 (255 marks an empty slot)");
            return VH_EINVAL;
        }
        if (L.type != VH_EMIT_POINT && L.type != VH_EMIT_SURFACE && L.type != VH_EMIT_SPOTLIGHT &&
            L.type != VH_EMIT_SKYLIGHT && L.type != VH_EMIT_SKYAMBIENT) {
            set_error("light " + std::to_string(l) + " has an unsupported type");
            return VH_ENOTSUP;
        }
    }
    int32_t ns = d->sun_samples > 0 ? d->sun_samples : 1;
    if (ns > 1 && !d->sun_jitter) {
        set_error("sun_samples > 1 needs sun_jitter");
        return VH_EINVAL;
    }
    if ((d->sky_dirs_count > 0 && !d->sky_dirs) || (d->ao_count > 0 && !d->ao_dirs)) {
        set_error("missing sky or ambient-occlusion directions");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    gather_release(c);
    GatherData *G = new GatherData();
    c->gather = G;
    G->nfaces = F;
    G->nsamples = S;
    G->nlights = NL;
    for (int32_t l = 0; l < NL; l++) G->n_sky_ambient += d->lights[l].type == VH_EMIT_SKYAMBIENT ? 1 : 0;
    G->pvs_bytes = NL > 0 ? d->pvs_bytes : 0;
    G->sun_extent = d->sun_extent;
    G->sun_samples = ns;
    G->sky_count = d->sky_dirs_count > 0 ? d->sky_dirs_count : 0;
    G->ao_count = d->ao_count > 0 ? d->ao_count : 0;
    G->ao_radius = d->ao_radius;
    G->h_face_ptr.assign(d->face_sample_ptr, d->face_sample_ptr + F + 1);
    G->h_normals.assign(d->face_normal_count, d->face_normal_count + F);
    // style groups: style 0 first (its slot always exists), then ascending
    std::vector<int32_t> gstyle{0};
    for (int32_t l = 0; l < NL; l++)
        if (d->lights[l].style != 0) gstyle.push_back(d->lights[l].style);
    std::sort(gstyle.begin() + 1, gstyle.end());
    gstyle.erase(std::unique(gstyle.begin() + 1, gstyle.end()), gstyle.end());
    const int32_t NG = (int32_t)gstyle.size();
    std::vector<int32_t> lindex, gptr(NG + 1, 0);
    std::vector<vh_light> sorted;
    for (int32_t g = 0; g < NG; g++) {
        for (int32_t l = 0; l < NL; l++)
            if (d->lights[l].style == gstyle[g]) {
                lindex.push_back(l);
                sorted.push_back(d->lights[l]);
            }
        gptr[g + 1] = (int32_t)lindex.size();
    }
    // the group offsets, then each group's style number
    std::vector<int32_t> gptr_styles(gptr);
    gptr_styles.insert(gptr_styles.end(), gstyle.begin(), gstyle.end());
    G->ngroups_style = NG;
    G->nslots = NG < 4 ? NG : 4;
    G->h_group_style = gstyle;
    G->h_face_styles.assign((size_t)F * 4, 255);
    for (int32_t f = 0; f < F; f++) G->h_face_styles[(size_t)f * 4] = 0;
    G->style_generation++;
    std::vector<int32_t> sface((size_t)S > 0 ? S : 1, 0), slot0((size_t)F * 4 + 4, -1);
    for (int32_t f = 0; f < F; f++) {
        for (int32_t x = d->face_sample_ptr[f]; x < d->face_sample_ptr[f + 1]; x++) sface[x] = f;
        slot0[(size_t)f * 4] = 0;
    }
    // chunks: 16 groups (64 samples) of one face each
    std::vector<int32_t> cf, cs;
    G->h_face_chunk_ptr.assign(F + 1, 0);
    for (int32_t f = 0; f < F; f++) {
        for (int32_t s = d->face_sample_ptr[f]; s < d->face_sample_ptr[f + 1]; s += 64) {
            cf.push_back(f);
            cs.push_back(s);
        }
        G->h_face_chunk_ptr[f + 1] = (int64_t)cf.size();
    }
    G->nchunks = (int64_t)cf.size();
    hipStream_t st = c->stream;
    int rc;
    if ((rc = upload(G->face_ptr, d->face_sample_ptr, (size_t)F + 1, "face samples", st)) ||
        (rc = upload(G->face_nc, d->face_normal_count, (size_t)F, "face normal counts", st)) ||
        (rc = upload(G->face_normals, d->face_normals, (size_t)F * 12, "face normals", st)) ||
        (rc = upload(G->pos, d->sample_pos, (size_t)S * 3, "sample positions", st)) ||
        (rc = upload(G->cluster, d->sample_cluster, (size_t)S, "sample clusters", st)) ||
        (rc = upload(G->lights, sorted.data(), sorted.size(), "lights", st)) ||
        (rc = upload(G->light_index, lindex.data(), lindex.size(), "light order", st)) ||
        (rc = upload(G->group_ptr, gptr_styles.data(), gptr_styles.size(), "style groups", st)) ||
        (rc = upload(G->sample_face, sface.data(), (size_t)S, "sample faces", st)) ||
        (rc = upload(G->slot_group, slot0.data(), slot0.size(), "style slots", st)) ||
        (rc = upload(G->face_styles, G->h_face_styles.data(), G->h_face_styles.size(), "face styles", st)) ||
        (NG > 1 && ((rc = G->first_touch.ensure((size_t)F * NG * 8 + 16, "first touches")) ||
                    (rc = G->stage_ldr.ensure((size_t)NG * S * 4 * 16 + 16, "LDR lighting per style")) ||
                    (rc = G->stage_hdr.ensure((size_t)NG * S * 4 * 16 + 16, "HDR lighting per style")))) ||
        (rc = G->dropped.ensure(16, "dropped styles")) ||
        (rc = upload(G->pvs, d->light_pvs, (size_t)NL * G->pvs_bytes, "light PVS", st)) ||
        (rc = upload(G->jitter, d->sun_jitter, (size_t)(ns - 1) * 3, "sun jitter", st)) ||
        (rc = upload(G->sky, d->sky_dirs, (size_t)G->sky_count * 3, "sky directions", st)) ||
        (rc = upload(G->ao_dirs, d->ao_dirs, (size_t)F * G->ao_count * 3, "AO directions", st)) ||
        (rc = upload(G->chunk_face, cf.data(), cf.size(), "chunks", st)) ||
        (rc = upload(G->chunk_first, cs.data(), cs.size(), "chunks", st)) ||
        (rc = G->out_ldr.ensure((size_t)G->nslots * S * 4 * 16 + 16, "LDR lighting")) ||
        (rc = G->out_hdr.ensure((size_t)G->nslots * S * 4 * 16 + 16, "HDR lighting")) ||
        (rc = G->out_ao.ensure((size_t)S * 4 + 16, "ambient occlusion")) ||
        (rc = G->ticket.ensure(16, "queue ticket")))
        return rc;
    VH_WAIT(c, st);
    return VH_OK;
}

int vh_gather_skybox(vh_ctx *c, const vh_skybox_desc *d)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!d || d->ncameras == 0) {
        c->sb_ncam = 0;
        return VH_OK;
    }
    const int32_t NC = d->ncameras, NN = d->nnodes, NP = d->nplanes, NLF = d->nleaves;
    if (NC < 0 || NC > VH_SKYBOX_MAX_CAMERAS) {
        set_error("a skybox holds 0 to " + std::to_string(VH_SKYBOX_MAX_CAMERAS) + " cameras");
        return VH_EINVAL;
    }
    if (NN < 0 || NP < 0 || NLF < 0 || !d->camera_origin || !d->camera_world_to_sky ||
        (NN > 0 && (!d->node_plane || !d->node_children)) ||
        (NP > 0 && (!d->plane_normal || !d->plane_dist || !d->plane_type)) || (NLF > 0 && !d->leaf_recurses)) {
        set_error("incomplete skybox descriptor");
        return VH_EINVAL;
    }
    std::vector<float4> cam(NC);
    for (int32_t k = 0; k < NC; k++) {
        const float *o = d->camera_origin + 3 * (size_t)k;
        const float w = d->camera_world_to_sky[k];
        if (!std::isfinite(o[0]) || !std::isfinite(o[1]) || !std::isfinite(o[2]) || !std::isfinite(w)) {
            set_error("sky camera " + std::to_string(k) + " is not finite");
            return VH_EINVAL;
        }
        cam[k] = make_float4(o[0], o[1], o[2], w);
    }
    // node records (children, plane type) and each node's plane; the walk
    // from node 0 must reach no node twice, so every descent ends
    std::vector<int4> nodes((size_t)NN > 0 ? NN : 1);
    std::vector<float4> planes((size_t)NN > 0 ? NN : 1);
    for (int32_t i = 0; i < NN; i++) {
        const int32_t p = d->node_plane[i];
        if (p < 0 || p >= NP) {
            set_error("skybox node " + std::to_string(i) + " names plane " + std::to_string(p) + " of " +
                      std::to_string(NP));
            return VH_EINVAL;
        }
        nodes[i] = make_int4(d->node_children[2 * (size_t)i], d->node_children[2 * (size_t)i + 1],
                             d->plane_type[p], 0);
        planes[i] = make_float4(d->plane_normal[3 * (size_t)p], d->plane_normal[3 * (size_t)p + 1],
                                d->plane_normal[3 * (size_t)p + 2], d->plane_dist[p]);
    }
    if (NN > 0) {
        std::vector<uint8_t> seen(NN, 0);
        std::vector<int32_t> todo{0};
        seen[0] = 1;
        while (!todo.empty()) {
            const int32_t i = todo.back();
            todo.pop_back();
            for (int k = 0; k < 2; k++) {
                const int32_t ch = k == 0 ? nodes[i].x : nodes[i].y;
                if (ch < 0 || ch >= NN) continue;
                if (seen[ch]) {
                    set_error("skybox node " + std::to_string(ch) + " is reached twice from node 0");
                    return VH_EINVAL;
                }
                seen[ch] = 1;
                todo.push_back(ch);
            }
        }
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    int rc;
    if ((rc = upload(c->sb_cam, cam.data(), cam.size(), "sky cameras", st)) ||
        (rc = upload(c->sb_nodes, nodes.data(), (size_t)NN, "skybox BSP nodes", st)) ||
        (rc = upload(c->sb_planes, planes.data(), (size_t)NN, "skybox BSP planes", st)) ||
        (rc = upload(c->sb_recurse, d->leaf_recurses, (size_t)NLF, "skybox leaves", st)))
        return rc;
    VH_WAIT(c, st);
    c->sb_ncam = NC;
    c->sb_nnodes = NN;
    c->sb_nleaves = NLF;
    return VH_OK;
}

int vh_gather_run(vh_ctx *c, int32_t schedule, int32_t waves_per_cu)
{
    return vh_gather_run_shaped(c, schedule, waves_per_cu, 1);
}

int vh_gather_run_shaped(vh_ctx *c, int32_t schedule, int32_t waves_per_cu, int32_t chunks_per_ticket)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->gather) {
        set_error("no gather loaded: call vh_gather_load first");
        return VH_ESTATE;
    }
    if (!c->have_scene) {
        set_error("no scene loaded: the gather traces against vh_scene_load's casters");
        return VH_ESTATE;
    }
    if (schedule < VH_SCHEDULE_QUEUE || schedule > VH_SCHEDULE_GRID) {
        set_error("bad schedule");
        return VH_EINVAL;
    }
    if (chunks_per_ticket < 1 || chunks_per_ticket > 64) {
        set_error("chunks_per_ticket runs from 1 to 64");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    GatherData *G = c->gather;
    hipStream_t st = c->stream;
    GatherArgs a = gather_args(c);
    const int32_t NG = G->ngroups_style;
    // the kernel writes every sample's light; AO only when it is traced
    VH_HIP(hipMemsetAsync(G->out_ao.ptr, 0, (size_t)G->nsamples * 4, st));
    VH_HIP(hipMemsetAsync(G->ticket.ptr, 0, 16, st));  // queue ticket, ray count
    if (NG > 1) {
        VH_HIP(hipMemsetAsync(G->first_touch.ptr, 0xFF, (size_t)G->nfaces * NG * 8, st));
        VH_HIP(hipMemsetAsync(G->dropped.ptr, 0, 8, st));
    }
    if (G->nchunks > 0) {
        KernelTimer timer(c, VH_KERNEL_FACELIGHT, st);
        if (schedule == VH_SCHEDULE_QUEUE) {
            int wpc = waves_per_cu > 0 ? waves_per_cu : 20;
            int64_t waves = (int64_t)c->cu_count * wpc;
            if (waves > G->nchunks) waves = G->nchunks;
            if (c->counting && c->counters.ptr)
                hipLaunchKernelGGL((k_facelight<true, true>), dim3((uint32_t)waves), dim3(kWave), 0, st, a,
                                   (int64_t)0, (unsigned long long *)G->ticket.ptr, chunks_per_ticket);
            else
                hipLaunchKernelGGL(k_facelight<true>, dim3((uint32_t)waves), dim3(kWave), 0, st, a,
                                   (int64_t)0, (unsigned long long *)G->ticket.ptr, chunks_per_ticket);
        } else if (schedule == VH_SCHEDULE_GRID) {
            hipLaunchKernelGGL(k_facelight<false>, dim3((uint32_t)G->nchunks), dim3(kWave), 0, st,
                               a, (int64_t)0, nullptr);
        } else {
            for (int32_t f = 0; f < G->nfaces; f++) {
                int64_t a0 = G->h_face_chunk_ptr[f], n = G->h_face_chunk_ptr[f + 1] - a0;
                if (n > 0)
                    hipLaunchKernelGGL(k_facelight<false>, dim3((uint32_t)n), dim3(kWave), 0, st, a,
                                       a0, nullptr);
            }
        }
        timer.done();
        VH_HIP(hipGetLastError());
    }
    unsigned long long dropped = 0;
    if (NG > 1) {
        if (G->nfaces > 0)
            hipLaunchKernelGGL(k_assign_slots, dim3((uint32_t)((G->nfaces + 255) / 256)), dim3(256), 0, st,
                               (const unsigned long long *)G->first_touch.ptr, G->nfaces, NG,
                               (const int32_t *)G->group_ptr.ptr + NG + 1, (int32_t *)G->slot_group.ptr,
                               (uint8_t *)G->face_styles.ptr, (unsigned long long *)G->dropped.ptr);
        const int64_t n = (int64_t)G->nslots * 4 * G->nsamples;
        if (n > 0)
            hipLaunchKernelGGL(k_fill_slots, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                               (const float4 *)G->stage_ldr.ptr, (const float4 *)G->stage_hdr.ptr,
                               (const int32_t *)G->slot_group.ptr, (const int32_t *)G->sample_face.ptr,
                               G->nsamples, G->nslots, (float4 *)G->out_ldr.ptr, (float4 *)G->out_hdr.ptr);
        VH_HIP(hipGetLastError());
    }
    // the bounded wait first: the copies below then find the work done
    VH_WAIT(c, st);
    if (NG > 1) {
        VH_HIP(hipMemcpy(G->h_face_styles.data(), G->face_styles.ptr, G->h_face_styles.size(),
                         hipMemcpyDeviceToHost));
        VH_HIP(hipMemcpy(&dropped, G->dropped.ptr, 8, hipMemcpyDeviceToHost));
    }
    unsigned long long rays = 0;
    VH_HIP(hipMemcpy(&rays, (unsigned long long *)G->ticket.ptr + 1, 8, hipMemcpyDeviceToHost));
    c->stats.rays_gather += (int64_t)rays;
    c->stats.calls++;
    G->dropped_styles = (int64_t)dropped;
    G->style_generation++;
    G->have_results = true;
    return VH_OK;
}

int vh_gather_supersample(vh_ctx *c, const vh_supersample_desc *d, int32_t *pass_counts)
{
    clear_error();
    if (!c || !d) {
        set_error("NULL context or descriptor");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->gather || !c->gather->have_results) {
        set_error("no gather results: call vh_gather_run first");
        return VH_ESTATE;
    }
    if (!c->have_scene) {
        set_error("no scene loaded: the resample traces against vh_scene_load's casters");
        return VH_ESTATE;
    }
    GatherData *G = c->gather;
    const int32_t F = G->nfaces, S = G->nsamples;
    if (d->which != 0 && d->which != 1) {
        set_error("which is 0 (LDR) or 1 (HDR)");
        return VH_EINVAL;
    }
    if (d->extra_passes < 0 || d->extra_passes > VH_SUPERSAMPLE_MAX_PASSES) {
        set_error("extra_passes runs from 0 to " + std::to_string(VH_SUPERSAMPLE_MAX_PASSES));
        return VH_EINVAL;
    }
    if (F > 0 && !d->faces) {
        set_error("no face frames");
        return VH_EINVAL;
    }
    if (pass_counts)
        for (int32_t k = 0; k < d->extra_passes; k++) pass_counts[k] = 0;
    // every sample on a luxel of its face's grid, one sample a luxel
    std::vector<int64_t> lux_ptr((size_t)F + 1, 0);
    for (int32_t f = 0; f < F; f++) {
        const int32_t w = d->faces[f].width, h = d->faces[f].height;
        const int32_t n = G->h_face_ptr[f + 1] - G->h_face_ptr[f];
        if (w < 0 || h < 0 || (n > 0 && (w == 0 || h == 0))) {
            set_error("face " + std::to_string(f) + " has samples but no luxel grid");
            return VH_EINVAL;
        }
        if (!d->sample_st && (int64_t)n > (int64_t)w * h) {
            set_error("face " + std::to_string(f) + " has more samples than luxels");
            return VH_EINVAL;
        }
        lux_ptr[f + 1] = lux_ptr[f] + (int64_t)w * h;
    }
    if (d->sample_st) {
        std::vector<uint8_t> used((size_t)lux_ptr[F], 0);
        for (int32_t f = 0; f < F; f++) {
            const int32_t w = d->faces[f].width, h = d->faces[f].height;
            for (int32_t s = G->h_face_ptr[f]; s < G->h_face_ptr[f + 1]; s++) {
                const int32_t x = d->sample_st[2 * (size_t)s], y = d->sample_st[2 * (size_t)s + 1];
                if (x < 0 || x >= w || y < 0 || y >= h || used[lux_ptr[f] + x + (int64_t)y * w]) {
                    set_error("sample " + std::to_string(s) + " is off its face's luxel grid or shares a luxel");
                    return VH_EINVAL;
                }
                used[lux_ptr[f] + x + (int64_t)y * w] = 1;
            }
        }
    }
    if (S == 0 || d->extra_passes == 0) return VH_OK;
    VH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    struct Bufs {
        DevBuf &frames, &lux_ptr, &st_in, &luxmap, &st, &inten, &processed, &flags, &counts, &off, &dptr, &aptr,
            &dcf, &dcs, &acf, &acs, &dpos, &dcl, &apos, &acl, &sel_sample, &sel_face, &dl, &dh, &al, &ah;
    } B{G->ss_frames, G->ss_lux_ptr, G->ss_st_in, G->ss_luxmap, G->ss_st, G->ss_inten, G->ss_processed,
        G->ss_flags, G->ss_counts, G->ss_off, G->ss_dptr, G->ss_aptr, G->ss_dcf, G->ss_dcs, G->ss_acf,
        G->ss_acs, G->ss_dpos, G->ss_dcl, G->ss_apos, G->ss_acl, G->ss_sel_sample, G->ss_sel_face, G->ss_dl,
        G->ss_dh, G->ss_al, G->ss_ah};
    int rc;
    if ((rc = upload(B.frames, d->faces, (size_t)F, "face frames", st)) ||
        (rc = upload(B.lux_ptr, lux_ptr.data(), lux_ptr.size(), "luxel offsets", st)) ||
        (d->sample_st && (rc = upload(B.st_in, d->sample_st, (size_t)S * 2, "sample luxels", st))) ||
        (rc = B.luxmap.ensure((size_t)lux_ptr[F] * 4 + 16, "luxel map")) ||
        (rc = B.st.ensure((size_t)S * 8 + 16, "sample luxels")) ||
        (rc = B.inten.ensure((size_t)S * 16 + 16, "intensities")) ||
        (rc = B.processed.ensure((size_t)S + 16, "processed flags")) ||
        (rc = B.flags.ensure((size_t)S + 16, "selection flags")) ||
        (rc = B.counts.ensure((size_t)F * 4 + 16, "selection counts")) ||
        (rc = G->ss_sface.ensure((size_t)S * 4 + 16, "sample faces")))
        return rc;
    VH_HIP(hipMemsetAsync(G->ticket.ptr, 0, 16, st));
    // one StyleSupersample per style slot (FaceLightJob.cs:688-700), each with
    // its own selection and passes, resampling only its own style's lights
    for (int32_t slot = 0; slot < G->nslots; slot++) {
    float4 *light = (float4 *)(d->which ? G->out_hdr.ptr : G->out_ldr.ptr) + (size_t)slot * 4 * S;
    const int32_t *face_group = (const int32_t *)G->slot_group.ptr + slot;
    VH_HIP(hipMemsetAsync(B.processed.ptr, 0, (size_t)S, st));
    hipLaunchKernelGGL(k_ss_init, dim3((uint32_t)F), dim3(kWave), 0, st, (const int32_t *)G->face_ptr.ptr,
                       (const int32_t *)G->face_nc.ptr, (const vh_final_face *)B.frames.ptr,
                       (const int64_t *)B.lux_ptr.ptr, (const int32_t *)(d->sample_st ? B.st_in.ptr : nullptr),
                       (int32_t *)B.luxmap.ptr, (int32_t *)B.st.ptr, (int32_t *)G->ss_sface.ptr, (const float4 *)light,
                       (float *)B.inten.ptr, S);
    VH_HIP(hipGetLastError());
    const GatherArgs base = gather_args(c);
    std::vector<int32_t> counts((size_t)F), off((size_t)F + 1), dptr((size_t)F + 1), aptr((size_t)F + 1);
    std::vector<int32_t> dcf, dcs, acf, acs;
    for (int32_t pass = 0; pass < d->extra_passes; pass++) {
        {
            VH_HIP(hipMemsetAsync(B.counts.ptr, 0, (size_t)F * 4, st));
            KernelTimer timer(c, VH_KERNEL_SUPERSAMPLE, st);
            hipLaunchKernelGGL(k_ss_select, dim3((uint32_t)((S + 255) / 256)), dim3(256), 0, st,
                               (const int32_t *)G->ss_sface.ptr, (const int32_t *)G->face_nc.ptr,
                               (const vh_final_face *)B.frames.ptr, (const int64_t *)B.lux_ptr.ptr,
                               (const int32_t *)B.luxmap.ptr, (const int32_t *)B.st.ptr, (const float *)B.inten.ptr,
                               S, (uint8_t *)B.processed.ptr, (uint8_t *)B.flags.ptr, (int32_t *)B.counts.ptr,
                               d->unset_is_garbage ? 1 : 0, face_group);
            timer.done();
        }
        VH_HIP(hipGetLastError());
        VH_HIP(hipMemcpyAsync(counts.data(), B.counts.ptr, (size_t)F * 4, hipMemcpyDeviceToHost, st));
        VH_WAIT(c, st);
        off[0] = 0;
        for (int32_t f = 0; f < F; f++) off[f + 1] = off[f] + counts[f];
        const int32_t total = off[F];
        if (total == 0) break;
        if ((int64_t)total * 16 > INT32_MAX) {
            set_error("too many samples selected in one pass for 32-bit point indices");
            return VH_EINVAL;
        }
        if (pass_counts) pass_counts[pass] += total;
        // chunks of 64 points within one face, as the gather's own
        dcf.clear(), dcs.clear(), acf.clear(), acs.clear();
        for (int32_t f = 0; f <= F; f++) {
            dptr[f] = 16 * off[f];
            aptr[f] = 4 * off[f];
        }
        for (int32_t f = 0; f < F; f++) {
            for (int32_t x = dptr[f]; x < dptr[f + 1]; x += 64) dcf.push_back(f), dcs.push_back(x);
            for (int32_t x = aptr[f]; x < aptr[f + 1]; x += 64) acf.push_back(f), acs.push_back(x);
        }
        const size_t nd = (size_t)16 * total, na = (size_t)4 * total;
        if ((rc = upload(B.off, off.data(), off.size(), "selection offsets", st)) ||
            (rc = upload(B.dptr, dptr.data(), dptr.size(), "direct point ranges", st)) ||
            (rc = upload(B.aptr, aptr.data(), aptr.size(), "ambient point ranges", st)) ||
            (rc = upload(B.dcf, dcf.data(), dcf.size(), "direct chunks", st)) ||
            (rc = upload(B.dcs, dcs.data(), dcs.size(), "direct chunks", st)) ||
            (rc = upload(B.acf, acf.data(), acf.size(), "ambient chunks", st)) ||
            (rc = upload(B.acs, acs.data(), acs.size(), "ambient chunks", st)) ||
            (rc = B.dpos.ensure(nd * 12 + 16, "direct points")) || (rc = B.dcl.ensure(nd * 4 + 16, "direct points")) ||
            (rc = B.apos.ensure(na * 12 + 16, "ambient points")) || (rc = B.acl.ensure(na * 4 + 16, "ambient points")) ||
            (rc = B.sel_sample.ensure((size_t)total * 4 + 16, "selection")) ||
            (rc = B.sel_face.ensure((size_t)total * 4 + 16, "selection")) ||
            (rc = B.dl.ensure(nd * 64 + 16, "direct point light")) || (rc = B.dh.ensure(nd * 64 + 16, "direct point light")) ||
            (rc = B.al.ensure(na * 64 + 16, "ambient point light")) || (rc = B.ah.ensure(na * 64 + 16, "ambient point light")))
            return rc;
        KernelTimer timer(c, VH_KERNEL_SUPERSAMPLE, st);
        hipLaunchKernelGGL(k_ss_points, dim3((uint32_t)F), dim3(kWave), 0, st, (const int32_t *)G->face_ptr.ptr,
                           (const vh_final_face *)B.frames.ptr, (const uint8_t *)B.flags.ptr,
                           (const int32_t *)B.off.ptr, (const float *)G->pos.ptr, (const int32_t *)G->cluster.ptr,
                           (float *)B.dpos.ptr, (int32_t *)B.dcl.ptr, (float *)B.apos.ptr, (int32_t *)B.acl.ptr,
                           (int32_t *)B.sel_sample.ptr, (int32_t *)B.sel_face.ptr);
        for (int kind = 0; kind < 2; kind++) {
            GatherArgs a = base;
            const bool direct = kind == 0;
            a.face_ptr = (const int32_t *)(direct ? B.dptr.ptr : B.aptr.ptr);
            a.pos = (const float *)(direct ? B.dpos.ptr : B.apos.ptr);
            a.cluster = (const int32_t *)(direct ? B.dcl.ptr : B.acl.ptr);
            a.nsamples = (int32_t)(direct ? nd : na);
            a.chunk_face = (const int32_t *)(direct ? B.dcf.ptr : B.acf.ptr);
            a.chunk_first = (const int32_t *)(direct ? B.dcs.ptr : B.acs.ptr);
            a.nchunks = (int64_t)(direct ? dcf.size() : acf.size());
            a.out_ldr = (float4 *)(direct ? B.dl.ptr : B.al.ptr);
            a.out_hdr = (float4 *)(direct ? B.dh.ptr : B.ah.ptr);
            a.out_ao = nullptr;
            a.ao_count = 0;
            a.light_filter = direct ? 2 : 1;
            a.no_skip = 1;
            a.face_group = face_group;
            a.first_touch = nullptr;
            const bool any = direct ? G->nlights > G->n_sky_ambient : G->n_sky_ambient > 0;
            if (a.nchunks > 0 && any)
                hipLaunchKernelGGL(k_facelight<false>, dim3((uint32_t)a.nchunks), dim3(kWave), 0, st, a,
                                   (int64_t)0, nullptr);
        }
        hipLaunchKernelGGL(k_ss_resolve, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, total,
                           (const int32_t *)B.sel_sample.ptr, (const int32_t *)B.sel_face.ptr,
                           (const int32_t *)G->face_nc.ptr, (const float4 *)(d->which ? B.dh.ptr : B.dl.ptr),
                           (const float4 *)(d->which ? B.ah.ptr : B.al.ptr),
                           G->nlights > G->n_sky_ambient ? 1 : 0, G->n_sky_ambient > 0 ? 1 : 0, light,
                           (float *)B.inten.ptr, S);
        timer.done();
        VH_HIP(hipGetLastError());
    }
    }
    unsigned long long rays = 0;
    VH_HIP(hipMemcpyAsync(&rays, (unsigned long long *)G->ticket.ptr + 1, 8, hipMemcpyDeviceToHost, st));
    VH_WAIT(c, st);
    c->stats.rays_gather += (int64_t)rays;
    return VH_OK;
}

int vh_gather_read(vh_ctx *c, float *ldr, float *hdr, float *ao)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->gather || !c->gather->have_results) {
        set_error("no gather results: call vh_gather_run first");
        return VH_ESTATE;
    }
    VH_HIP(hipSetDevice(c->device));
    GatherData *G = c->gather;
    size_t outb = (size_t)G->nsamples * 4 * 16;
    if (ldr) VH_HIP(hipMemcpy(ldr, G->out_ldr.ptr, outb, hipMemcpyDeviceToHost));
    if (hdr) VH_HIP(hipMemcpy(hdr, G->out_hdr.ptr, outb, hipMemcpyDeviceToHost));
    if (ao) VH_HIP(hipMemcpy(ao, G->out_ao.ptr, (size_t)G->nsamples * 4, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_gather_styles(vh_ctx *c, uint8_t *face_styles, int32_t *slots, int64_t *dropped)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->gather || !c->gather->have_results) {
        set_error("no gather results: call vh_gather_run first");
        return VH_ESTATE;
    }
    GatherData *G = c->gather;
    if (face_styles && !G->h_face_styles.empty())
        std::memcpy(face_styles, G->h_face_styles.data(), G->h_face_styles.size());
    if (slots) *slots = G->nslots;
    if (dropped) *dropped = G->dropped_styles;
    return VH_OK;
}

int vh_gather_read_slot(vh_ctx *c, int32_t slot, float *ldr, float *hdr)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->gather || !c->gather->have_results) {
        set_error("no gather results: call vh_gather_run first");
        return VH_ESTATE;
    }
    GatherData *G = c->gather;
    if (slot < 0 || slot >= 4) {
        set_error("a style slot runs from 0 to 3");
        return VH_EINVAL;
    }
    size_t outb = (size_t)G->nsamples * 4 * 16;
    if (slot >= G->nslots) {
        // no face has this slot: all zeros, as the reference's empty slots
        if (ldr) std::memset(ldr, 0, outb);
        if (hdr) std::memset(hdr, 0, outb);
        return VH_OK;
    }
    VH_HIP(hipSetDevice(c->device));
    const size_t off = (size_t)slot * G->nsamples * 4;
    if (ldr) VH_HIP(hipMemcpy(ldr, (float4 *)G->out_ldr.ptr + off, outb, hipMemcpyDeviceToHost));
    if (hdr) VH_HIP(hipMemcpy(hdr, (float4 *)G->out_hdr.ptr + off, outb, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_encode_rgbe(vh_ctx *c, const float *rgb, int64_t n, int32_t stride, uint8_t *out)
{
    clear_error();
    if (!c || (n > 0 && (!rgb || !out)) || n < 0 || stride < 3) {
        set_error("bad encode arguments");
        return VH_EINVAL;
    }
    if (n == 0) return VH_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    VH_HIP(hipSetDevice(c->device));
    DevBuf in, o;
    int rc;
    if ((rc = in.ensure((size_t)n * stride * 4, "colours")) || (rc = o.ensure((size_t)n * 4, "RGBE"))) {
        in.release();
        o.release();
        return rc;
    }
    hipStream_t st = c->stream;
    hipError_t e = hipMemcpyAsync(in.ptr, rgb, (size_t)n * stride * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_encode, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                           (const float *)in.ptr, n, stride, (uint32_t *)o.ptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, o.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    in.release();
    o.release();
    if (e != hipSuccess) {
        set_error(std::string("encode failed: ") + hipGetErrorString(e));
        return VH_EDEVICE;
    }
    return VH_OK;
}

int vh_lightmap_encode(vh_ctx *c, int32_t which, uint8_t *out, int64_t capacity)
{
    clear_error();
    if (!c || !out || (which != 0 && which != 1)) {
        set_error("bad lump encode arguments");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->gather || !c->gather->have_results) {
        set_error("no gather results: call vh_gather_run first");
        return VH_ESTATE;
    }
    GatherData *G = c->gather;
    std::vector<int64_t> base(G->nfaces + 1, 0);
    std::vector<int32_t> face_of(G->nsamples > 0 ? G->nsamples : 1);
    for (int32_t f = 0; f < G->nfaces; f++) {
        int32_t n = G->h_face_ptr[f + 1] - G->h_face_ptr[f];
        base[f + 1] = base[f] + 4 * (int64_t)G->h_normals[f] * n;
        for (int32_t s = G->h_face_ptr[f]; s < G->h_face_ptr[f + 1]; s++) face_of[s] = f;
    }
    int64_t bytes = base[G->nfaces];
    if (capacity < bytes) {
        set_error(std::to_string(bytes) + " lump bytes need as much room, and " +
                  std::to_string(capacity) + " were given");
        return VH_EINVAL;
    }
    if (bytes == 0) return VH_OK;
    VH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DevBuf fo, fb, o;
    int rc;
    if ((rc = upload(fo, face_of.data(), (size_t)G->nsamples, "sample faces", st)) ||
        (rc = upload(fb, base.data(), base.size(), "face offsets", st)) ||
        (rc = o.ensure((size_t)bytes, "lump"))) {
        fo.release();
        fb.release();
        o.release();
        return rc;
    }
    int64_t n = (int64_t)G->nsamples * 4;
    hipLaunchKernelGGL(k_encode_lump, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                       (const float4 *)(which ? G->out_hdr.ptr : G->out_ldr.ptr), G->nsamples,
                       (const int32_t *)fo.ptr, (const int32_t *)G->face_ptr.ptr,
                       (const int32_t *)G->face_nc.ptr, (const int64_t *)fb.ptr,
                       (uint32_t *)o.ptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, o.ptr, (size_t)bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    fo.release();
    fb.release();
    o.release();
    if (e != hipSuccess) {
        set_error(std::string("lump encode failed: ") + hipGetErrorString(e));
        return VH_EDEVICE;
    }
    return VH_OK;
}

}  // extern "C"
