/*
 * vradcpu -- CPU oracle for the vrad lighting hot path: the direct-light
 * gather, the directional sampler and the RGBE encode.
 *
 * TEST INFRASTRUCTURE ONLY (see vradcpu.h).  Built with -ffp-contract=off.
 * Restates (paths under the reference's src/SourceSharp.MapTools/):
 *   Rad/Light/DirectLightGatherer.cs -- GatherStandardLane (:1169-1306),
 *       EmitSkyLight/ResolveSkyLight (:716-845), AmbientSky (:847-1012),
 *       Resolve's bump rule (:441-479), InverseQuadratic (:1579-1587)
 *   Rad/Light/FaceLightJob.cs  -- EmitGroup / ResolveGroup: 4-sample groups,
 *       the per-lane PVS mask, (dot * mask) * falloff, AddLight per normal
 *   Rad/Light/LightingValue.cs -- AddLight
 *   Rad/Light/DirectLightGatherer.cs -- TestLineDoesHitSky's 3D-skybox
 *       recursion (:510-714, Correct side), LaneRecurses (:1611-1621);
 *       Rad/Light/CompiledBspTree.cs -- LeafFromPoint
 *   Rad/Light/DirectionalSampler.cs -- HaltonSequence, DirectionalSampler
 *   Rad/Ambient/StockLightColor.cs:136-185 -- Encode, CalcExponent
 * plus the build's own ambient-occlusion term (no reference counterpart).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "vradcpu.h"

#define LANES 4
#define MAX_TRACE_LENGTH ((float)(1.732050807569 * (2 * 16384))) /* LightConstants.cs */
#define DIST_EPSILON 0.03125f
#define EQUAL_EPSILON 0.001f
#define CONSTANT_DOT (0.7f / 2.0f)

int vc_bvh_blocked(const vc_bvh *b, const float o[3], const float d[3], float reach, float tmin);
int vc_bvh_sky_occluded(const vc_bvh *b, const float o[3], const float d[3], float reach);

static inline float maxps(float a, float b) { return a > b ? a : b; }
static inline float minps(float a, float b) { return a < b ? a : b; }
static inline float dot3(const float *a, const float *b)
{
    return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

/* ---- DirectionalSampler (Rad/Light/DirectionalSampler.cs) ---- */

static float halton_next(int base, int *seed)
{
    (*seed)++;
    int tmpseed = *seed;
    float ret = 0.0f;
    float fbase = (float)base;
    float base_inv = (float)(1.0 / fbase);
    while (tmpseed != 0) {
        int dig = tmpseed % base;
        ret += (float)dig * base_inv;
        base_inv /= fbase;
        tmpseed /= base;
    }
    return ret;
}

/* The first n directions of DirectionalSampler.  DetMathF's correctly rounded
 * acos/sin/cos are evaluated in double and rounded once. */
void vc_directional_sampler(int32_t n, float *out)
{
    int s2 = 1, s3 = 1;
    for (int32_t i = 0; i < n; i++) {
        float z = halton_next(2, &s2);
        z = (float)((2 * z) - 1.0);
        float phi = (float)acos((double)z);
        float theta = (float)(2.0 * M_PI * halton_next(3, &s3));
        float sinp = (float)sin((double)phi);
        out[3 * i] = (float)cos((double)theta) * sinp;
        out[3 * i + 1] = (float)sin((double)theta) * sinp;
        out[3 * i + 2] = z;
    }
}

float vc_halton(int32_t base, int32_t index)
{
    int seed = index;
    return halton_next(base, &seed);
}

/* ---- StockLightColor.Encode (Rad/Ambient/StockLightColor.cs:136-185) ---- */

static inline int calc_exponent(float f)
{
    if (f == 0.0f) return 0;
    uint32_t bits;
    memcpy(&bits, &f, 4);
    int exp_component = (int)((bits & 0x7F800000u) >> 23);
    return exp_component - (7 + 127);
}

/* .NET's float -> int conversion saturates and maps NaN to 0. */
static inline int32_t to_int_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)x;
}

void vc_encode_rgbe(const float *rgb, int64_t n, int32_t stride, uint8_t *out)
{
    for (int64_t i = 0; i < n; i++) {
        const float *v = rgb + i * stride;
        float mx;
        if (v[0] > v[1])
            mx = v[0] > v[2] ? v[0] : v[2];
        else
            mx = v[1] > v[2] ? v[1] : v[2];
        int e = calc_exponent(mx);
        uint32_t fbits = (uint32_t)(127 - e) << 23;
        float scalar;
        memcpy(&scalar, &fbits, 4);
        out[4 * i] = (uint8_t)to_int_sat(v[0] * scalar);
        out[4 * i + 1] = (uint8_t)to_int_sat(v[1] * scalar);
        out[4 * i + 2] = (uint8_t)to_int_sat(v[2] * scalar);
        out[4 * i + 3] = (uint8_t)(int8_t)e;
    }
}

/* ---- the gather ---- */

/* Planted mistakes (vc_gather_set_mutation): each breaks one rule of the
 * gather, so a test can show that its closed form catches that mistake. */
enum {
    MUT_SURFACE_EPS_SIGN = 1, /* the surface light's ray end moved into the emitter */
    MUT_SPOT_NO_DOT2 = 2,     /* the spot falloff without its cone dot */
    MUT_FRINGE_STOPDOT = 3,   /* the fringe's numerator with stopdot for stopdot2 */
    MUT_BUMP_NOT_ZEROED = 4,  /* bump dots kept where the flat dot is zero */
    MUT_TRACE_ZERO_DOT = 5,   /* a ray even when the dot is zero */
    MUT_TRACE_PADDING = 6,    /* padding lanes traced too */
    MUT_NO_POW = 7,           /* the spot exponent ignored */
    MUT_HARD_END_START = 8,   /* the hard falloff cut at the start distance */
    MUT_SKY_AMBIENT_OPEN = 9, /* sky ambient ignores occlusion */
    MUT_NO_DIST_CLAMP = 10,   /* the distance not clamped at one unit */
    MUT_SUN_FORWARD = 11,     /* the sun's rays cast along its direction, not back */
    MUT_SUN_HALF_LENGTH = 12, /* the sun's rays half MaxTraceLength long */
    MUT_STYLE_MERGED = 13,    /* every style's light added to slot 0 */
    MUT_SKYBOX_LANE_ZERO = 14, /* every lane recurses as lane zero does (stock's quirk) */
    MUT_SKYBOX_UNSCALED = 15,  /* skybox rays start at origin + start, not scaled */
    MUT_SKYBOX_NO_FULLY = 16,  /* skybox rays cast for a group whose first rays are all blocked */
    MUT_SKYBOX_NOT_ADDED = 17, /* a skybox hit replaces the first ray's answer instead of adding */
};
static int g_mut = 0;

void vc_gather_set_mutation(int32_t which) { g_mut = which; }

typedef struct {
    const vc_gather_desc *d;
    const vc_bvh *bvh;
    float *out_ldr, *out_hdr, *out_ao;
    uint8_t *face_styles; /* NULL: one output slot */
    int slots;
    int64_t next;
    int64_t rays, dropped;
    pthread_mutex_t mu;
} gjob;

static inline int pvs_check(const vc_gather_desc *d, int32_t light, int32_t cluster)
{
    if (cluster < 0 || d->pvs_bytes == 0) return 1;
    if (cluster >= d->pvs_bytes * 8) return 1;
    return (d->light_pvs[(size_t)light * d->pvs_bytes + (cluster >> 3)] >> (cluster & 7)) & 1;
}

/* InverseQuadratic: ((d*d)*q + l*d) + c, then its reciprocal. */
static inline float inverse_quadratic(const vc_light *l, float dd)
{
    float f = dd * dd;
    f *= l->quadratic_attn;
    f += l->linear_attn * dd;
    f += l->constant_attn;
    return 1.0f / f;
}

/* GatherStandardLane (DirectLightGatherer.cs:1169-1306): returns 1 when the
 * lane needs a ray (pos -> src). */
static int standard_lane(const vc_light *l, const float *pos, const float *nrm, int normals,
                         int needed, float *dot_out /* [normals] */, float *falloff_out,
                         float *src_out)
{
    float src[3] = {l->origin[0], l->origin[1], l->origin[2]};
    float delta[3] = {src[0] - pos[0], src[1] - pos[1], src[2] - pos[2]};
    float dist2 = (delta[0] * delta[0]) + (delta[1] * delta[1]) + (delta[2] * delta[2]);
    float rpc = 1.0f / sqrtf(dist2);
    delta[0] = delta[0] * rpc;
    delta[1] = delta[1] * rpc;
    delta[2] = delta[2] * rpc;
    float dist = sqrtf(dist2);
    float dot = dot3(delta, nrm);
    dot = maxps(0.0f, dot);
    int hard = l->end_fade > l->start_fade;
    float end = g_mut == MUT_HARD_END_START ? l->start_fade : l->end_fade;
    if (hard && !(dist <= end)) return 0;
    if (g_mut != MUT_NO_DIST_CLAMP) dist = maxps(dist, 1.0f);
    float eval = minps(dist, l->cap_dist);
    float falloff;
    if (l->type == VC_EMIT_POINT) {
        falloff = inverse_quadratic(l, eval);
    } else if (l->type == VC_EMIT_SURFACE) {
        float dot2 = -dot3(delta, l->normal);
        dot2 = maxps(0.0f, dot2);
        if (dot == 0.0f) return 0;
        falloff = dot2 / dist2;
        float eps = g_mut == MUT_SURFACE_EPS_SIGN ? -DIST_EPSILON : DIST_EPSILON;
        src[0] = src[0] + (l->normal[0] * eps);
        src[1] = src[1] + (l->normal[1] * eps);
        src[2] = src[2] + (l->normal[2] * eps);
    } else if (l->type == VC_EMIT_SPOTLIGHT) {
        float dot2 = -dot3(delta, l->normal);
        if (!(dot2 > l->stopdot2)) return 0;
        falloff = inverse_quadratic(l, eval);
        if (g_mut != MUT_SPOT_NO_DOT2) falloff = falloff * dot2;
        int in_fringe = dot2 <= l->stopdot;
        float lowc = g_mut == MUT_FRINGE_STOPDOT ? l->stopdot : l->stopdot2;
        float mult = (dot2 - lowc) / (l->stopdot - l->stopdot2);
        mult = minps(mult, 1.0f);
        mult = maxps(mult, 0.0f);
        /* DetMathF.Pow, correctly rounded (vradcpu_math.c) */
        if (l->exponent != 0.0f && l->exponent != 1.0f && g_mut != MUT_NO_POW)
            mult = vc_pow_f32(mult, l->exponent);
        mult = in_fringe ? mult : 1.0f;
        falloff = mult * falloff;
    } else {
        return 0;
    }
    if (hard) {
        float t = (dist - l->start_fade) / (l->end_fade - l->start_fade);
        t = minps(t, 1.0f);
        t = maxps(t, 0.0f);
        t = 1.0f - t;
        float mult = (6.0f * t) - 15.0f;
        mult = (mult * t) + 10.0f;
        mult = (t * t) * mult;
        mult = t * mult;
        falloff = mult * falloff;
    }
    int trace = needed && (dot != 0.0f || g_mut == MUT_TRACE_ZERO_DOT) && falloff != 0.0f;
    dot_out[0] = dot;
    *falloff_out = falloff;
    for (int n = 1; n < normals; n++) dot_out[n] = maxps(0.0f, dot3(nrm + 3 * n, delta));
    memcpy(src_out, src, 12);
    return trace;
}

/* FaceLight.FindOrAllocateStyle (Rad/Light/FaceLight.cs:204-222): the first
 * slot holding the style or the first empty one; -1 when all four are taken. */
static int find_or_allocate_style(uint8_t *styles, int style)
{
    for (int k = 0; k < 4; k++) {
        if (styles[k] == style) return k;
        if (styles[k] == 255) {
            styles[k] = (uint8_t)style;
            return k;
        }
    }
    return -1;
}

/* ---- the 3D skybox: TestLineDoesHitSky's recursion (DirectLightGatherer.cs:510-714,
 * Correct compliance), SkyCameras (SkyCameras.cs), CompiledBspTree.LeafFromPoint ---- */

#define SKYBOX_MAX_CAMERAS 64

/* CompiledBspTree.LeafFromPoint: from node 0, an axial plane (type < 3)
 * compares the coordinate it names (0 x, 1 y, any other z), any other the
 * dot product; a point on the plane goes to the front child (children[0]); a
 * node index past the last node answers leaf 0.  The walk is cut after
 * nnodes steps (-1), which only a cyclic node table reaches. */
int32_t vc_bsp_leaf_from_point(const vc_skybox *s, const float p[3])
{
    int32_t node = 0;
    for (int32_t step = 0; node >= 0; step++) {
        if (node >= s->nnodes) return 0;
        if (step > s->nnodes) return -1;
        const int32_t pl = s->node_plane[node];
        const int32_t type = s->plane_type[pl];
        float dist;
        if (type < 3)
            dist = (type == 0 ? p[0] : (type == 1 ? p[1] : p[2])) - s->plane_dist[pl];
        else
            dist = dot3(s->plane_normal + 3 * (size_t)pl, p) - s->plane_dist[pl];
        node = dist < 0.0f ? s->node_children[2 * (size_t)node + 1] : s->node_children[2 * (size_t)node];
    }
    return -1 - node;
}

/* DirectLightGatherer.LaneRecurses (:1611-1621) with the area test folded
 * into the per-leaf flag. */
static int lane_recurses(const vc_skybox *s, const float p[3])
{
    int32_t leaf = vc_bsp_leaf_from_point(s, p);
    return leaf >= 0 && leaf < s->nleaves && s->leaf_recurses[leaf] != 0;
}

/* ResolveSkyTest's recursion (:661-706): occ holds the 4 lanes' first-ray
 * occlusions (padding copies), inner ncam x 4 skybox answers; a group whose
 * lanes are all blocked takes none. */
static void sky_resolve(float occ[LANES], int32_t ncam, const float *inner, int mask)
{
    int fully = 1;
    for (int lane = 0; lane < LANES; lane++) fully &= occ[lane] >= 1.0f;
    if (mask == 0 || fully) return;
    for (int32_t c = 0; c < ncam; c++)
        for (int lane = 0; lane < LANES; lane++) {
            float fraction = 1.0f - minps(maxps(inner[c * LANES + lane], 0.0f), 1.0f);
            if ((mask >> lane) & 1) {
                if (g_mut == MUT_SKYBOX_NOT_ADDED)
                    occ[lane] = 1.0f - fraction;
                else
                    occ[lane] = occ[lane] + 1.0f - fraction;
            }
        }
}

void vc_sky_resolve(int32_t count, const float *first_occ, int32_t ncam, const float *inner, int32_t mask,
                    float *fraction)
{
    float occ[LANES], in[SKYBOX_MAX_CAMERAS * LANES];
    if (count < 1) count = 1;
    if (count > LANES) count = LANES;
    if (ncam > SKYBOX_MAX_CAMERAS) ncam = SKYBOX_MAX_CAMERAS;
    for (int lane = 0; lane < LANES; lane++) occ[lane] = first_occ[lane < count ? lane : count - 1];
    for (int32_t c = 0; c < ncam; c++)
        for (int lane = 0; lane < LANES; lane++) in[c * LANES + lane] = inner[c * LANES + (lane < count ? lane : count - 1)];
    sky_resolve(occ, ncam, in, mask);
    for (int lane = 0; lane < LANES; lane++) fraction[lane] = 1.0f - minps(maxps(occ[lane], 0.0f), 1.0f);
}

/* TestLineDoesHitSky for one group (EmitSkyTest + ResolveSkyTest): every real
 * lane's first ray p -> stop (LightRayLog.Occlusion: a hit short of the end
 * that is not sky), then with a skybox the recursion: the lanes whose start
 * recurses, unless every lane was blocked, cast per camera a ray from
 * origin + start * world_to_sky along their first ray's direction normalised
 * (Vec3.Normalise: three divides by the length; zero for a zero length),
 * MaxTraceLength long (EmitRecursion, :587-607).  Only those lanes' skybox
 * rays are traced and counted: the others' answers are never read.  occ gets
 * the 4 lanes' occlusions, unclamped, padding copying the last real lane. */
static void sky_test(gjob *j, float p[LANES][3], float stop[LANES][3], int count, float occ[LANES],
                     int64_t *rays)
{
    for (int lane = 0; lane < count; lane++) {
        float dir[3] = {stop[lane][0] - p[lane][0], stop[lane][1] - p[lane][1], stop[lane][2] - p[lane][2]};
        occ[lane] = vc_bvh_sky_occluded(j->bvh, p[lane], dir, 1.0f) ? 1.0f : 0.0f;
        ++*rays;
    }
    for (int lane = count; lane < LANES; lane++) occ[lane] = occ[count - 1];
    const vc_skybox *s = j->d->skybox;
    if (!s || s->ncameras <= 0) return;
    const int32_t ncam = s->ncameras < SKYBOX_MAX_CAMERAS ? s->ncameras : SKYBOX_MAX_CAMERAS;
    int mask = 0;
    for (int lane = 0; lane < LANES; lane++)
        if (lane_recurses(s, g_mut == MUT_SKYBOX_LANE_ZERO ? p[0] : p[lane])) mask |= 1 << lane;
    int fully = 1;
    for (int lane = 0; lane < LANES; lane++) fully &= occ[lane] >= 1.0f;
    if (mask == 0 || (fully && g_mut != MUT_SKYBOX_NO_FULLY)) return;
    float u[LANES][3];
    for (int lane = 0; lane < count; lane++) {
        float v[3] = {stop[lane][0] - p[lane][0], stop[lane][1] - p[lane][1], stop[lane][2] - p[lane][2]};
        float len = sqrtf((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]));
        for (int k = 0; k < 3; k++) u[lane][k] = len == 0.0f ? 0.0f : v[k] / len;
    }
    float inner[SKYBOX_MAX_CAMERAS * LANES];
    for (int32_t c = 0; c < ncam; c++) {
        const float *o = s->camera_origin + 3 * (size_t)c;
        const float w = g_mut == MUT_SKYBOX_UNSCALED ? 1.0f : s->camera_world_to_sky[c];
        for (int lane = 0; lane < count; lane++) {
            inner[c * LANES + lane] = 0.0f;
            if (!((mask >> lane) & 1)) continue;
            float ss[3], dir[3];
            for (int k = 0; k < 3; k++) {
                ss[k] = o[k] + (p[lane][k] * w);
                dir[k] = (ss[k] + (u[lane][k] * MAX_TRACE_LENGTH)) - ss[k];
            }
            inner[c * LANES + lane] = vc_bvh_sky_occluded(j->bvh, ss, dir, 1.0f) ? 1.0f : 0.0f;
            ++*rays;
        }
        for (int lane = count; lane < LANES; lane++) inner[c * LANES + lane] = inner[c * LANES + count - 1];
    }
    sky_resolve(occ, ncam, inner, mask);
}

/* One SampleGroup: EmitGroup + ResolveGroup (FaceLightJob.cs:329-463) for
 * the four samples of face f from `first`, adding into the face's style
 * slots `styles` as lights first touch it. */
static void gather_group(gjob *j, int32_t f, int32_t first, uint8_t *styles, int64_t *rays,
                         int64_t *dropped)
{
    const vc_gather_desc *d = j->d;
    int32_t fend = d->face_sample_ptr[f + 1];
    int count = fend - first < LANES ? fend - first : LANES;
    int normals = d->face_normal_count[f];
    const float *fn = d->face_normals + (size_t)f * 12;
    int32_t want_style = d->face_style_filter ? d->face_style_filter[f] : -1;
    float p[LANES][3];
    int32_t cl[LANES];
    for (int i = 0; i < LANES; i++) {
        int32_t s = first + (i < count ? i : count - 1);
        const float *ps = d->sample_pos + 3 * (size_t)s;
        p[i][0] = ps[0] + fn[0];
        p[i][1] = ps[1] + fn[1];
        p[i][2] = ps[2] + fn[2];
        cl[i] = d->sample_cluster[s];
    }
    for (int32_t li = 0; li < d->nlights; li++) {
        const vc_light *l = &d->lights[li];
        if (d->light_filter == 1 && l->type != VC_EMIT_SKYAMBIENT) continue;
        if (d->light_filter == 2 && l->type == VC_EMIT_SKYAMBIENT) continue;
        if (want_style >= 0 && l->style != want_style) continue;
        int mask = 0;
        for (int s = 0; s < LANES; s++)
            if (s < count && pvs_check(d, li, cl[s])) mask |= 1 << s;
        if (mask == 0) continue;
        float dot[4][LANES];
        float falloff[LANES], sun[LANES];
        memset(dot, 0, sizeof dot);
        memset(falloff, 0, sizeof falloff);
        memset(sun, 0, sizeof sun);
        if (l->type == VC_EMIT_POINT || l->type == VC_EMIT_SURFACE ||
            l->type == VC_EMIT_SPOTLIGHT) {
            for (int lane = 0; lane < count; lane++) {
                float dl[4] = {0, 0, 0, 0}, fo = 0, src[3];
                int trace = standard_lane(l, p[lane], fn, normals, (mask >> lane) & 1, dl, &fo, src);
                for (int n = 0; n < normals; n++) dot[n][lane] = dl[n];
                falloff[lane] = fo;
                if (trace) {
                    /* LightRayLog.MakeRay: the whole segment, reach 1 */
                    float dir[3] = {src[0] - p[lane][0], src[1] - p[lane][1], src[2] - p[lane][2]};
                    float vis = vc_bvh_blocked(j->bvh, p[lane], dir, 1.0f, 0.0f) ? 0.0f : 1.0f;
                    ++*rays;
                    dot[0][lane] = vis * dot[0][lane];
                    if (g_mut == MUT_TRACE_PADDING && lane == count - 1) *rays += LANES - count;
                }
            }
            /* padding lanes copy the last real lane and are not traced */
            for (int lane = count; lane < LANES; lane++) {
                for (int n = 0; n < normals; n++) dot[n][lane] = dot[n][count - 1];
                falloff[lane] = falloff[count - 1];
            }
        } else if (l->type == VC_EMIT_SKYLIGHT) {
            float dl[LANES];
            int all_zero = 1;
            for (int lane = 0; lane < LANES; lane++) {
                float dd = -dot3(fn, l->normal);
                dd = maxps(dd, 0.0f);
                dl[lane] = dd;
                all_zero &= dd == 0.0f;
            }
            if (!all_zero) {
                int nsamples = d->sun_samples;
                float jitter_scale = (float)(1.732050807569 * (2 * 16384) * (double)d->sun_extent);
                float len = g_mut == MUT_SUN_FORWARD ? MAX_TRACE_LENGTH : -MAX_TRACE_LENGTH;
                if (g_mut == MUT_SUN_HALF_LENGTH) len = -MAX_TRACE_LENGTH * 0.5f;
                float total[LANES] = {0, 0, 0, 0};
                for (int s = 0; s < nsamples; s++) {
                    float delta[3] = {l->normal[0] * len, l->normal[1] * len, l->normal[2] * len};
                    if (s != 0) {
                        const float *jt = d->sun_jitter + 3 * (size_t)(s - 1);
                        delta[0] = delta[0] + (jt[0] * jitter_scale);
                        delta[1] = delta[1] + (jt[1] * jitter_scale);
                        delta[2] = delta[2] + (jt[2] * jitter_scale);
                    }
                    /* the sun traces every real lane (:780-790) */
                    float occ[LANES], stop[LANES][3];
                    for (int lane = 0; lane < LANES; lane++)
                        for (int k = 0; k < 3; k++) stop[lane][k] = delta[k] + p[lane][k];
                    sky_test(j, p, stop, count, occ, rays);
                    for (int lane = 0; lane < LANES; lane++) {
                        float o = minps(maxps(occ[lane], 0.0f), 1.0f);
                        total[lane] += 1.0f - o;
                    }
                }
                float inv = 1.0f / (float)nsamples;
                for (int lane = 0; lane < LANES; lane++) {
                    float see = total[lane] * inv;
                    dot[0][lane] = dl[lane] * see;
                    falloff[lane] = 1.0f;
                    sun[lane] = see * 10000.0f;
                    for (int n = 1; n < normals; n++) dot[n][lane] = (-dot3(fn + 3 * n, l->normal)) * see;
                }
            }
        } else if (l->type == VC_EMIT_SKYAMBIENT) {
            float sumdot[LANES] = {0, 0, 0, 0};
            float ambient[4][LANES], possible[4][LANES], dots[4][LANES];
            memset(ambient, 0, sizeof ambient);
            memset(possible, 0, sizeof possible);
            for (int32_t a = 0; a < d->sky_dirs_count; a++) {
                const float *an = d->sky_dirs + 3 * (size_t)a;
                int valid[LANES], any = 0, tr[LANES];
                for (int lane = 0; lane < LANES; lane++) {
                    float dd = -dot3(fn, an);
                    valid[lane] = dd > EQUAL_EPSILON;
                    any |= valid[lane];
                    dots[0][lane] = valid[lane] ? dd : 0.0f;
                }
                if (!any) continue;
                for (int lane = 0; lane < LANES; lane++) {
                    sumdot[lane] = dots[0][lane] + sumdot[lane];
                    possible[0][lane] = (valid[lane] ? 1.0f : 0.0f) + possible[0][lane];
                    tr[lane] = valid[lane];
                    for (int n = 1; n < normals; n++) {
                        float dd = -dot3(fn + 3 * n, an);
                        int v2 = dd > EQUAL_EPSILON;
                        dots[n][lane] = v2 ? dd : 0.0f;
                        possible[n][lane] = (valid[lane] && v2 ? 1.0f : 0.0f) + possible[n][lane];
                        tr[lane] |= v2;
                    }
                }
                float frac[LANES];
                const vc_skybox *sb = d->skybox;
                if (sb && sb->ncameras > 0) {
                    /* EmitAmbientRays / ResolveAmbientRays with traceAll
                     * (:1014-1058): every real lane through the sky test */
                    float occ[LANES], stop[LANES][3];
                    for (int lane = 0; lane < LANES; lane++)
                        for (int k = 0; k < 3; k++) stop[lane][k] = (an[k] * -MAX_TRACE_LENGTH) + p[lane][k];
                    sky_test(j, p, stop, count, occ, rays);
                    for (int lane = 0; lane < LANES; lane++) frac[lane] = 1.0f - minps(maxps(occ[lane], 0.0f), 1.0f);
                } else {
                    for (int lane = 0; lane < LANES; lane++) {
                        if (lane >= count) {
                            frac[lane] = frac[count - 1];
                            continue;
                        }
                        int need = tr[lane] && ((mask >> lane) & 1);
                        if (need) {
                            float stop[3] = {(an[0] * -MAX_TRACE_LENGTH) + p[lane][0],
                                             (an[1] * -MAX_TRACE_LENGTH) + p[lane][1],
                                             (an[2] * -MAX_TRACE_LENGTH) + p[lane][2]};
                            float dir[3] = {stop[0] - p[lane][0], stop[1] - p[lane][1], stop[2] - p[lane][2]};
                            float occ = vc_bvh_sky_occluded(j->bvh, p[lane], dir, 1.0f) ? 1.0f : 0.0f;
                            ++*rays;
                            if (g_mut == MUT_SKY_AMBIENT_OPEN) occ = 0.0f;
                            frac[lane] = 1.0f - occ;
                        } else {
                            frac[lane] = 1.0f;
                        }
                    }
                }
                for (int n = 0; n < normals; n++)
                    for (int lane = 0; lane < LANES; lane++)
                        ambient[n][lane] = ambient[n][lane] + (frac[lane] * dots[n][lane]);
            }
            for (int lane = 0; lane < LANES; lane++) {
                falloff[lane] = 1.0f;
                for (int n = 0; n < normals; n++) {
                    float factor = possible[n][lane] / possible[0][lane];
                    dot[n][lane] = ambient[n][lane] / (factor * sumdot[lane]);
                }
            }
        } else {
            continue;
        }
        /* Resolve (DirectLightGatherer.cs:441-479): a light behind the face
         * lights no bump direction */
        for (int lane = 0; lane < LANES; lane++) {
            float d0 = maxps(dot[0][lane], 0.0f);
            dot[0][lane] = d0;
            int nz = d0 > 0.0f || g_mut == MUT_BUMP_NOT_ZEROED;
            for (int n = 1; n < normals; n++) {
                float dn = maxps(dot[n][lane], 0.0f);
                dot[n][lane] = nz ? dn : 0.0f;
            }
        }
        /* ResolveGroup: (dot * mask) * falloff, skip when all zero */
        float fx[4][LANES];
        int skip = 1;
        for (int n = 0; n < normals; n++)
            for (int lane = 0; lane < LANES; lane++) {
                float masked = ((mask >> lane) & 1) ? dot[n][lane] : 0.0f;
                fx[n][lane] = masked * falloff[lane];
                skip &= fx[n][lane] == 0.0f;
            }
        if (skip && !d->no_skip) continue;
        /* the light's style slot on this face; one-slot output takes every
         * light the caller passed (a resample of one style) */
        int slot = 0;
        if (j->slots > 1 && g_mut != MUT_STYLE_MERGED) {
            slot = find_or_allocate_style(styles, l->style);
            if (slot < 0) {
                ++*dropped;
                continue;
            }
        }
        for (int n = 0; n < normals; n++)
            for (int i = 0; i < count; i++) {
                size_t o = (((size_t)slot * 4 + n) * d->nsamples + first + i) * 4;
                float amt = fx[n][i];
                for (int pass = 0; pass < 2; pass++) {
                    float *out = pass ? j->out_hdr : j->out_ldr;
                    const float *col = pass ? l->intensity_hdr : l->intensity;
                    if (!out) continue;
                    out[o] = out[o] + (amt * col[0]);
                    out[o + 1] = out[o + 1] + (amt * col[1]);
                    out[o + 2] = out[o + 2] + (amt * col[2]);
                    out[o + 3] += sun[i];
                }
            }
    }
    /* ambient occlusion (the build's own term): cosine-weighted directions in
     * the face's frame, any hit within ao_radius; ao = unblocked fraction. */
    if (j->out_ao && d->ao_count > 0) {
        for (int i = 0; i < count; i++) {
            int unblocked = 0;
            for (int32_t a = 0; a < d->ao_count; a++) {
                const float *dir = d->ao_dirs + ((size_t)f * d->ao_count + a) * 3;
                float seg[3] = {dir[0] * d->ao_radius, dir[1] * d->ao_radius, dir[2] * d->ao_radius};
                if (!vc_bvh_blocked(j->bvh, p[i], seg, 1.0f, 0.0f)) unblocked++;
                ++*rays;
            }
            j->out_ao[first + i] = (float)unblocked / (float)d->ao_count;
        }
    }
}

/* A face's groups run in order on one thread: style slots fill in the order
 * lights first touch the face, group by group (FaceLightJob.cs:435). */
static void *gather_worker(void *arg)
{
    gjob *j = (gjob *)arg;
    const vc_gather_desc *d = j->d;
    int64_t rays = 0, dropped = 0;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t f = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (f >= d->nfaces) break;
        uint8_t styles[4] = {0, 255, 255, 255}; /* style 0 always exists (FaceLightJob.cs:185) */
        for (int32_t s = d->face_sample_ptr[f]; s < d->face_sample_ptr[f + 1]; s += LANES)
            gather_group(j, (int32_t)f, s, styles, &rays, &dropped);
        if (j->face_styles) memcpy(j->face_styles + 4 * f, styles, 4);
    }
    pthread_mutex_lock(&j->mu);
    j->rays += rays;
    j->dropped += dropped;
    pthread_mutex_unlock(&j->mu);
    return NULL;
}

static int64_t run_gather(const vc_gather_desc *d, const vc_bvh *bvh, float *out_ldr, float *out_hdr,
                          float *out_ao, uint8_t *face_styles, int64_t *dropped, int32_t threads)
{
    gjob j;
    memset(&j, 0, sizeof j);
    j.d = d;
    j.bvh = bvh;
    j.out_ldr = out_ldr;
    j.out_hdr = out_hdr;
    j.out_ao = out_ao;
    j.face_styles = face_styles;
    j.slots = face_styles ? 4 : 1;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int i = 1; i < threads; i++) pthread_create(&th[i], NULL, gather_worker, &j);
    gather_worker(&j);
    for (int i = 1; i < threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&j.mu);
    if (dropped) *dropped = j.dropped;
    return j.rays;
}

/* The gather into one slot: every face's samples in groups of four.
 * out_ldr/out_hdr hold 4 floats (RGB, direct sun amount) per (normal,
 * sample), normal-major, zeroed by the caller; out_ao one float per sample
 * (may be NULL).  Every light the descriptor passes adds to that one slot, so
 * the caller passes one style's lights (or a style filter).  Returns the rays
 * traced. */
int64_t vc_gather(const vc_gather_desc *d, const vc_bvh *bvh, float *out_ldr, float *out_hdr,
                  float *out_ao, int32_t threads)
{
    return run_gather(d, bvh, out_ldr, out_hdr, out_ao, NULL, NULL, threads);
}

int64_t vc_gather_styles(const vc_gather_desc *d, const vc_bvh *bvh, float *out_ldr, float *out_hdr,
                         float *out_ao, uint8_t *face_styles, int64_t *dropped, int32_t threads)
{
    return run_gather(d, bvh, out_ldr, out_hdr, out_ao, face_styles, dropped, threads);
}
