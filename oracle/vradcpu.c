/*
 * vradcpu -- CPU oracle for the vrad lighting hot path: the ray part.
 *
 * TEST INFRASTRUCTURE ONLY (see vradcpu.h).  Built with -O2
 * -ffp-contract=off so every float operation rounds once, in the order the
 * reference writes it.  Reference citations are paths under the reference's
 * src/SourceSharp.MapTools/.
 */
#include "vradcpu.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Stock's FourZeros, which is 1e-10 and not zero: the near clip, the
 * parallel-ray reject and the edge threshold (Tracing/KdRayTracer.cs:1211-1251). */
#define VC_EPS 1.0e-10f
/* Correct-mode zero-direction substitute 2^-60 (Tracing/KdRayTracer.cs:193). */
#define VC_ZERO_SUB 8.6736174e-19f
#define VC_SKY 0x01000000

/* ------------------------------------------------------------------------ */
/* Triangle conversion                                                       */
/* ------------------------------------------------------------------------ */

/* KdTreeBuilder.GetEdgeEquation (Tracing/KdTreeBuilder.cs:1531-1555). */
static void edge_equation(const float p1[3], const float p2[3], int c1, int c2,
                          const float inside[3], float out[3])
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

/* KdTreeBuilder.ToIntersectionFormat, Correct side (exact Vec3.Normalise)
 * (Tracing/KdTreeBuilder.cs:1568-1612; MapFormats/Geometry/Vec3.cs:130-190). */
int vc_tri_convert(const float *v9, const int32_t *ids, const uint8_t *flags,
                   int32_t n, vc_tri *out)
{
    for (int32_t i = 0; i < n; i++) {
        const float *p1 = v9 + (int64_t)i * 9;
        const float *p2 = p1 + 3;
        const float *p3 = p1 + 6;
        float e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        float e2[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
        float c[3] = {(e1[1] * e2[2]) - (e1[2] * e2[1]),
                      (e1[2] * e2[0]) - (e1[0] * e2[2]),
                      (e1[0] * e2[1]) - (e1[1] * e2[0])};
        float len = sqrtf((c[0] * c[0]) + (c[1] * c[1]) + (c[2] * c[2]));
        float nv[3];
        if (len == 0.0f) {
            nv[0] = nv[1] = nv[2] = 0.0f;
        } else {
            nv[0] = c[0] / len;
            nv[1] = c[1] / len;
            nv[2] = c[2] / len;
        }
        int drop = 0;
        for (int k = 1; k < 3; k++)
            if (fabsf(nv[k]) > fabsf(nv[drop])) drop = k;
        int cs0 = (drop + 1) % 3;
        int cs1 = (drop + 2) % 3;
        float ea[3], eb[3];
        edge_equation(p1, p2, cs0, cs1, p3, ea);
        edge_equation(p2, p3, cs0, cs1, p1, eb);
        vc_tri *t = &out[i];
        t->nx = nv[0];
        t->ny = nv[1];
        t->nz = nv[2];
        t->d = (nv[0] * p1[0]) + (nv[1] * p1[1]) + (nv[2] * p1[2]);
        t->id = ids ? ids[i] : i;
        t->e[0] = ea[0];
        t->e[1] = ea[1];
        t->e[2] = ea[2];
        t->e[3] = eb[0];
        t->e[4] = eb[1];
        t->e[5] = eb[2];
        t->cs0 = (uint8_t)cs0;
        t->cs1 = (uint8_t)cs1;
        t->flags = flags ? flags[i] : 0;
        t->unused = 0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The exact per-triangle test                                               */
/* ------------------------------------------------------------------------ */

/* KdRayTracer.IntersectTriangle (Tracing/KdRayTracer.cs:1194-1267), one lane:
 * plane, parallel reject at +-1e-10, t > 1e-10, then the two pre-scaled edge
 * equations >= 1e-10 and their sum <= 1.  The "nearer than the best so far"
 * comparison is left to the caller.  Returns 1 and t on a hit. */
static inline int tri_hit(const vc_tri *tri, const float o[3], const float d[3], float *t_out)
{
    float ddotn = d[0] * tri->nx;
    ddotn = (d[1] * tri->ny) + ddotn;
    ddotn = (d[2] * tri->nz) + ddotn;
    if (!(ddotn > VC_EPS || ddotn < -VC_EPS)) return 0;
    float odotn = o[0] * tri->nx;
    odotn = (o[1] * tri->ny) + odotn;
    odotn = (o[2] * tri->nz) + odotn;
    float t = (tri->d - odotn) / ddotn;
    if (!(t > VC_EPS)) return 0;
    float h1 = o[tri->cs0] + (t * d[tri->cs0]);
    float h2 = o[tri->cs1] + (t * d[tri->cs1]);
    float b0 = tri->e[0] * h1;
    b0 = b0 + (tri->e[1] * h2);
    b0 = b0 + tri->e[2];
    if (!(b0 >= VC_EPS)) return 0;
    float b1 = tri->e[3] * h1;
    b1 = b1 + (tri->e[4] * h2);
    b1 = b1 + tri->e[5];
    if (!(b1 >= VC_EPS)) return 0;
    if (!((b1 + b0) <= 1.0f)) return 0;
    *t_out = t;
    return 1;
}

/* A ray with a non-finite origin or direction hits nothing: the KD walk's
 * scene-box clip fails for it (Tracing/KdRayTracer.cs:1023-1029) and the GPU
 * seam answers it a miss (MapTools.Gpu/Interop/Kernels.cs:111-115). */
static inline int ray_finite(const float o[3], const float d[3])
{
    for (int k = 0; k < 3; k++)
        if (!isfinite(o[k]) || !isfinite(d[k])) return 0;
    return 1;
}

static inline void load_ray(const float *rays, int32_t rw, float shared_reach, int64_t i,
                            float o[3], float d[3], float *reach)
{
    const float *r = rays + i * rw;
    o[0] = r[0];
    o[1] = r[1];
    o[2] = r[2];
    d[0] = r[3];
    d[1] = r[4];
    d[2] = r[5];
    *reach = rw == 7 ? r[6] : shared_reach;
}

/* Nearest hit, ties to the lower triangle index (the KD test replaces only on
 * a strictly nearer t, Tracing/KdRayTracer.cs:1227).  limit_mode 0: unclipped
 * (KD's initial best of 1e23, :981); 1: t <= reach; 2: t < reach. */
static inline int32_t closest_brute_one(const vc_tri *tris, int32_t ntri, const float o[3],
                                        const float d[3], float reach, float tmin,
                                        int32_t skip_id, int limit_mode, float *t_out)
{
    int32_t best = -1;
    float bt = 1.0e23f;
    if (!ray_finite(o, d)) {
        *t_out = bt;
        return -1;
    }
    for (int32_t k = 0; k < ntri; k++) {
        if (skip_id != -1 && tris[k].id == skip_id) continue;
        float t;
        if (!tri_hit(&tris[k], o, d, &t)) continue;
        if (!(t >= tmin)) continue;
        if (limit_mode == 1 && !(t <= reach)) continue;
        if (limit_mode == 2 && !(t < reach)) continue;
        if (t < bt) {
            bt = t;
            best = k;
        }
    }
    *t_out = bt;
    return best;
}

void vc_trace_closest_brute(const vc_tri *tris, int32_t ntri, const float *rays,
                            int32_t record_words, float shared_reach, int64_t nrays,
                            float tmin, int32_t skip_id, int32_t clip,
                            int32_t *tri_index, float *t_out)
{
    for (int64_t i = 0; i < nrays; i++) {
        float o[3], d[3], reach, t;
        load_ray(rays, record_words, shared_reach, i, o, d, &reach);
        int32_t k = closest_brute_one(tris, ntri, o, d, reach, tmin, skip_id, clip ? 1 : 0, &t);
        tri_index[i] = k;
        t_out[i] = k >= 0 ? t : 0.0f;
    }
}

/* IRayTracer.TraceVisibilityAsync's rule (Tracing/IRayTracer.cs:130-141,
 * KdRayTracer.cs:718-735): blocked when a hit lies strictly short of the end;
 * with sky passing, the first hit short of the end must not be a sky
 * triangle (KdRayTracer.cs:625-627). */
void vc_trace_visibility_brute(const vc_tri *tris, int32_t ntri, const float *rays,
                               int32_t record_words, float shared_reach, int64_t nrays,
                               float tmin, int32_t skip_id, int32_t sky_passes,
                               uint64_t *bits)
{
    memset(bits, 0, (size_t)((nrays + 63) / 64) * 8);
    for (int64_t i = 0; i < nrays; i++) {
        float o[3], d[3], reach, t;
        load_ray(rays, record_words, shared_reach, i, o, d, &reach);
        int32_t k = closest_brute_one(tris, ntri, o, d, reach, tmin, skip_id, 2, &t);
        int blocked = k >= 0;
        if (blocked && sky_passes && (tris[k].id & VC_SKY)) blocked = 0;
        if (blocked) bits[i >> 6] |= 1ULL << (i & 63);
    }
}

/* ------------------------------------------------------------------------ */
/* CPU BVH (binned SAH).  Conservative: every node box is widened so that no */
/* hit the exact test accepts can fall outside it.                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    float lo[3], hi[3];
    int32_t left;  /* internal: left child (right = left + 1); leaf: first tri */
    int32_t count; /* 0 for internal */
} bnode;

struct vc_bvh {
    bnode *nodes;
    int32_t nnodes;
    vc_tri *tris; /* reordered */
    int32_t *order; /* reordered position -> original index */
    int32_t ntri;
};

typedef struct {
    float lo[3], hi[3], c[3];
} tbox;

static void pad_box(float lo[3], float hi[3])
{
    for (int k = 0; k < 3; k++) {
        float m = fmaxf(fabsf(lo[k]), fabsf(hi[k]));
        float pad = m * 3.8146973e-6f + 1.0e-6f; /* 2^-18 relative */
        lo[k] -= pad;
        hi[k] += pad;
    }
}

static float area_of(const float lo[3], const float hi[3])
{
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
}

typedef struct {
    tbox *tb;
    int32_t *idx;
    bnode *nodes;
    int32_t nnodes;
    int32_t cap;
} bstate;

static int32_t push_node(bstate *s)
{
    if (s->nnodes == s->cap) {
        s->cap = s->cap ? s->cap * 2 : 1024;
        s->nodes = (bnode *)realloc(s->nodes, (size_t)s->cap * sizeof(bnode));
    }
    return s->nnodes++;
}

#define NBINS 16

static void build_rec(bstate *s, int32_t node, int32_t first, int32_t count, int depth)
{
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t i = first; i < first + count; i++) {
        tbox *b = &s->tb[s->idx[i]];
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], b->lo[k]);
            hi[k] = fmaxf(hi[k], b->hi[k]);
            clo[k] = fminf(clo[k], b->c[k]);
            chi[k] = fmaxf(chi[k], b->c[k]);
        }
    }
    bnode *nd = &s->nodes[node];
    memcpy(nd->lo, lo, sizeof lo);
    memcpy(nd->hi, hi, sizeof hi);
    pad_box(nd->lo, nd->hi);
    if (count <= 4 || depth > 60) {
        nd->left = first;
        nd->count = count;
        return;
    }
    float best_cost = INFINITY;
    int best_axis = -1, best_split = 0;
    for (int ax = 0; ax < 3; ax++) {
        float ext = chi[ax] - clo[ax];
        if (!(ext > 0)) continue;
        int cnt[NBINS] = {0};
        float blo[NBINS][3], bhi[NBINS][3];
        for (int b = 0; b < NBINS; b++)
            for (int k = 0; k < 3; k++) {
                blo[b][k] = INFINITY;
                bhi[b][k] = -INFINITY;
            }
        for (int32_t i = first; i < first + count; i++) {
            tbox *tb = &s->tb[s->idx[i]];
            int b = (int)((tb->c[ax] - clo[ax]) / ext * NBINS);
            if (b >= NBINS) b = NBINS - 1;
            if (b < 0) b = 0;
            cnt[b]++;
            for (int k = 0; k < 3; k++) {
                blo[b][k] = fminf(blo[b][k], tb->lo[k]);
                bhi[b][k] = fmaxf(bhi[b][k], tb->hi[k]);
            }
        }
        float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float rarea[NBINS];
        int rcnt[NBINS];
        int acc = 0;
        for (int b = NBINS - 1; b > 0; b--) {
            acc += cnt[b];
            for (int k = 0; k < 3; k++) {
                rlo[k] = fminf(rlo[k], blo[b][k]);
                rhi[k] = fmaxf(rhi[k], bhi[b][k]);
            }
            rarea[b] = area_of(rlo, rhi);
            rcnt[b] = acc;
        }
        float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        int lacc = 0;
        for (int b = 0; b < NBINS - 1; b++) {
            lacc += cnt[b];
            for (int k = 0; k < 3; k++) {
                llo[k] = fminf(llo[k], blo[b][k]);
                lhi[k] = fmaxf(lhi[k], bhi[b][k]);
            }
            if (lacc == 0 || rcnt[b + 1] == 0) continue;
            float cost = area_of(llo, lhi) * lacc + rarea[b + 1] * rcnt[b + 1];
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = ax;
                best_split = b;
            }
        }
    }
    int32_t mid;
    if (best_axis < 0) {
        mid = first + count / 2; /* all centroids equal: split by position */
    } else {
        float ext = chi[best_axis] - clo[best_axis];
        int32_t i = first, j = first + count - 1;
        while (i <= j) {
            tbox *tb = &s->tb[s->idx[i]];
            int b = (int)((tb->c[best_axis] - clo[best_axis]) / ext * NBINS);
            if (b >= NBINS) b = NBINS - 1;
            if (b < 0) b = 0;
            if (b <= best_split) {
                i++;
            } else {
                int32_t t = s->idx[i];
                s->idx[i] = s->idx[j];
                s->idx[j] = t;
                j--;
            }
        }
        mid = i;
        if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int32_t l = push_node(s);
    int32_t r = push_node(s);
    (void)r;
    s->nodes[node].left = l;
    s->nodes[node].count = 0;
    build_rec(s, l, first, mid - first, depth + 1);
    build_rec(s, l + 1, mid, first + count - mid, depth + 1);
}

vc_bvh *vc_bvh_build(const float *v9, const vc_tri *tris, int32_t ntri)
{
    vc_bvh *b = (vc_bvh *)calloc(1, sizeof(vc_bvh));
    b->ntri = ntri;
    bstate s = {0};
    s.tb = (tbox *)malloc((size_t)(ntri > 0 ? ntri : 1) * sizeof(tbox));
    s.idx = (int32_t *)malloc((size_t)(ntri > 0 ? ntri : 1) * sizeof(int32_t));
    for (int32_t i = 0; i < ntri; i++) {
        const float *v = v9 + (int64_t)i * 9;
        for (int k = 0; k < 3; k++) {
            float a = v[k], c = v[3 + k], e = v[6 + k];
            s.tb[i].lo[k] = fminf(a, fminf(c, e));
            s.tb[i].hi[k] = fmaxf(a, fmaxf(c, e));
            s.tb[i].c[k] = 0.5f * (s.tb[i].lo[k] + s.tb[i].hi[k]);
        }
        s.idx[i] = i;
    }
    int32_t root = push_node(&s);
    if (ntri > 0) {
        build_rec(&s, root, 0, ntri, 0);
    } else {
        bnode *nd = &s.nodes[root];
        for (int k = 0; k < 3; k++) {
            nd->lo[k] = 1.0f;
            nd->hi[k] = -1.0f;
        }
        nd->left = 0;
        nd->count = 0;
    }
    b->nodes = s.nodes;
    b->nnodes = s.nnodes;
    b->tris = (vc_tri *)malloc((size_t)(ntri > 0 ? ntri : 1) * sizeof(vc_tri));
    b->order = s.idx;
    for (int32_t i = 0; i < ntri; i++) b->tris[i] = tris[s.idx[i]];
    free(s.tb);
    return b;
}

void vc_bvh_free(vc_bvh *b)
{
    if (!b) return;
    free(b->nodes);
    free(b->tris);
    free(b->order);
    free(b);
}

int32_t vc_bvh_node_count(const vc_bvh *b) { return b->nnodes; }

static inline float inv_sat(float d)
{
    if (d == 0.0f) d = copysignf(VC_ZERO_SUB, d);
    return 1.0f / d;
}

static inline int box_hit(const bnode *n, const float o[3], const float inv[3], float tlo, float thi)
{
    float tn = tlo, tf = thi;
    for (int k = 0; k < 3; k++) {
        float a = (n->lo[k] - o[k]) * inv[k];
        float b = (n->hi[k] - o[k]) * inv[k];
        float mn = a < b ? a : b;
        float mx = a < b ? b : a;
        tn = mn > tn ? mn : tn;
        tf = mx < tf ? mx : tf;
    }
    return tn <= tf;
}

/* mode: 0 closest (t <= reach, ties to the lower ORIGINAL index); 1 any hit
 * strictly short of reach; 2 closest strictly short of reach (sky test). */
static int32_t bvh_trace_one(const vc_bvh *b, const float o[3], const float d[3], float reach,
                             float tmin, int32_t skip_id, int mode, float *t_out,
                             int32_t *hit_id)
{
    *t_out = 0.0f;
    if (!ray_finite(o, d) || b->ntri == 0) return -1;
    float inv[3] = {inv_sat(d[0]), inv_sat(d[1]), inv_sat(d[2])};
    int32_t stack[128];
    int sp = 0;
    int32_t node = 0;
    int32_t best = -1, best_orig = 0x7fffffff;
    float bt = reach;
    /* The box interval is widened relative to the exact test's range so a
     * hit on a box face is never culled. */
    float tlo = tmin < 0 ? tmin : tmin * 0.99999f;
    for (;;) {
        const bnode *n = &b->nodes[node];
        float thi = (mode == 1 ? reach : bt);
        thi = thi * 1.00001f + 1e-30f;
        if (box_hit(n, o, inv, tlo, thi)) {
            if (n->count == 0) {
                if (sp >= 126) return -2;
                stack[sp++] = n->left + 1;
                node = n->left;
                continue;
            }
            for (int32_t i = n->left; i < n->left + n->count; i++) {
                const vc_tri *tri = &b->tris[i];
                if (skip_id != -1 && tri->id == skip_id) continue;
                float t;
                if (!tri_hit(tri, o, d, &t)) continue;
                if (!(t >= tmin)) continue;
                if (mode == 1) {
                    if (t < reach) {
                        *t_out = t;
                        if (hit_id) *hit_id = tri->id;
                        return b->order[i];
                    }
                    continue;
                }
                int ok = mode == 0 ? (t <= reach) : (t < reach);
                if (!ok) continue;
                int32_t orig = b->order[i];
                if (best < 0 || t < bt || (t == bt && orig < best_orig)) {
                    bt = t;
                    best = i;
                    best_orig = orig;
                    if (hit_id) *hit_id = tri->id;
                }
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    if (best < 0) return -1;
    *t_out = bt;
    return best_orig;
}

/* Single-ray helpers for the gather oracle. */
int vc_bvh_blocked(const vc_bvh *b, const float o[3], const float d[3], float reach, float tmin)
{
    float t;
    return bvh_trace_one(b, o, d, reach, tmin, -1, 1, &t, NULL) >= 0;
}

/* LightRayLog.Occlusion (Rad/Light/LightRayLog.cs:655-659): 1 when the first
 * hit short of the end is not a sky triangle. */
int vc_bvh_sky_occluded(const vc_bvh *b, const float o[3], const float d[3], float reach)
{
    float t;
    int32_t id = 0;
    int32_t k = bvh_trace_one(b, o, d, reach, 0.0f, -1, 2, &t, &id);
    return k >= 0 && !(id & VC_SKY);
}

typedef struct {
    const vc_bvh *b;
    const float *rays;
    int32_t rw;
    float shared_reach;
    int64_t n;
    float tmin;
    int32_t skip_id;
    int32_t sky_passes;
    int mode; /* 0 vis, 1 closest, 2 point lights */
    uint64_t *bits;
    int32_t *idx;
    float *t;
    const float *points;
    const float *lights;
    int32_t nlights;
    int64_t next; /* in units of 64 rays (vis) or 256 (closest) */
    pthread_mutex_t mu;
} job;

static int64_t claim(job *j, int64_t step)
{
    pthread_mutex_lock(&j->mu);
    int64_t v = j->next;
    j->next += step;
    pthread_mutex_unlock(&j->mu);
    return v;
}

static void *worker(void *arg)
{
    job *j = (job *)arg;
    for (;;) {
        if (j->mode == 0) {
            int64_t w = claim(j, 1);
            if (w * 64 >= j->n) break;
            uint64_t word = 0;
            for (int64_t i = w * 64; i < j->n && i < w * 64 + 64; i++) {
                float o[3], d[3], reach, t;
                load_ray(j->rays, j->rw, j->shared_reach, i, o, d, &reach);
                int blocked;
                if (j->sky_passes) {
                    int32_t id = 0;
                    int32_t k = bvh_trace_one(j->b, o, d, reach, j->tmin, j->skip_id, 2, &t, &id);
                    blocked = k >= 0 && !(id & VC_SKY);
                } else {
                    blocked = bvh_trace_one(j->b, o, d, reach, j->tmin, j->skip_id, 1, &t, NULL) >= 0;
                }
                if (blocked) word |= 1ULL << (i & 63);
            }
            j->bits[w] = word;
        } else if (j->mode == 1) {
            int64_t s = claim(j, 256);
            if (s >= j->n) break;
            for (int64_t i = s; i < j->n && i < s + 256; i++) {
                float o[3], d[3], reach, t;
                load_ray(j->rays, j->rw, j->shared_reach, i, o, d, &reach);
                int32_t k = bvh_trace_one(j->b, o, d, reach, j->tmin, j->skip_id, 0, &t, NULL);
                j->idx[i] = k;
                j->t[i] = k >= 0 ? t : 0.0f;
            }
        } else {
            /* one point, every light: 64 lights make one word when nlights == 64 */
            int64_t p = claim(j, 1);
            if (p >= j->n) break;
            const float *pt = j->points + p * 3;
            for (int32_t l = 0; l < j->nlights; l++) {
                const float *lt = j->lights + (int64_t)l * 3;
                float o[3] = {pt[0], pt[1], pt[2]};
                float d[3] = {lt[0] - pt[0], lt[1] - pt[1], lt[2] - pt[2]};
                float t;
                int blocked = bvh_trace_one(j->b, o, d, 1.0f, j->tmin, -1, 1, &t, NULL) >= 0;
                int64_t r = p * j->nlights + l;
                if (blocked)
                    __atomic_fetch_or(&j->bits[r >> 6], 1ULL << (r & 63), __ATOMIC_RELAXED);
            }
        }
    }
    return NULL;
}

static void run_job(job *j, int32_t threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_mutex_init(&j->mu, NULL);
    pthread_t th[256];
    for (int i = 1; i < threads; i++) pthread_create(&th[i], NULL, worker, j);
    worker(j);
    for (int i = 1; i < threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&j->mu);
}

void vc_bvh_visibility(const vc_bvh *b, const float *rays, int32_t record_words,
                       float shared_reach, int64_t nrays, float tmin, int32_t skip_id,
                       int32_t sky_passes, uint64_t *bits, int32_t threads)
{
    job j;
    memset(&j, 0, sizeof j);
    j.b = b;
    j.rays = rays;
    j.rw = record_words;
    j.shared_reach = shared_reach;
    j.n = nrays;
    j.tmin = tmin;
    j.skip_id = skip_id;
    j.sky_passes = sky_passes;
    j.mode = 0;
    j.bits = bits;
    run_job(&j, threads);
}

void vc_bvh_closest(const vc_bvh *b, const float *rays, int32_t record_words,
                    float shared_reach, int64_t nrays, float tmin, int32_t skip_id,
                    int32_t *tri_index, float *t_out, int32_t threads)
{
    job j;
    memset(&j, 0, sizeof j);
    j.b = b;
    j.rays = rays;
    j.rw = record_words;
    j.shared_reach = shared_reach;
    j.n = nrays;
    j.tmin = tmin;
    j.skip_id = skip_id;
    j.mode = 1;
    j.idx = tri_index;
    j.t = t_out;
    run_job(&j, threads);
}

void vc_bvh_point_lights(const vc_bvh *b, const float *points, int64_t npoints,
                         const float *lights, int32_t nlights, float tmin,
                         uint64_t *bits, int32_t threads)
{
    job j;
    memset(&j, 0, sizeof j);
    memset(bits, 0, (size_t)((npoints * nlights + 63) / 64) * 8);
    j.b = b;
    j.n = npoints;
    j.tmin = tmin;
    j.mode = 2;
    j.bits = bits;
    j.points = points;
    j.lights = lights;
    j.nlights = nlights;
    run_job(&j, threads);
}
