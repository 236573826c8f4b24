/*
 * vradcpu_final -- the final light on the CPU: FinalLightFace.Run with
 * LuxelRadial (SourceSharp.MapTools/Rad/Final/FinalLightFace.cs:93-299,
 * Rad/Final/LuxelRadial.cs).  TEST INFRASTRUCTURE ONLY (see vradcpu.h).
 *
 * The same statement as oracle/final.py, in C for maps of realistic size:
 * every light style of a face in turn (the loop over its slots; neighbours
 * matched by style number; bounced light into style 0 only), no
 * displacements, macro textures the identity, the Correct
 * sides of SampleRadialEdgeOffByOne and PatchRadialNeighbourBumpFromSelf.
 * Single precision in the reference's order (built -ffp-contract=off); the
 * reference's double promotions in double.  Faces are independent (each
 * writes its own bytes), so they are shared out over threads.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "vradcpu.h"

#define ONE_OVER_SQRT3 0.57735025882720947f /* OO_SQRT_3, LuxelRadial.cs:51 */
#define WEIGHT_EPS 0.00001f                 /* WEIGHT_EPS, :48 */
#define RADIAL_DIST 1.42                    /* RADIALDIST (double), :45 */
#define EQUAL_EPSILON_D 0.001               /* LightConstants.EqualEpsilonDouble */

/* FaceLightInfo.WorldToLuxel (Rad/Light/FaceLightInfo.cs:264-270). */
static void w2l(const vc_final_face *f, const float *p, float *s, float *t)
{
    float px = p[0] - f->origin[0], py = p[1] - f->origin[1], pz = p[2] - f->origin[2];
    float a = px * f->w2l_s[0] + py * f->w2l_s[1] + pz * f->w2l_s[2];
    float b = px * f->w2l_t[0] + py * f->w2l_t[1] + pz * f->w2l_t[2];
    *s = a - (float)f->min_s;
    *t = b - (float)f->min_t;
}

/* FaceLightInfo.LuxelToWorld (:285-291): origin, plus s, then plus t. */
static void l2w(const vc_final_face *f, float s, float t, float *out)
{
    s += (float)f->min_s;
    t += (float)f->min_t;
    for (int k = 0; k < 3; k++) {
        float p = f->origin[k] + f->l2w_s[k] * s;
        out[k] = p + f->l2w_t[k] * t;
    }
}

typedef struct {
    const vc_final_face *f;
    int32_t w, h, nb;
    float *wt; /* w * h */
    float *lt; /* 4 * w * h * 3: [b][cell][k] */
} radial;

static void radial_reset(radial *r, const vc_final_face *f, int32_t nb)
{
    r->f = f;
    r->w = f->width;
    r->h = f->height;
    r->nb = nb;
    int64_t n = (int64_t)r->w * r->h;
    memset(r->wt, 0, (size_t)n * sizeof(float));
    memset(r->lt, 0, (size_t)n * 12 * sizeof(float));
}

static inline void add_weighted(radial *r, int b, int64_t i, const float *c, float wgt)
{
    float *L = r->lt + ((int64_t)b * r->w * r->h + i) * 3;
    for (int k = 0; k < 3; k++) L[k] = L[k] + wgt * c[k];
}

/* LuxelRadial.AddDirect (LuxelRadial.cs:140-208).  light[b] = 3 floats. */
static void add_direct(radial *r, const float *point, float min_s, float min_t, float max_s,
                       float max_t, const float *const *light, int has_bump, int nbr_bump)
{
    float cs, ct;
    w2l(r->f, point, &cs, &ct);
    int32_t smin = (int32_t)min_s, tmin = (int32_t)min_t;
    int32_t smax = (int32_t)(max_s + 0.9999f) + 1, tmax = (int32_t)(max_t + 0.9999f) + 1;
    if (smin < 0) smin = 0;
    if (tmin < 0) tmin = 0;
    if (smax > r->w) smax = r->w;
    if (tmax > r->h) tmax = r->h;
    for (int32_t s = smin; s < smax; s++) {
        for (int32_t t = tmin; t < tmax; t++) {
            double x;
            x = (double)(min_s - (float)s);
            float s0 = (float)(x > -1.0 ? x : -1.0);
            x = (double)(min_t - (float)t);
            float t0 = (float)(x > -1.0 ? x : -1.0);
            x = (double)(max_s - (float)s);
            float s1 = (float)(x < 1.0 ? x : 1.0);
            x = (double)(max_t - (float)t);
            float t1 = (float)(x < 1.0 ? x : 1.0);
            float area = (s1 - s0) * (t1 - t0);
            if (!((double)area > EQUAL_EPSILON_D)) continue;
            float ds = fabsf(cs - (float)s), dt = fabsf(ct - (float)t);
            float rr = ds > dt ? ds : dt;
            rr = (double)rr < 0.1 ? (float)((double)area / 0.1) : area / rr;
            int64_t i = s + (int64_t)t * r->w;
            if (has_bump) {
                if (nbr_bump) {
                    for (int b = 0; b < 4; b++) add_weighted(r, b, i, light[b], rr);
                } else {
                    add_weighted(r, 0, i, light[0], rr);
                    float bw = rr * ONE_OVER_SQRT3;
                    for (int b = 1; b < 4; b++) add_weighted(r, b, i, light[0], bw);
                }
            } else {
                add_weighted(r, 0, i, light[0], rr);
            }
            r->wt[i] = r->wt[i] + rr;
        }
    }
}

static inline void add_scaled(radial *r, int b, int64_t i, const float *c, float wgt)
{
    float *L = r->lt + ((int64_t)b * r->w * r->h + i) * 3;
    for (int k = 0; k < 3; k++) L[k] = L[k] + c[k] * wgt;
}

/* LuxelRadial.AddBounced (LuxelRadial.cs:230-293). */
static void add_bounced(radial *r, const float *point, float min_s, float min_t, float max_s,
                        float max_t, const float *const *light, int has_bump, int nbr_bump)
{
    float cs, ct;
    w2l(r->f, point, &cs, &ct);
    float dist_s = max_s - min_s, dist_t = max_t - min_t;
    dist_s = 1.0 > (double)dist_s ? 1.0f : dist_s;
    dist_t = 1.0 > (double)dist_t ? 1.0f : dist_t;
    int32_t smin = (int32_t)((double)cs - (double)dist_s * RADIAL_DIST);
    int32_t tmin = (int32_t)((double)ct - (double)dist_t * RADIAL_DIST);
    int32_t smax = (int32_t)((double)cs + (double)dist_s * RADIAL_DIST + 1.0);
    int32_t tmax = (int32_t)((double)ct + (double)dist_t * RADIAL_DIST + 1.0);
    if (smin < 0) smin = 0;
    if (tmin < 0) tmin = 0;
    if (smax > r->w) smax = r->w;
    if (tmax > r->h) tmax = r->h;
    for (int32_t s = smin; s < smax; s++) {
        for (int32_t t = tmin; t < tmax; t++) {
            float ds = (cs - (float)s) / dist_s, dt = (ct - (float)t) / dist_t;
            float rr = 2.0f - (ds * ds + dt * dt);
            if (!(rr > 0.0f)) continue;
            int64_t i = s + (int64_t)t * r->w;
            if (has_bump) {
                if (nbr_bump) {
                    for (int b = 0; b < 4; b++) add_scaled(r, b, i, light[b], rr);
                } else {
                    add_scaled(r, 0, i, light[0], rr);
                    float bw = rr * ONE_OVER_SQRT3;
                    for (int b = 1; b < 4; b++) add_scaled(r, b, i, light[0], bw);
                }
            } else {
                add_scaled(r, 0, i, light[0], rr);
            }
            r->wt[i] = r->wt[i] + rr;
        }
    }
}

/* LuxelRadial.Sample (LuxelRadial.cs:314-359), the Correct edge test.
 * out = nb x 3; returns baseSampleOk. */
static int sample(const radial *r, const float *point, int red_errors, float *out)
{
    float cs, ct;
    w2l(r->f, point, &cs, &ct);
    int32_t u = (int32_t)(cs + 0.5f), v = (int32_t)(ct + 0.5f);
    if (u < 0 || u >= r->w || v < 0 || v >= r->h) {
        for (int b = 0; b < r->nb; b++) {
            out[3 * b] = 2550.0f;
            out[3 * b + 1] = 0.0f;
            out[3 * b + 2] = 0.0f;
        }
        return 0;
    }
    int64_t i = u + (int64_t)v * r->w;
    int ok = 1;
    for (int b = 0; b < r->nb; b++) {
        float wgt = r->wt[i];
        if (wgt > WEIGHT_EPS) {
            float sc = (float)(1.0 / (double)wgt);
            const float *L = r->lt + ((int64_t)b * r->w * r->h + i) * 3;
            for (int k = 0; k < 3; k++) out[3 * b + k] = L[k] * sc;
        } else {
            out[3 * b] = red_errors ? 2550.0f : 0.0f;
            out[3 * b + 1] = 0.0f;
            out[3 * b + 2] = 0.0f;
            if (b == 0) ok = 0;
        }
    }
    return ok;
}

static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* The face's light styles: its slots up to the first empty one. */
static int face_style_count(const vc_final_desc *d, int32_t f)
{
    if (!d->face_styles) return 1;
    int k = 0;
    while (k < 4 && d->face_styles[4 * (int64_t)f + k] != 255) k++;
    return k;
}

/* LightmapOffsets.Compute (Rad/Light/LightmapOffsets.cs:60-123): per lit
 * face, a median word per style, then 4 bytes per luxel per normal per style. */
int64_t vc_final_layout(const vc_final_desc *d, int32_t *offsets)
{
    int64_t size = 0;
    for (int32_t f = 0; f < d->nfaces; f++) {
        const vc_final_face *x = &d->faces[f];
        const int ns = face_style_count(d, f);
        if (!x->lit || ns == 0) {
            if (offsets) offsets[f] = -1;
            continue;
        }
        size += 4 * ns;
        if (offsets) offsets[f] = (int32_t)size;
        int nb = d->face_normal_count[f] == 4 ? 4 : 1;
        size += 4 * (int64_t)nb * x->width * x->height * ns;
    }
    return size;
}

typedef struct {
    const vc_final_desc *d;
    const int64_t *lux_base, *light_ofs;
    const int32_t *pptr, *plist; /* each face's leaf patches, highest index first */
    int32_t max_cells;
    uint8_t *lump;
    int32_t next;
    pthread_mutex_t mu;
} final_job;

/* The slot of face f holding light style `style`, or -1 (the neighbour
 * match of BuildLuxelRadial, FinalLightFace.cs:340-360). */
static int slot_of_style(const vc_final_desc *d, int32_t f, int style)
{
    if (!d->face_styles) return style == 0 ? 0 : -1;
    for (int k = 0; k < 4; k++)
        if (d->face_styles[4 * (int64_t)f + k] == style) return k;
    return -1;
}

/* FinalLightFace.Run's body for style slot `slot` of face fn. */
static void light_face_style(final_job *j, int32_t fn, int slot, radial *rad, radial *prad, float *lb,
                             float *pv, float *vals)
{
    const vc_final_desc *d = j->d;
    const vc_final_face *face = &d->faces[fn];
    const int nb = d->face_normal_count[fn] == 4 ? 4 : 1, bumped = nb == 4;
    const int32_t S = d->nsamples;
    const int style = d->face_styles ? d->face_styles[4 * (int64_t)fn + slot] : 0;
    const float *lp[4];
    radial_reset(rad, face, nb);
    /* BuildLuxelRadial (FinalLightFace.cs:301-389) */
    const float *own = d->light + (int64_t)slot * 4 * S * 4;
    for (int32_t k = d->face_sample_ptr[fn]; k < d->face_sample_ptr[fn + 1]; k++) {
        for (int b = 0; b < 4; b++) lp[b] = own + ((int64_t)b * S + k) * 4;
        const float *bd = d->sample_bounds + (int64_t)k * 4;
        add_direct(rad, d->sample_pos + (int64_t)k * 3, bd[0], bd[1], bd[2], bd[3], lp, bumped, bumped);
    }
    for (int32_t q = d->nbr_ptr[fn]; q < d->nbr_ptr[fn + 1]; q++) {
        int32_t n = d->nbr[q];
        const vc_final_face *nf = &d->faces[n];
        const int nslot = slot_of_style(d, n, style);  /* matched by style number */
        if (nslot < 0 || !nf->lit || d->face_sample_ptr[n + 1] == d->face_sample_ptr[n]) continue;
        const float *nl = d->light + (int64_t)nslot * 4 * S * 4;
        int nbump = d->face_normal_count[n] == 4;
        for (int32_t k = d->face_sample_ptr[n]; k < d->face_sample_ptr[n + 1]; k++) {
            for (int b = 0; b < 4; b++) lp[b] = nl + ((int64_t)b * S + k) * 4;
            const float *bd = d->sample_bounds + (int64_t)k * 4;
            float tmp[3], mns, mnt, mxs, mxt;
            l2w(nf, bd[0], bd[1], tmp);
            w2l(face, tmp, &mns, &mnt);
            l2w(nf, bd[2], bd[3], tmp);
            w2l(face, tmp, &mxs, &mxt);
            add_direct(rad, d->sample_pos + (int64_t)k * 3, mns, mnt, mxs, mxt, lp, bumped, nbump);
        }
    }
    /* bounced light goes into style 0 only (FinalLightFace.cs:182) */
    const int bounce = d->patch_light != NULL && slot == 0;
    if (bounce) {
        /* BuildPatchRadial (FinalLightFace.cs:404-416, 489-536) */
        radial_reset(prad, face, nb);
        for (int32_t q = d->nbr_ptr[fn] - 1; q < d->nbr_ptr[fn + 1]; q++) {
            int32_t src = q < d->nbr_ptr[fn] ? fn : d->nbr[q];
            int nbump = src == fn ? bumped : d->face_normal_count[src] == 4;
            for (int32_t e = j->pptr[src]; e < j->pptr[src + 1]; e++) {
                int32_t p = j->plist[e];
                float mns = 1e30f, mnt = 1e30f, mxs = -1e30f, mxt = -1e30f;
                for (int32_t w = d->winding_ptr[p]; w < d->winding_ptr[p + 1]; w++) {
                    float s, t;
                    w2l(face, d->winding + (int64_t)w * 3, &s, &t);
                    mns = mns < s ? mns : s;
                    mxs = mxs > s ? mxs : s;
                    mnt = mnt < t ? mnt : t;
                    mxt = mxt > t ? mxt : t;
                }
                for (int b = 0; b < 4; b++) lp[b] = d->patch_light + ((int64_t)p * 4 + b) * 3;
                add_bounced(prad, d->patch_origin + (int64_t)p * 3, mns, mnt, mxs, mxt, lp, bumped, nbump);
            }
        }
    }
    const int64_t n = (int64_t)face->width * face->height;
    const int no_samples = d->face_sample_ptr[fn + 1] == d->face_sample_ptr[fn];
    const float ml = face->minlight;
    uint8_t *out = j->lump + j->light_ofs[fn] + (int64_t)slot * nb * n * 4;
    int64_t cnt = 0;
    for (int64_t jx = 0; jx < n; jx++) {
        float grid[3];
        const float *pt = grid;
        if (d->luxel_pos)
            pt = d->luxel_pos + (j->lux_base[fn] + jx) * 3;
        else /* BuildFaceLuxels (Rad/Light/FaceSampleBuilder.cs:242-259) */
            l2w(face, (float)(jx % face->width), (float)(jx / face->width), grid);
        int ok = sample(rad, pt, d->red_errors, lb);
        if (bounce) {
            sample(prad, pt, d->red_errors, pv);
            for (int b = 0; b < nb; b++)
                for (int k = 0; k < 3; k++) lb[3 * b + k] = lb[3 * b + k] + pv[3 * b + k];
        }
        if (no_samples) {
            for (int b = 0; b < nb; b++) {
                lb[3 * b] = 255.0f;
                lb[3 * b + 1] = 0.0f;
                lb[3 * b + 2] = 0.0f;
            }
            ok = 0;
        }
        for (int b = 0; b < nb; b++) {
            float c[3];
            for (int k = 0; k < 3; k++) c[k] = lb[3 * b + k] > ml ? lb[3 * b + k] : ml;
            if (b == 0 && ok) {
                vals[cnt] = c[0];
                vals[n + cnt] = c[1];
                vals[2 * n + cnt] = c[2];
                cnt++;
            }
            vc_encode_rgbe(c, 1, 3, out + ((int64_t)b * n + jx) * 4);
        }
    }
    float med[3] = {0.0f, 0.0f, 0.0f};
    if (cnt > 0) {
        /* FinalLightFace.Median (:466-472) */
        for (int k = 0; k < 3; k++) {
            qsort(vals + k * n, (size_t)cnt, sizeof(float), cmp_float);
            med[k] = vals[k * n + (cnt >> 1)];
        }
    }
    /* stored before lightofs, in reverse style order (:270-279) */
    vc_encode_rgbe(med, 1, 3, j->lump + j->light_ofs[fn] - 4 * (slot + 1));
}

static void light_face(final_job *j, int32_t fn, radial *rad, radial *prad, float *lb, float *pv,
                       float *vals)
{
    const int ns = face_style_count(j->d, fn);
    for (int k = 0; k < ns; k++) light_face_style(j, fn, k, rad, prad, lb, pv, vals);
}

static void *final_worker(void *arg)
{
    final_job *j = (final_job *)arg;
    int64_t cells = j->max_cells > 0 ? j->max_cells : 1;
    radial rad = {0}, prad = {0};
    rad.wt = malloc((size_t)cells * sizeof(float));
    rad.lt = malloc((size_t)cells * 12 * sizeof(float));
    prad.wt = malloc((size_t)cells * sizeof(float));
    prad.lt = malloc((size_t)cells * 12 * sizeof(float));
    float *vals = malloc((size_t)cells * 3 * sizeof(float));
    float lb[12], pv[12];
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int32_t f = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (f >= j->d->nfaces) break;
        if (!j->d->faces[f].lit || j->light_ofs[f] < 0) continue;
        light_face(j, f, &rad, &prad, lb, pv, vals);
    }
    free(rad.wt);
    free(rad.lt);
    free(prad.wt);
    free(prad.lt);
    free(vals);
    return NULL;
}

int vc_final_light(const vc_final_desc *d, uint8_t *lump, int64_t capacity, int32_t threads)
{
    const int32_t F = d->nfaces;
    int64_t *lux_base = malloc(sizeof(int64_t) * (size_t)(F + 1));
    int64_t *light_ofs = malloc(sizeof(int64_t) * (size_t)(F + 1));
    int32_t *pptr = calloc((size_t)F + 1, sizeof(int32_t));
    int32_t *plist = malloc(sizeof(int32_t) * (size_t)(d->npatch + 1));
    int32_t max_cells = 0;
    int64_t size = 0, lux = 0;
    for (int32_t f = 0; f < F; f++) {
        const vc_final_face *x = &d->faces[f];
        int32_t n = x->width * x->height;
        lux_base[f] = lux;
        lux += n;
        light_ofs[f] = -1;
        if (n > max_cells) max_cells = n;
        const int ns = face_style_count(d, f);
        if (!x->lit || ns == 0) continue;
        size += 4 * ns;
        light_ofs[f] = size;
        size += 4 * (int64_t)(d->face_normal_count[f] == 4 ? 4 : 1) * n * ns;
    }
    if (size > capacity) {
        free(lux_base);
        free(light_ofs);
        free(pptr);
        free(plist);
        return -1;
    }
    for (int32_t p = 0; p < d->npatch; p++)
        if (d->patch_child1[p] < 0) pptr[d->patch_face[p] + 1]++;
    for (int32_t f = 0; f < F; f++) pptr[f + 1] += pptr[f];
    int32_t *fill = malloc(sizeof(int32_t) * (size_t)(F + 1));
    memcpy(fill, pptr, sizeof(int32_t) * (size_t)(F + 1));
    for (int32_t p = d->npatch - 1; p >= 0; p--)
        if (d->patch_child1[p] < 0) plist[fill[d->patch_face[p]]++] = p;
    free(fill);
    final_job j;
    memset(&j, 0, sizeof j);
    j.d = d;
    j.lux_base = lux_base;
    j.light_ofs = light_ofs;
    j.pptr = pptr;
    j.plist = plist;
    j.max_cells = max_cells;
    j.lump = lump;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 1; i < threads; i++) pthread_create(&th[i], NULL, final_worker, &j);
    final_worker(&j);
    for (int i = 1; i < threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&j.mu);
    free(th);
    free(lux_base);
    free(light_ofs);
    free(pptr);
    free(plist);
    return 0;
}
