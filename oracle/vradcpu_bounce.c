/*
 * vradcpu -- CPU oracle for the vrad lighting hot path: transfers and bounce.
 *
 * TEST INFRASTRUCTURE ONLY (see vradcpu.h).  Built with -ffp-contract=off.
 * Restates, in the reference's operand order (paths under the reference's
 * src/SourceSharp.MapTools/):
 *   Rad/Bounce/VisMatrix.cs      -- candidate enumeration (Enumerator.Run,
 *                                   TestPatchToFace, TestPatchToPatch), the
 *                                   transfer ray, MakeTransfer, MakeScales
 *   Rad/Bounce/FormFactors.cs    -- DiffToDiff, PolyToDiff (Correct side)
 *   Rad/Bounce/Radiosity.cs      -- MoveDirectLightToEmission, GatherLight
 *                                   (flat), CollectLight, the bounce loop
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "vradcpu.h"

#define PLANE_TEST_EPSILON 0.01   /* VisMatrix.cs PlaneTestEpsilon (a double) */
#define TRANSFER_EPSILON 0.0000001 /* VisMatrix.cs TransferEpsilon (a double) */
#define MAX_PATCHES (4 * 65536)    /* VisMatrix.cs MaxPatches */

static inline float dot3(const float *a, const float *b)
{
    return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

/* Vec3.Normalise (MapFormats/Geometry/Vec3.cs:178-190): exact sqrt of the
 * left-to-right sum of squares and three divides; zero stays zero. */
static inline float normalise(const float *v, float *out)
{
    float len = sqrtf((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]));
    if (len == 0.0f) {
        out[0] = out[1] = out[2] = 0.0f;
        return 0.0f;
    }
    out[0] = v[0] / len;
    out[1] = v[1] / len;
    out[2] = v[2] / len;
    return len;
}

/* DetMathF.Asin is correctly rounded; the double evaluation rounded once
 * to float agrees with it except in cases of double rounding. */
static inline float asin_f(float x) { return (float)asin((double)x); }

/* FormFactors.DiffToDiff (Rad/Bounce/FormFactors.cs, Correct normalise). */
float vc_diff_to_diff(const float *o1, const float *n1, const float *o2, const float *n2)
{
    float d[3] = {o1[0] - o2[0], o1[1] - o2[1], o1[2] - o2[2]};
    float delta[3];
    float length = normalise(d, delta);
    return -dot3(delta, n1) * dot3(delta, n2) / (length * length);
}

/* FormFactors.PolyToDiff, Correct side (a sine above 1 is clamped). */
float vc_poly_to_diff(const float *poly, int32_t npts, float area, const float *o, const float *n)
{
    float ff = 0.0f;
    for (int32_t p = 0; p < npts; p++) {
        int32_t q = p < npts - 1 ? p + 1 : 0;
        float a[3] = {poly[3 * p] - o[0], poly[3 * p + 1] - o[1], poly[3 * p + 2] - o[2]};
        float b[3] = {poly[3 * q] - o[0], poly[3 * q + 1] - o[1], poly[3 * q + 2] - o[2]};
        float v1[3], v2[3];
        normalise(a, v1);
        normalise(b, v2);
        float c[3] = {(v1[1] * v2[2]) - (v1[2] * v2[1]), (v1[2] * v2[0]) - (v1[0] * v2[2]),
                      (v1[0] * v2[1]) - (v1[1] * v2[0])};
        float gamma[3];
        float sin_alpha = normalise(c, gamma);
        if (sin_alpha < -1.0f || sin_alpha > 1.0f) sin_alpha = sin_alpha > 1.0f ? 1.0f : -1.0f;
        float as = asin_f(sin_alpha);
        gamma[0] = gamma[0] * as;
        gamma[1] = gamma[1] * as;
        gamma[2] = gamma[2] * as;
        ff += dot3(gamma, n);
    }
    ff *= 0.5f / area;
    return ff;
}

/* VisMatrix.MakeScales: total > pi -> 1/total, else (float)(1/pi). */
void vc_make_scales(float *w, int32_t n)
{
    if (n == 0) return;
    float total = 0;
    for (int32_t j = 0; j < n; j++) total += w[j];
    total = (double)total > M_PI ? 1.0f / total : (float)(1.0f / M_PI);
    for (int32_t j = 0; j < n; j++) w[j] = w[j] * total;
}

/* ------------------------------------------------------------------------ */
/* Candidate enumeration (VisMatrix.Enumerator)                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    const vc_patchset *ps;
    int32_t *out;
    int64_t count;
    int64_t cap;
    float r_origin[3], r_normal[3];
    float r_plane_dist;
} sink_t;

static void test_patch_to_patch(sink_t *s, int32_t src)
{
    const vc_patchset *ps = s->ps;
    const float *so = ps->origin + 3 * (size_t)src;
    if (ps->child1[src] != -1) {
        float tmp[3] = {s->r_origin[0] - so[0], s->r_origin[1] - so[1], s->r_origin[2] - so[2]};
        if ((double)dot3(tmp, tmp) * 0.0625 < (double)ps->area[src]) {
            test_patch_to_patch(s, ps->child1[src]);
            test_patch_to_patch(s, ps->child2[src]);
            return;
        }
    }
    if ((double)dot3(so, s->r_normal) > (double)s->r_plane_dist + PLANE_TEST_EPSILON) {
        if (s->out && s->count < s->cap) s->out[s->count] = src;
        s->count++;
    }
}

/* Enumerator.Run (VisMatrix.cs): the faces of the receiver's PVS row, its own
 * face skipped, each face's root chain behind the "emitter behind the face
 * plane" test, hierarchical descent. */
int64_t vc_candidates(const vc_patchset *ps, int32_t receiver, int32_t *out, int64_t cap)
{
    sink_t s;
    s.ps = ps;
    s.out = out;
    s.count = 0;
    s.cap = cap;
    memcpy(s.r_origin, ps->origin + 3 * (size_t)receiver, 12);
    memcpy(s.r_normal, ps->normal + 3 * (size_t)receiver, 12);
    s.r_plane_dist = ps->plane_dist[receiver];
    int32_t cluster = ps->cluster[receiver];
    int32_t myface = ps->face[receiver];
    for (int32_t k = ps->row_faces_ptr[cluster]; k < ps->row_faces_ptr[cluster + 1]; k++) {
        int32_t f = ps->row_faces[k];
        if (f == myface) continue;
        int32_t head = ps->face_parent[f];
        if (head == -1) continue;
        const float *fn = ps->normal + 3 * (size_t)head;
        if (!((double)dot3(s.r_origin, fn) > (double)ps->plane_dist[head] + PLANE_TEST_EPSILON))
            continue;
        for (int32_t p = head; p != -1; p = ps->next_parent[p]) test_patch_to_patch(&s, p);
    }
    return s.count;
}

/* VisMatrix.TransferRay, exact divide: p1 = receiver origin + normal,
 * p2 = source origin + normal; direction = delta / len, reach = len. */
static void transfer_ray(const vc_patchset *ps, int32_t r, int32_t s, float *ray)
{
    const float *ro = ps->origin + 3 * (size_t)r, *rn = ps->normal + 3 * (size_t)r;
    const float *so = ps->origin + 3 * (size_t)s, *sn = ps->normal + 3 * (size_t)s;
    float p1[3] = {ro[0] + rn[0], ro[1] + rn[1], ro[2] + rn[2]};
    float p2[3] = {so[0] + sn[0], so[1] + sn[1], so[2] + sn[2]};
    float d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    float len = sqrtf((d[0] * d[0]) + (d[1] * d[1]) + (d[2] * d[2]));
    float safe = len == 0.0f ? 1.1920929E-07f : len;
    ray[0] = p1[0];
    ray[1] = p1[1];
    ray[2] = p1[2];
    ray[3] = d[0] / safe;
    ray[4] = d[1] / safe;
    ray[5] = d[2] / safe;
    ray[6] = len;
}

/* VisMatrix.MakeTransfer (Correct side). */
int vc_make_transfer(const vc_patchset *ps, int32_t r, int32_t s, float *trans)
{
    *trans = 0.0f;
    if (ps->face_sky[ps->face[s]]) return 0;
    float area2 = ps->area[s];
    if (area2 <= 0) return 0;
    const float *ro = ps->origin + 3 * (size_t)r, *rn = ps->normal + 3 * (size_t)r;
    const float *so = ps->origin + 3 * (size_t)s, *sn = ps->normal + 3 * (size_t)s;
    float scale = vc_diff_to_diff(so, sn, ro, rn);
    if (scale <= 0) return 0;
    float delta[3] = {ro[0] - so[0], ro[1] - so[1], ro[2] - so[2]};
    float threshold = (float)((M_PI * 0.04) * (double)dot3(delta, delta));
    if (threshold < area2) {
        scale = vc_poly_to_diff(ps->winding + 12 * (size_t)s, 4, area2, ro, rn);
        if ((double)scale <= 0.0) return 0;
    }
    *trans = area2 * scale;
    return (double)*trans > TRANSFER_EPSILON;
}

typedef struct {
    const vc_patchset *ps;
    const vc_bvh *bvh;
    const int64_t *cand_ptr; /* per patch, candidate capacity prefix */
    int64_t *made;           /* per patch */
    int32_t *src;
    float *w;
    int64_t next;
    pthread_mutex_t mu;
} tjob;

static void *transfer_worker(void *arg)
{
    tjob *j = (tjob *)arg;
    const vc_patchset *ps = j->ps;
    int32_t *cand = NULL;
    float *rays = NULL;
    int64_t cap = 0;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t k = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (k >= ps->nreceivers) break;
        int32_t r = ps->receivers[k];
        int64_t base = j->cand_ptr[r];
        int64_t n = j->cand_ptr[r + 1] - base;
        if (n > cap) {
            cap = n;
            cand = (int32_t *)realloc(cand, (size_t)cap * 4);
            rays = (float *)realloc(rays, (size_t)cap * 28);
        }
        int64_t got = vc_candidates(ps, r, cand, n);
        if (got != n) {
            j->made[r] = -1; /* the count pass and the fill pass disagree */
            continue;
        }
        for (int64_t i = 0; i < n; i++) transfer_ray(ps, r, cand[i], rays + 7 * i);
        uint64_t *bits = (uint64_t *)calloc((size_t)((n + 63) / 64) + 1, 8);
        if (n > 0) vc_bvh_visibility(j->bvh, rays, 7, 1.0f, n, 0.0f, -1, 0, bits, 1);
        int64_t made = 0;
        for (int64_t i = 0; i < n; i++) {
            if ((bits[i >> 6] >> (i & 63)) & 1) continue;
            if (made >= MAX_PATCHES) continue;
            float t;
            if (vc_make_transfer(ps, r, cand[i], &t)) {
                j->src[base + made] = cand[i];
                j->w[base + made] = t;
                made++;
            }
        }
        free(bits);
        vc_make_scales(j->w + base, (int32_t)made);
        j->made[r] = made;
    }
    free(cand);
    free(rays);
    return NULL;
}

/* The whole transfer build (VisMatrix.BuildAsync's count, fill, trace,
 * MakeTransfers): cand_ptr is the per-patch prefix of candidate counts
 * (vc_candidates), the output is staged at those offsets with made[r] kept. */
void vc_transfers_build(const vc_patchset *ps, const vc_bvh *bvh, const int64_t *cand_ptr,
                        int64_t *made, int32_t *src, float *w, int32_t threads)
{
    tjob j;
    memset(&j, 0, sizeof j);
    j.ps = ps;
    j.bvh = bvh;
    j.cand_ptr = cand_ptr;
    j.made = made;
    j.src = src;
    j.w = w;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int i = 1; i < threads; i++) pthread_create(&th[i], NULL, transfer_worker, &j);
    transfer_worker(&j);
    for (int i = 1; i < threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&j.mu);
}

/* ------------------------------------------------------------------------ */
/* Radiosity                                                                 */
/* ------------------------------------------------------------------------ */

/* The bounce of Rad/Bounce/Radiosity.cs, flat patches, in T (float: the
 * reference's arithmetic, left to right; double: the FP64 oracle).
 *
 * mode 0 (progressive, the reference): emit starts as the direct light;
 *   each iteration add = sum shoot[src]*w, TotalLight += add, emit = add.
 * mode 1 (from a start x0): total starts at x0 and emit = e0 + x0; each
 *   iteration add = sum shoot[src]*w is the new total and emit = e0 + add.
 *   With x0 = 0 the two produce the same sequence of totals.
 * Stop: tol > 0 stops when sum|add - delta| / sum|total| <= tol (L1 over
 * leaf RGB, in double); tol <= 0 uses the reference's rule, every channel of
 * the summed leaf emission below 1.0 (Radiosity.cs:163).  Either way at most
 * max_iter iterations.  Parents' totals are pulled up from their children at
 * the end, as CollectLight leaves them. */
/* GatherBumpedScalar / GatherBumpedFour (Rad/Bounce/Radiosity.cs:406-512) for
 * receiver j: per transfer the direction to the source normalised (exact
 * length, true divides, the zero vector for a zero length), the receiver's
 * cosine divided out, and shoot * (w * scale) added along each of the four
 * normals (normals4: the plane normal, then the three bump normals) unless the
 * source is behind it (Accumulate, :523-531).  The geometry is float; the sums
 * are in the shooters' precision (dbl: double).  out: 4 x 3 sums. */
static void gather_bumped_row(int32_t j, const int64_t *row_ptr, const int32_t *src, const float *w,
                              const float *origin, const float *normals4, const void *shoot, int dbl,
                              void *out)
{
    const float *n = normals4 + (int64_t)j * 12;
    double sd[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    float sf[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int64_t k = row_ptr[j]; k < row_ptr[j + 1]; k++) {
        int32_t s = src[k];
        float dx = origin[3 * s] - origin[3 * j];
        float dy = origin[3 * s + 1] - origin[3 * j + 1];
        float dz = origin[3 * s + 2] - origin[3 * j + 2];
        float len = sqrtf((dx * dx + dy * dy) + dz * dz);
        if (len != 0.0f) {
            dx = dx / len;
            dy = dy / len;
            dz = dz / len;
        } else {
            dx = dy = dz = 0.0f;
        }
        float scale = 1.0f / ((dx * n[0] + dy * n[1]) + dz * n[2]);
        float ws = w[k] * scale;
        for (int q = 0; q < 4; q++) {
            float d = (dx * n[3 * q] + dy * n[3 * q + 1]) + dz * n[3 * q + 2];
            if (d <= 0.0f) continue;
            for (int c = 0; c < 3; c++) {
                if (dbl) {
                    double v = ((const double *)shoot)[3 * s + c] * (double)ws;
                    sd[q][c] = sd[q][c] + v * (double)d;
                } else {
                    float v = ((const float *)shoot)[3 * s + c] * ws;
                    sf[q][c] = sf[q][c] + v * d;
                }
            }
        }
    }
    if (dbl)
        memcpy(out, sd, sizeof sd);
    else
        memcpy(out, sf, sizeof sf);
}

#define BOUNCE_BODY(T)                                                                        \
    {                                                                                         \
        T *emit = (T *)calloc((size_t)P * 3, sizeof(T));                                     \
        T *shoot = (T *)calloc((size_t)P * 3, sizeof(T));                                    \
        T *add = (T *)calloc((size_t)P * 3, sizeof(T));                                      \
        T *tot = (T *)calloc((size_t)P * 3, sizeof(T));                                      \
        T *addb = (T *)calloc((size_t)P * 9, sizeof(T));                                     \
        T *totb = (T *)calloc((size_t)P * 9, sizeof(T));                                     \
        for (int32_t i = 0; i < P; i++)                                                       \
            for (int c = 0; c < 3; c++) {                                                     \
                T x0 = start ? (T)start[3 * i + c] : (T)0;                                    \
                tot[3 * i + c] = mode == 1 ? x0 : (T)0;                                       \
                emit[3 * i + c] = (T)emit0[3 * i + c] + (mode == 1 ? x0 : (T)0);              \
                shoot[3 * i + c] = emit[3 * i + c] * (T)refl[3 * i + c];                      \
            }                                                                                 \
        /* from a start: a parent's emission is its children's, pulled up */                  \
        for (int32_t i = mode == 1 ? P - 1 : -1; i >= 0; i--) {                               \
            if (sky[i]) {                                                                     \
                for (int c = 0; c < 3; c++) emit[3 * i + c] = shoot[3 * i + c] = (T)0;        \
            } else if (child1[i] != -1) {                                                     \
                int32_t a = child1[i], b = child2[i];                                         \
                T s1 = (T)area[a] / ((T)area[a] + (T)area[b]);                                \
                T s2 = (T)area[b] / ((T)area[a] + (T)area[b]);                                \
                for (int c = 0; c < 3; c++) {                                                 \
                    emit[3 * i + c] = (emit[3 * a + c] * s1) + (emit[3 * b + c] * s2);        \
                    shoot[3 * i + c] = emit[3 * i + c] * (T)refl[3 * i + c];                  \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
        int32_t it = 0;                                                                       \
        for (;;) {                                                                            \
            for (int32_t j = 0; j < P; j++) {                                                 \
                if (bumped && bumped[j]) {                                                    \
                    T sb[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};                \
                    gather_bumped_row(j, row_ptr, src, w, origin, normals4,                   \
                                      (const void *)shoot, sizeof(T) == 8, (void *)sb);       \
                    for (int c = 0; c < 3; c++) {                                             \
                        add[3 * j + c] = sb[0][c];                                            \
                        for (int n = 1; n < 4; n++) addb[9 * j + 3 * (n - 1) + c] = sb[n][c]; \
                    }                                                                         \
                    continue;                                                                 \
                }                                                                             \
                T sx = 0, sy = 0, sz = 0;                                                     \
                for (int64_t k = row_ptr[j]; k < row_ptr[j + 1]; k++) {                       \
                    int32_t s = src[k];                                                       \
                    T wk = (T)w[k];                                                           \
                    sx = sx + (shoot[3 * s] * wk);                                            \
                    sy = sy + (shoot[3 * s + 1] * wk);                                        \
                    sz = sz + (shoot[3 * s + 2] * wk);                                        \
                }                                                                             \
                add[3 * j] = sx;                                                              \
                add[3 * j + 1] = sy;                                                          \
                add[3 * j + 2] = sz;                                                          \
            }                                                                                 \
            double num = 0, den = 0, etot[3] = {0, 0, 0};                                     \
            for (int32_t i = P - 1; i >= 0; i--) {                                            \
                if (sky[i]) {                                                                 \
                    for (int c = 0; c < 3; c++) emit[3 * i + c] = shoot[3 * i + c] = (T)0;    \
                } else if (child1[i] == -1) {                                                 \
                    if (bumped && bumped[i])                                                  \
                        for (int q = 0; q < 9; q++)                                           \
                            totb[9 * i + q] = mode == 0 ? totb[9 * i + q] + addb[9 * i + q]   \
                                                        : addb[9 * i + q];                    \
                    for (int c = 0; c < 3; c++) {                                             \
                        T a = add[3 * i + c];                                                 \
                        if (mode == 0) {                                                      \
                            num += fabs((double)a);                                           \
                            tot[3 * i + c] = tot[3 * i + c] + a;                              \
                            emit[3 * i + c] = a;                                              \
                        } else {                                                              \
                            num += fabs((double)a - (double)tot[3 * i + c]);                  \
                            tot[3 * i + c] = a;                                               \
                            emit[3 * i + c] = (T)emit0[3 * i + c] + a;                        \
                        }                                                                     \
                        den += fabs((double)tot[3 * i + c]);                                  \
                        etot[c] += (double)emit[3 * i + c];                                   \
                        shoot[3 * i + c] = emit[3 * i + c] * (T)refl[3 * i + c];              \
                    }                                                                         \
                } else {                                                                      \
                    int32_t a = child1[i], b = child2[i];                                     \
                    T s1 = (T)area[a] / ((T)area[a] + (T)area[b]);                            \
                    T s2 = (T)area[b] / ((T)area[a] + (T)area[b]);                            \
                    for (int c = 0; c < 3; c++) {                                             \
                        emit[3 * i + c] = (emit[3 * a + c] * s1) + (emit[3 * b + c] * s2);    \
                        shoot[3 * i + c] = emit[3 * i + c] * (T)refl[3 * i + c];              \
                    }                                                                         \
                }                                                                             \
            }                                                                                 \
            double res = den > 0 ? num / den : 0.0;                                           \
            if (residuals && it < max_iter) residuals[it] = res;                              \
            if (energies && it < max_iter) {                                                  \
                energies[3 * it] = etot[0];                                                   \
                energies[3 * it + 1] = etot[1];                                               \
                energies[3 * it + 2] = etot[2];                                               \
            }                                                                                 \
            it++;                                                                             \
            int stop = it >= max_iter;                                                        \
            if (tol > 0) stop = stop || res <= tol;                                           \
            else stop = stop || (etot[0] < 1.0 && etot[1] < 1.0 && etot[2] < 1.0);            \
            if (stop) break;                                                                  \
        }                                                                                     \
        for (int32_t i = P - 1; i >= 0; i--) {                                                \
            if (child1[i] != -1 && !sky[i]) {                                                 \
                int32_t a = child1[i], b = child2[i];                                         \
                T s1 = (T)area[a] / ((T)area[a] + (T)area[b]);                                \
                T s2 = (T)area[b] / ((T)area[a] + (T)area[b]);                                \
                for (int c = 0; c < 3; c++)                                                   \
                    tot[3 * i + c] = (tot[3 * a + c] * s1) + (tot[3 * b + c] * s2);           \
            }                                                                                 \
        }                                                                                     \
        for (int32_t i = P - 1; bumped && i >= 0; i--) {                                     \
            if (child1[i] != -1 && !sky[i] && bumped[i]) {                                    \
                int32_t a = child1[i], b = child2[i];                                         \
                T s1 = (T)area[a] / ((T)area[a] + (T)area[b]);                                \
                T s2 = (T)area[b] / ((T)area[a] + (T)area[b]);                                \
                for (int q = 0; q < 9; q++)                                                   \
                    totb[9 * i + q] = (totb[9 * a + q] * s1) + (totb[9 * b + q] * s2);        \
            }                                                                                 \
        }                                                                                     \
        for (int64_t i = 0; i < (int64_t)P * 3; i++) total_out[i] = (double)tot[i];           \
        if (total_bump_out)                                                                   \
            for (int64_t i = 0; i < (int64_t)P * 9; i++) total_bump_out[i] = (double)totb[i];  \
        *iters = it;                                                                          \
        free(emit);                                                                           \
        free(shoot);                                                                          \
        free(add);                                                                            \
        free(tot);                                                                            \
        free(addb);                                                                           \
        free(totb);                                                                           \
    }

void vc_bounce(int32_t P, const int64_t *row_ptr, const int32_t *src, const float *w,
               const float *refl, const float *emit0, const uint8_t *sky, const int32_t *child1,
               const int32_t *child2, const float *area, const float *start, int32_t mode,
               int32_t max_iter, double tol, int32_t use_double, double *total_out,
               double *residuals, double *energies, int32_t *iters, const float *origin,
               const float *normals4, const uint8_t *bumped, double *total_bump_out)
{
    if (use_double)
        BOUNCE_BODY(double)
    else
        BOUNCE_BODY(float)
}

/* ---- GatherLight alone, row-parallel (the bounce's CPU baseline) ---- */

typedef struct {
    int32_t P;
    const int64_t *row_ptr;
    const int32_t *src;
    const float *w, *shoot;
    float *add;
    int32_t next;
    pthread_mutex_t mu;
} gl_job;

static void *gl_worker(void *arg)
{
    gl_job *j = (gl_job *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int32_t a = j->next;
        j->next += 1024;
        pthread_mutex_unlock(&j->mu);
        if (a >= j->P) break;
        int32_t b = a + 1024 < j->P ? a + 1024 : j->P;
        for (int32_t r = a; r < b; r++) {
            float sx = 0, sy = 0, sz = 0;
            for (int64_t k = j->row_ptr[r]; k < j->row_ptr[r + 1]; k++) {
                int32_t s = j->src[k];
                float wk = j->w[k];
                sx = sx + (j->shoot[3 * s] * wk);
                sy = sy + (j->shoot[3 * s + 1] * wk);
                sz = sz + (j->shoot[3 * s + 2] * wk);
            }
            j->add[3 * r] = sx;
            j->add[3 * r + 1] = sy;
            j->add[3 * r + 2] = sz;
        }
    }
    return NULL;
}

/* Radiosity.GatherLight (Rad/Bounce/Radiosity.cs:216-252) for flat rows, the
 * bounce loop's per-row sum above in FP32 and transfer order: add_r = sum over
 * the row of shoot[src] * w.  Rows are independent, so they are shared out
 * over `threads` in blocks of 1,024. */
void vc_gather_light(int32_t P, const int64_t *row_ptr, const int32_t *src, const float *w,
                     const float *shoot, float *add, int32_t threads)
{
    gl_job j;
    memset(&j, 0, sizeof j);
    j.P = P;
    j.row_ptr = row_ptr;
    j.src = src;
    j.w = w;
    j.shoot = shoot;
    j.add = add;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int i = 1; i < threads; i++) pthread_create(&th[i], NULL, gl_worker, &j);
    gl_worker(&j);
    for (int i = 1; i < threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&j.mu);
}
