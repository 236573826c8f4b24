/*
 * vradcpu -- CPU oracle for the vrad lighting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libvradhip.so and the
 * hl2sdk_tools_sharp_amd package) links, loads or calls this code.  Only the
 * tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, and
 * only as the checker or the timed CPU restatement, never as the thing
 * measured on the GPU.
 *
 * Every function restates the reference's method (C#, muflub/hl2sdk_tools_sharp)
 * in plain C with IEEE single precision, no FMA contraction (built with
 * -ffp-contract=off) and the reference's operand order.  The reference file
 * and line each one follows is cited beside it in vradcpu.c.
 *
 * Parity pins (what checks each function against something other than itself):
 *   vc_tri_convert          -- Correct-mode SHA-256 digests of kd-scene/kd-boxes
 *                              (triangle section); stock planes in *.tree.bin
 *                              agree to the estimate's last bits.
 *   vc_trace_closest_brute  -- stock's hit triangle per ray in *.answers.bin
 *                              (0 differing) and the Correct-mode digests
 *                              (hit distances bit for bit).
 *   vc_trace_visibility_brute -- stock's "hit short of the end" bits derived
 *                              from *.answers.bin (0 differing); swapped
 *                              endpoints give the same bit; a closed box room.
 *   vc_bvh_*                -- equal to the brute force on every ray of the
 *                              fixtures and of seeded random scenes.
 */
#ifndef VRADCPU_H
#define VRADCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* KdRayTracer's intersection-format triangle, 48 bytes
 * (Tracing/KdTree.cs:121-167). */
typedef struct vc_tri {
    float nx, ny, nz, d;
    int32_t id;
    float e[6];
    uint8_t cs0, cs1, flags, unused;
} vc_tri;

/* Ray record: origin xyz, direction xyz, reach (Tracing/IRayTracer.cs:32-39).
 * record_words is 7, or 6 with shared_reach giving every ray's reach. */

int vc_tri_convert(const float *v9, const int32_t *ids, const uint8_t *flags,
                   int32_t n, vc_tri *out);

/* Closest hit by brute force over every triangle in index order.
 * clip != 0: only t <= reach counts (the GPU seam's closest-hit rule).
 * clip == 0: any t counts (the unclipped KD rule, used for fixtures).
 * tri_index receives the triangle INDEX (-1 = miss), t the hit distance. */
void vc_trace_closest_brute(const vc_tri *tris, int32_t ntri, const float *rays,
                            int32_t record_words, float shared_reach, int64_t nrays,
                            float tmin, int32_t skip_id, int32_t clip,
                            int32_t *tri_index, float *t_out);

/* Visibility bits (LSB-first uint64 words), brute force. */
void vc_trace_visibility_brute(const vc_tri *tris, int32_t ntri, const float *rays,
                               int32_t record_words, float shared_reach, int64_t nrays,
                               float tmin, int32_t skip_id, int32_t sky_passes,
                               uint64_t *bits);

/* CPU BVH over the same triangles (binned SAH, conservative boxes). */
typedef struct vc_bvh vc_bvh;
vc_bvh *vc_bvh_build(const float *v9, const vc_tri *tris, int32_t ntri);
void vc_bvh_free(vc_bvh *b);
int32_t vc_bvh_node_count(const vc_bvh *b);
void vc_bvh_visibility(const vc_bvh *b, const float *rays, int32_t record_words,
                       float shared_reach, int64_t nrays, float tmin, int32_t skip_id,
                       int32_t sky_passes, uint64_t *bits, int32_t threads);
void vc_bvh_closest(const vc_bvh *b, const float *rays, int32_t record_words,
                    float shared_reach, int64_t nrays, float tmin, int32_t skip_id,
                    int32_t *tri_index, float *t_out, int32_t threads);
/* Shadow rays from points to point lights, correct-mode parameterisation
 * (direction = light - point, reach 1): bit (p*nlights + l). */
void vc_bvh_point_lights(const vc_bvh *b, const float *points, int64_t npoints,
                         const float *lights, int32_t nlights, float tmin,
                         uint64_t *bits, int32_t threads);

/* ---- transfers and bounce (vradcpu_bounce.c) ---- */

/* The patch system the transfer build walks (Rad/Light/Patch.cs, PatchSet.cs):
 * per patch its origin, plane normal, plane distance, area, children (-1),
 * next root patch of the same face (-1), face, cluster and a 4-corner
 * winding; per face its sky flag and first root patch; per cluster the faces
 * of its PVS row in order (VisMatrix.FacesInRow); the receivers (leaf
 * patches) in receiver order. */
typedef struct vc_patchset {
    int32_t npatch, nface, ncluster, nreceivers;
    const float *origin, *normal, *plane_dist, *area;
    const int32_t *child1, *child2, *next_parent, *face, *cluster;
    const float *winding; /* 12 per patch */
    const uint8_t *face_sky;
    const int32_t *face_parent;
    const int32_t *row_faces_ptr, *row_faces;
    const int32_t *receivers;
} vc_patchset;

float vc_diff_to_diff(const float *o1, const float *n1, const float *o2, const float *n2);
float vc_poly_to_diff(const float *poly, int32_t npts, float area, const float *o, const float *n);
void vc_make_scales(float *w, int32_t n);
int64_t vc_candidates(const vc_patchset *ps, int32_t receiver, int32_t *out, int64_t cap);
int vc_make_transfer(const vc_patchset *ps, int32_t r, int32_t s, float *trans);
void vc_transfers_build(const vc_patchset *ps, const vc_bvh *bvh, const int64_t *cand_ptr,
                        int64_t *made, int32_t *src, float *w, int32_t threads);
void vc_bounce(int32_t P, const int64_t *row_ptr, const int32_t *src, const float *w,
               const float *refl, const float *emit0, const uint8_t *sky, const int32_t *child1,
               const int32_t *child2, const float *area, const float *start, int32_t mode,
               int32_t max_iter, double tol, int32_t use_double, double *total_out,
               double *residuals, double *energies, int32_t *iters, const float *origin,
               const float *normals4, const uint8_t *bumped, double *total_bump_out);
/* (vc_bounce's last four: bumped receivers, may all be NULL; origin 3 and
 * normals4 12 floats per patch -- the plane normal, then the bump basis --
 * bumped one flag per patch, total_bump_out 9 doubles per patch.) */

/* ---- direct-light gather, sampler, encode (vradcpu_gather.c) ---- */

/* EmitType (MapFormats/Bsp/Structs/BspStructs.cs:907-925). */
#define VC_EMIT_SURFACE 0
#define VC_EMIT_POINT 1
#define VC_EMIT_SPOTLIGHT 2
#define VC_EMIT_SKYLIGHT 3
#define VC_EMIT_SKYAMBIENT 5

/* A 3D skybox (Rad/Light/SkyCameras.cs; the BSP as CompiledBspTree.LeafFromPoint,
 * Rad/Light/CompiledBspTree.cs, walks it): the cameras, as origin and 1 / scale,
 * the nodes (a plane index and two children, >= 0 a node, < 0 the leaf
 * -1 - child), the planes (normal, dist, type) and per leaf whether a sky
 * test starting in it recurses.  The same layout as the product's
 * vh_skybox_desc. */
typedef struct vc_skybox {
    int32_t ncameras;
    const float *camera_origin;
    const float *camera_world_to_sky;
    int32_t nnodes;
    const int32_t *node_plane;
    const int32_t *node_children;
    int32_t nplanes;
    const float *plane_normal;
    const float *plane_dist;
    const int32_t *plane_type;
    int32_t nleaves;
    const uint8_t *leaf_recurses;
} vc_skybox;

/* CompiledBspTree.LeafFromPoint: the leaf index of p. */
int32_t vc_bsp_leaf_from_point(const vc_skybox *s, const float p[3]);
/* ResolveSkyTest's arithmetic (DirectLightGatherer.cs:661-714, Correct
 * compliance) over given answers, as the reference's replay tests drive it:
 * first_occ 4 first-ray occlusions (lanes past count are ignored and copy
 * the last real one), inner ncam x 4 skybox-ray occlusions, mask the lanes
 * that recurse; writes 4 visible fractions. */
void vc_sky_resolve(int32_t count, const float *first_occ, int32_t ncam, const float *inner, int32_t mask,
                    float *fraction);

/* A DirectLight (Rad/Light/DirectLight.cs) as the gather reads it. */
typedef struct vc_light {
    int32_t type, style;
    float origin[3], normal[3], intensity[3], intensity_hdr[3];
    float constant_attn, linear_attn, quadratic_attn, cap_dist;
    float start_fade, end_fade, stopdot, stopdot2, exponent;
} vc_light;

typedef struct vc_gather_desc {
    int32_t nfaces, nsamples, nlights, pvs_bytes;
    const int32_t *face_sample_ptr;   /* nfaces + 1 */
    const int32_t *face_normal_count; /* 1, or 4 for a bumped face */
    const float *face_normals;        /* 12 per face: the face normal, then 3 bump normals */
    const float *sample_pos;          /* 3 per sample (on the face; +normal is the point) */
    const int32_t *sample_cluster;
    const vc_light *lights;
    const uint8_t *light_pvs; /* nlights x pvs_bytes cluster bits */
    float sun_extent;
    int32_t sun_samples;
    const float *sun_jitter; /* 3 x (sun_samples - 1) */
    int32_t sky_dirs_count;
    const float *sky_dirs;
    int32_t ao_count;
    float ao_radius;
    const float *ao_dirs; /* 3 x ao_count per face */
    /* The supersampling resample (FaceLightJob.EmitResample/ResolveResample,
     * FaceLightJob.cs:465-555): light_filter 1 keeps only the sky-ambient
     * lights (the 2 x 2 ambient points), 2 every other light (the 4 x 4
     * direct points), 0 all; no_skip adds every light's record, as
     * ResolveResample does (no all-zero skip).  Zero for the primary gather. */
    int32_t light_filter;
    int32_t no_skip;
    /* EmitResample's style test (FaceLightJob.cs:486): per face, the light
     * style the resample lights (-1 or NULL: every style). */
    const int32_t *face_style_filter;
    /* The 3D skybox behind the sky tests, or NULL (vh_gather_skybox). */
    const vc_skybox *skybox;
} vc_gather_desc;

/* Radiosity.GatherLight for flat rows, FP32, row-parallel (vradcpu_bounce.c). */
void vc_gather_light(int32_t P, const int64_t *row_ptr, const int32_t *src, const float *w,
                     const float *shoot, float *add, int32_t threads);

/* Correctly rounded powers (vradcpu_math.c): DetMathF.Pow and
 * DetMath.PowToSingle. */
float vc_pow_f32(float x, float y);
float vc_pow_to_single(double x, double y);
int32_t vc_pow_needs_exact(int32_t kind, double x, double y);
void vc_pow_f32_array(const float *x, const float *y, int64_t n, float *out);
void vc_pow_to_single_array(const double *x, const double *y, int64_t n, float *out);
int32_t vc_pow_scan(int32_t kind, double y, uint32_t first_bits, int64_t count, int32_t misrounded,
                    uint32_t *found, int32_t max_found);

void vc_directional_sampler(int32_t n, float *out);
float vc_halton(int32_t base, int32_t index);
void vc_encode_rgbe(const float *rgb, int64_t n, int32_t stride, uint8_t *out);
int64_t vc_gather(const vc_gather_desc *d, const vc_bvh *bvh, float *out_ldr, float *out_hdr,
                  float *out_ao, int32_t threads);
/* The gather with light styles: out_ldr/out_hdr hold 4 style slots x 4
 * normals x nsamples x 4 floats, face_styles 4 style numbers per face (255
 * for an empty slot).  Returns the rays traced; *dropped counts the
 * (face, light) records lost to a fifth style. */
int64_t vc_gather_styles(const vc_gather_desc *d, const vc_bvh *bvh, float *out_ldr, float *out_hdr,
                         float *out_ao, uint8_t *face_styles, int64_t *dropped, int32_t threads);
/* Plants one mistake in the gather (test infrastructure for the tests that
 * show each closed form catches it); 0 for none. */
void vc_gather_set_mutation(int32_t which);

/* The final light (vradcpu_final.c): FinalLightFace.Run with LuxelRadial,
 * one style.  Faces as FaceLightInfo frames; luxel_pos width * height points
 * per face; per sample its bounds in its own face's luxel space; neighbour
 * CSR; patches (origin, face, child1, winding CSR); the gather's samples and
 * light ([4][nsamples][4] floats); patch_light npatch x 4 x 3 (TotalLight per
 * normal) or NULL for no bounce.  luxel_pos NULL: BuildFaceLuxels' grid points.  Pinned to oracle/final.py (itself pinned by
 * the reference's LuxelRadial and FinalLightFace closed forms) on seeded
 * random faces, byte for byte. */
typedef struct vc_final_face {
    float origin[3], w2l_s[3], w2l_t[3], l2w_s[3], l2w_t[3];
    int32_t min_s, min_t, width, height;
    float minlight;
    int32_t lit;
} vc_final_face;

typedef struct vc_final_desc {
    int32_t nfaces;
    const vc_final_face *faces;
    const float *luxel_pos;
    const float *sample_bounds;
    const int32_t *nbr_ptr, *nbr;
    int32_t npatch;
    const float *patch_origin;
    const int32_t *patch_face, *patch_child1;
    const int32_t *winding_ptr;
    const float *winding;
    const int32_t *face_sample_ptr, *face_normal_count;
    const float *sample_pos;
    int32_t nsamples;
    const float *light;  /* light_slots x [4][nsamples][4] */
    const float *patch_light;
    int32_t red_errors;
    /* Light styles: 4 style numbers per face (255 empty; slot 0 style 0),
     * or NULL for style 0 alone; light then holds light_slots slots. */
    const uint8_t *face_styles;
    int32_t light_slots;
} vc_final_desc;

/* The lump size; offsets (nfaces, may be NULL) each lit face's light offset, -1 otherwise. */
int64_t vc_final_layout(const vc_final_desc *d, int32_t *offsets);
/* The lump (capacity >= vc_final_layout's size, else -1). */
int vc_final_light(const vc_final_desc *d, uint8_t *lump, int64_t capacity, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
