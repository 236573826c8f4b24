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

#ifdef __cplusplus
}
#endif
#endif
