/*
 * vradhip.h -- C ABI of libvradhip.so, the MI355X (gfx950) backend for vrad's
 * lighting hot path.
 *
 * The reference's host is C# (.NET); it reaches this library through
 * P/Invoke.  Every signature uses plain pointers, sizes and status codes; no
 * C++ exception and no torch type crosses it.  Each declaration names the
 * reference interface it replaces (paths under the reference's src/).
 *
 * Conventions
 *   - Every function returning int returns a VH_* status.  On failure the
 *     message is available from vh_last_error() on the same thread.
 *   - Host-pointer entry points (no _dev suffix) read the caller's arrays
 *     during the call only and return when every output is written; no
 *     pointer is retained (SourceSharp.MapTools.Gpu/VulkanRayTracer.cs:246-271).
 *   - _dev entry points take device pointers (HBM-resident inputs), enqueue on
 *     `stream` (a hipStream_t; NULL = the context's own stream) and return
 *     before the work finishes; vh_sync() waits for it.
 *   - Calls on one context are serialised by a per-context mutex; results are
 *     a pure function of each call's inputs.
 */
#ifndef VRADHIP_H
#define VRADHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Errors never throw across the seam; a device that cannot
 * prove itself declines (SourceSharp.MapTools/Rad/GpuTracing.cs:28,
 * SourceSharp.MapTools.Gpu/VulkanRayTracer.cs:706-719). */
#define VH_OK 0
#define VH_DECLINED 1  /* no usable device / self-test failed: use the CPU tracer */
#define VH_EINVAL 2    /* bad argument: short output, bad record size (ArgumentException) */
#define VH_EDEVICE 3   /* a HIP call failed after open */
#define VH_ETIMEOUT 4
#define VH_ECANCELED 5
#define VH_ENOTSUP 6   /* options this kernel cannot honour (NotSupportedException) */
#define VH_ESTATE 7    /* no scene loaded, or a closed context */

/* Sky bit of a triangle id (SourceSharp.MapTools/Rad/TraceId.cs:31). */
#define VH_TRACE_ID_SKY 0x01000000

typedef struct vh_ctx vh_ctx;

typedef struct vh_open_opts {
    int32_t device;     /* HIP device ordinal; -1 = the current device */
    int32_t self_test;  /* 1 = run the known-answer self-test at open (default) */
    int32_t fail_stage; /* fault hook for tests: 0 = off; N = fail at open stage N
                           (1 device query, 2 stream, 3 self-test) without using the GPU */
    int32_t reserved;
} vh_open_opts;

/* Opens a context on one GPU.  Replaces IGpuTracerFactory.TryCreateAsync's
 * device half: VulkanRayTracer.TryCreateAsync -> TryOne (open, construct,
 * self-test), SourceSharp.MapTools.Gpu/VulkanRayTracer.cs:366,595-704.
 * Returns VH_DECLINED with the reason in vh_last_error() when no device is
 * usable; *out is then NULL. */
int vh_open(const vh_open_opts *opts, vh_ctx **out);

/* Waits for in-flight work, then frees the device buffers.  Replaces
 * VulkanRayTracer.Dispose (SourceSharp.MapTools.Gpu/VulkanRayTracer.cs:1030-1041).
 * NULL is ignored; a second close of the same pointer is the caller's bug. */
void vh_close(vh_ctx *ctx);

/* The last error message on this thread ("" if none). */
const char *vh_last_error(void);

/* IRayTracer.TracerIdentity: backend, device and kernel revision
 * (SourceSharp.MapTools/Tracing/IRayTracer.cs:221-229). */
const char *vh_tracer_identity(const vh_ctx *ctx);

/* IRayTracer.Supports (SourceSharp.MapTools/Tracing/IRayTracer.cs:193-219):
 * 1 when a visibility batch with these options gets the answer they ask
 * for.  Skip ids and sky pass-through are honoured in-kernel; only a NaN
 * minimum distance is refused. */
int vh_supports(const vh_ctx *ctx, float min_distance, int32_t skip_id, int32_t sky_passes);

/* Loads the shadow casters: n triangles as 9 floats (three vertices), the
 * caller's id per triangle and FCACHETRI_* flags (may be NULL).  Builds the
 * intersection-format triangles (KdTreeBuilder.ToIntersectionFormat, Correct
 * side, SourceSharp.MapTools/Tracing/KdTreeBuilder.cs:1568-1612) and a BVH over
 * them.  Replaces VulkanDevice.LoadScene (BLAS/TLAS,
 * SourceSharp.MapTools.Gpu/Interop/VulkanDevice.cs:1380-1608) and
 * ShadowCasterSet.BuildTracerAsync (SourceSharp.MapTools/Rad/ShadowCasterSet.cs:181). */
int vh_scene_load(vh_ctx *ctx, const float *v9, const int32_t *ids, const uint8_t *flags,
                  int32_t n);

/* Scene statistics: triangles, BVH nodes, BVH depth. */
int vh_scene_info(const vh_ctx *ctx, int32_t *triangles, int32_t *nodes, int32_t *depth);

/* The builder vh_scene_load uses: on the device (the host builder's binned
 * SAH, level by level on the GPU, the same tree) from VH_BUILD_DEVICE_MIN
 * triangles up, on host threads below; vh_scene_load_with picks one
 * explicitly (the device builder needs at least 16 triangles).  The answers
 * of every trace do not depend on the builder. */
#define VH_BUILD_AUTO 0
#define VH_BUILD_HOST 1
#define VH_BUILD_DEVICE 2
#define VH_BUILD_DEVICE_MIN 65536
int vh_scene_load_with(vh_ctx *ctx, const float *v9, const int32_t *ids, const uint8_t *flags, int32_t n,
                       int32_t builder);
/* The loaded scene as the kernels see it: nodes (16 floats each, node_floats
 * >= 16 * nodes) and triangles (12 floats each, tri_floats >= 12 * triangles,
 * in leaf order); either may be NULL. */
int vh_scene_read(vh_ctx *ctx, float *nodes, int64_t node_floats, float *tris, int64_t tri_floats);

/* IRayTracer.TraceVisibilityAsync (SourceSharp.MapTools/Tracing/IRayTracer.cs:162-166).
 * rays: n records of record_words floats (7: origin, direction, reach; 6:
 * origin, direction with shared_reach for all; SourceSharp.MapTools.Gpu/RayRecord.cs:78-144).
 * bits: LSB-first, bit i set when ray i is blocked strictly short of its
 * end; needs bit_words >= (n + 63) / 64.  min_distance: RayTraceOptions.MinDistance;
 * skip_id: RayTraceOptions.SkipId (-1 = none); sky_passes:
 * RayTraceOptions.SkyDoesNotBlock. */
int vh_trace_visibility(vh_ctx *ctx, const float *rays, int32_t record_words, float shared_reach,
                        int64_t n, float min_distance, int32_t skip_id, int32_t sky_passes,
                        uint64_t *bits, int64_t bit_words);

/* IRayTracer.TraceClosestAsync (SourceSharp.MapTools/Tracing/IRayTracer.cs:187-191):
 * per ray the caller's id of the nearest triangle with min_distance <= t <=
 * reach and its fraction (t, divided by the reach unless the reach is 1;
 * SourceSharp.MapTools.Gpu/SlabBatcher.cs:1369); a miss writes (-1, 1.0f)
 * (SourceSharp.MapTools/Tracing/HitId.cs:86).  Equal distances go to the
 * lower triangle index. */
int vh_trace_closest(vh_ctx *ctx, const float *rays, int32_t record_words, float shared_reach,
                     int64_t n, float min_distance, int32_t skip_id, int32_t *surface,
                     float *fraction, int64_t hit_slots);

/* Shadow rays from sample points to point lights, generated in-kernel in
 * vrad's correct-mode parameterisation (direction = light - point, reach 1;
 * SourceSharp.MapTools/Rad/Light/LightRayLog.cs:686-690): bit (p * nlights + l)
 * is set when the segment from point p to light l is blocked.  The same
 * answer as vh_trace_visibility over those npoints * nlights rays, without
 * materialising them (28 B a ray). */
int vh_trace_point_lights(vh_ctx *ctx, const float *points, int64_t npoints,
                          const float *lights, int32_t nlights, float min_distance,
                          uint64_t *bits, int64_t bit_words);

/* How vh_trace_point_lights and vh_trace_point_lights_dev find the
 * triangles a ray may hit (no counterpart in the reference, whose
 * LightRayLog batches go through the generic tracer; the answers are the same
 * bits either way, the exact triangle test's):
 *  - VH_POINT_LIGHTS_BVH: the rays of a wave walk the BVH together;
 *  - VH_POINT_LIGHTS_LIGHT_BUFFER: per light a cube of `cells` x `cells`
 *    direction cells on each face, each listing the triangles seen through it
 *    by depth, built on the device for each call; a ray tests its cell's list
 *    up to its own depth;
 *  - VH_POINT_LIGHTS_AUTO (the default): light buffers for batches of at least
 *    65,536 points and at least as many points as triangles, else the BVH.
 * `cells` is 4..120 (0: 64).  vh_trace_point_lights_noalloc always walks the
 * BVH. */
#define VH_POINT_LIGHTS_AUTO 0
#define VH_POINT_LIGHTS_BVH 1
#define VH_POINT_LIGHTS_LIGHT_BUFFER 2
int vh_set_point_light_method(vh_ctx *ctx, int32_t method, int32_t cells);

/* Device-pointer forms of the three traces above, enqueued on `stream`. */
int vh_trace_visibility_dev(vh_ctx *ctx, const float *d_rays, int32_t record_words,
                            float shared_reach, int64_t n, float min_distance, int32_t skip_id,
                            int32_t sky_passes, uint64_t *d_bits, void *stream);
int vh_trace_closest_dev(vh_ctx *ctx, const float *d_rays, int32_t record_words,
                         float shared_reach, int64_t n, float min_distance, int32_t skip_id,
                         int32_t *d_surface, float *d_fraction, void *stream);
int vh_trace_point_lights_dev(vh_ctx *ctx, const float *d_points, int64_t npoints,
                              const float *d_lights, int32_t nlights, float min_distance,
                              uint64_t *d_bits, void *stream);

/* vh_trace_point_lights_dev without any allocation.  No counterpart in the
 * reference: this is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same rays and the same hit bits as
 * vh_trace_point_lights_dev; its working memory (the points' spatial order)
 * is a fixed array in the library, 16.1 MiB on each GPU the library runs on,
 * held until the process ends.  At most VH_NOALLOC_MAX_POINTS points per
 * call (VH_EINVAL above that).  Every call on one GPU shares that array:
 * calls that overlap on one GPU race (use one stream, or wait between them).
 * Clears what it needs itself with fill kernels (no memset). */
#define VH_NOALLOC_MAX_POINTS 2097152
int vh_trace_point_lights_noalloc(vh_ctx *ctx, const float *d_points, int64_t npoints,
                                  const float *d_lights, int32_t nlights, float min_distance,
                                  uint64_t *d_bits, void *stream);

/* Bounded waits and cancellation, the seam's CancellationToken and fence
 * bound (SourceSharp.MapTools/Tracing/IRayTracer.cs:162-191 take a
 * CancellationToken; SourceSharp.MapTools.Gpu/VulkanRayTracer.cs:39-49 bounds
 * each fence wait at 120 s).  Every blocking call waits on the device through
 * a deadline of `ms` milliseconds (default 120000; negative: none) and gives up
 * with VH_ETIMEOUT when it passes.  vh_cancel (any thread, no lock: meant for a
 * token's callback) makes the calls in flight on the context return
 * VH_ECANCELED at their next wait, and later calls fail at once, until
 * vh_cancel_clear.  Either way the work already enqueued runs to its end on
 * the device, writing only the context's own buffers (results reach the
 * caller's arrays only after a completed wait), and the context stays usable:
 * later calls queue behind it. */
int vh_set_timeout(vh_ctx *ctx, int64_t ms);
int vh_cancel(vh_ctx *ctx);
int vh_cancel_clear(vh_ctx *ctx);

/* Waits for the work enqueued on `stream` (NULL = the context's stream),
 * bounded as above. */
int vh_sync(vh_ctx *ctx, void *stream);

/* Cumulative counters (IGpuTraceStatistics, SourceSharp.MapTools.Gpu/SlabBatcher.cs:387). */
typedef struct vh_stats {
    int64_t rays_visibility;
    int64_t rays_closest;
    int64_t calls;
    int64_t rays_gather; /* rays traced by vh_gather_run: light, sun, sky and AO */
    int64_t reserved[4];
} vh_stats;
int vh_get_stats(const vh_ctx *ctx, vh_stats *out);

/* ---- The bounce: IVradBounceStage (SourceSharp.MapTools/Rad/VradStages.cs:112-122) ---- */

/* The patch system of one lighting pass (SourceSharp.MapTools/Rad/Light/Patch.cs,
 * PatchSet.cs), as the transfer build and the bounce read it:
 *   per patch: origin (3), plane normal (3), plane distance, area, child1 and
 *     child2 (-1 for a leaf; children have higher indices than parents), the
 *     next root patch of the same face (-1), face, cluster, a four-corner
 *     winding (12 floats, wound as vrad's), reflectivity (3);
 *   per face: sky flag, first root patch (-1 if none);
 *   per cluster: the faces of its PVS row, in order (VisMatrix.FacesInRow),
 *     as row_faces_ptr (ncluster + 1 offsets) into row_faces;
 *   receivers: the leaf patches that gather light, in receiver order. */
typedef struct vh_patch_desc {
    int32_t npatch, nface, ncluster, nreceivers;
    const float *origin, *normal, *plane_dist, *area;
    const int32_t *child1, *child2, *next_parent, *face, *cluster;
    const float *winding;
    const float *reflectivity;
    const uint8_t *face_sky;
    const int32_t *face_parent;
    const int32_t *row_faces_ptr, *row_faces;
    const int32_t *receivers;
    /* Bumped patches (Patch.NeedsBumpmap; may be NULL: none): one flag per
     * patch and, per patch, the three bump basis normals (9 floats) the
     * gather re-applies the cosine on (Radiosity.PrepareNormals,
     * Rad/Bounce/Radiosity.cs:322-356; the flat slot uses the plane normal). */
    const uint8_t *bumped;
    const float *bump_normals;
} vh_patch_desc;

/* Uploads the patch system; replaces any previous one and its transfers. */
int vh_patches_load(vh_ctx *ctx, const vh_patch_desc *desc);

/* Builds every receiver's transfer list on the GPU against the loaded scene:
 * VisMatrix.BuildAsync (SourceSharp.MapTools/Rad/Bounce/VisMatrix.cs:219-411) --
 * candidate enumeration (Enumerator.Run, TestPatchToFace, TestPatchToPatch),
 * one visibility ray per candidate (VisMatrix.TransferRay, StockExact),
 * MakeTransfer with FormFactors.DiffToDiff / PolyToDiff, MakeScales.  Rows keep
 * the reference's build order.  Outputs may be NULL. */
int vh_transfers_build(vh_ctx *ctx, int64_t *transfers, int64_t *candidates, int64_t *blocked);

/* Copies the transfer lists out as CSR: row_ptr (npatch + 1), src and weight
 * (row_ptr[npatch] entries; capacity is their length).  With src and weight
 * NULL only row_ptr is written, so a caller can size the arrays first. */
int vh_transfers_read(vh_ctx *ctx, int64_t *row_ptr, int32_t *src, float *weight,
                      int64_t capacity);

/* Loads transfer lists built elsewhere (a TransferSet, or the TRF1 cache,
 * SourceSharp.MapTools/Rad/Bounce/TransferSetCodec.cs:23-80): row_ptr has
 * npatch + 1 entries. */
/* The transfers in the reference's transfer-cache layout
 * (TransferSetCodec.Index / Chunks, Rad/Bounce/TransferSetCodec.cs:23-98):
 * counts[p] = patch p's list length (npatch entries), and records = every
 * transfer, patch after patch, as 8 little-endian bytes (int32 source patch,
 * float32 weight), written on the device.  Either may be NULL; records needs
 * capacity >= the transfer count. */
int vh_transfers_export(vh_ctx *ctx, int32_t *counts, void *records, int64_t capacity);
int vh_transfers_load(vh_ctx *ctx, const int64_t *row_ptr, const int32_t *src,
                      const float *weight);

#define VH_WEIGHTS_F32 0
#define VH_WEIGHTS_F16 1 /* FP16 weights with a per-row FP32 scale, FP32 accumulation */
#define VH_BOUNCE_PROGRESSIVE 0 /* Radiosity.BounceAsync: emit = the last bounce's light */
#define VH_BOUNCE_FROM_START 1  /* start from an estimate of the bounced light */
#define VH_BOUNCE_FROM_PREDICTION 2 /* from a start: the device-resident vh_mlp_predict result */

typedef struct vh_bounce_opts {
    int32_t weights;   /* VH_WEIGHTS_* */
    int32_t mode;      /* VH_BOUNCE_* */
    int32_t max_iter;  /* the bounce count (VradOptions.Bounces, default 100) */
    int32_t use_graph; /* replay one captured iteration (hipGraph) */
    double tol;        /* > 0: stop at relative residual <= tol (L1, FP64);
                          <= 0: the reference's rule, every channel of the
                          summed emission below 1.0 (Radiosity.cs:163) */
} vh_bounce_opts;

/* Radiosity.BounceAsync (SourceSharp.MapTools/Rad/Bounce/Radiosity.cs:132-170)
 * for flat patches: MoveDirectLightToEmission, then per iteration GatherLight
 * over every receiver's transfers and CollectLight (leaves take the gathered
 * light, parents the area-weighted light of their children, sky emits 0).
 * emit0: each patch's direct light (3 per patch; NULL = the device-resident
 * result of the last vh_patch_lights); start: the bounced-light
 * estimate for VH_BOUNCE_FROM_START (3 per patch; NULL = zero).  Writes
 * total_out (3 per patch: the bounced light, parents pulled up), and per
 * iteration the residual and the summed leaf emission (3 doubles). */
int vh_bounce(vh_ctx *ctx, const vh_bounce_opts *opts, const float *emit0, const float *start,
              float *total_out, double *residuals, double *energies, int32_t *iters);

/* The last vh_bounce's TotalLight in the three bump directions (9 floats per
 * patch, zero for a flat patch).  A bumped receiver gathers as
 * GatherBumpedScalar / GatherBumpedFour do (Rad/Bounce/Radiosity.cs:406-512):
 * per transfer the direction to the source normalised, the receiver's cosine
 * divided out, and the light added along each of its four normals that the
 * source is not behind; its flat slot is vh_bounce's total_out.  Leaves
 * collect, bumped parents take their children's area-weighted blend. */
int vh_bounce_bump_totals(vh_ctx *ctx, float *out);

/* ---- The direct-light gather: RadWorld.LightFacesAsync (SourceSharp.MapTools/Rad/Light/RadWorld.Facelights.cs:194) ---- */

/* EmitType (SourceSharp.MapFormats/Bsp/Structs/BspStructs.cs:907-925). */
#define VH_EMIT_SURFACE 0
#define VH_EMIT_POINT 1
#define VH_EMIT_SPOTLIGHT 2
#define VH_EMIT_SKYLIGHT 3
#define VH_EMIT_SKYAMBIENT 5

/* A DirectLight (SourceSharp.MapTools/Rad/Light/DirectLight.cs) as the gather
 * reads it; intensity is the LDR pass's, intensity_hdr the HDR pass's (the
 * two passes differ only there, DirectLightBuilder.cs:218-231).  style is the
 * light style, 0-254 (the slots are described at vh_gather_styles). */
typedef struct vh_light {
    int32_t type, style;
    float origin[3], normal[3], intensity[3], intensity_hdr[3];
    float constant_attn, linear_attn, quadratic_attn, cap_dist;
    float start_fade, end_fade, stopdot, stopdot2, exponent;
} vh_light;

/* The faces to light.  Samples of face f are [face_sample_ptr[f],
 * face_sample_ptr[f+1]); each sample's point is its position plus the face
 * normal (FaceLightJob.ComputeIlluminationPointAndNormals,
 * SourceSharp.MapTools/Rad/Light/FaceLightJob.cs:237-327, flat faces);
 * face_normals holds the face normal and, for a bumped face
 * (face_normal_count 4), its three bump basis normals.  light_pvs holds one
 * row of cluster bits per light (pvs_bytes each; 0 = every light reaches
 * every cluster).  Sun: sun_samples rays per sample (SunAreaLightSamples = 30
 * in the reference), jittered by sun_jitter (3 x (sun_samples - 1)) scaled by
 * the angular extent.  Sky ambient: sky_dirs_count directions (162 x
 * SkySampleScale, from DirectionalSampler).  Ambient occlusion (no reference
 * counterpart): ao_count directions per face (ao_dirs, 3 x ao_count per face),
 * any hit within ao_radius. */
typedef struct vh_gather_desc {
    int32_t nfaces, nsamples, nlights, pvs_bytes;
    const int32_t *face_sample_ptr;
    const int32_t *face_normal_count;
    const float *face_normals;
    const float *sample_pos;
    const int32_t *sample_cluster;
    const vh_light *lights;
    const uint8_t *light_pvs;
    float sun_extent;
    int32_t sun_samples;
    const float *sun_jitter;
    int32_t sky_dirs_count;
    const float *sky_dirs;
    int32_t ao_count;
    float ao_radius;
    const float *ao_dirs;
} vh_gather_desc;

/* Uploads the faces, samples, lights and tables (kept on the device). */
int vh_gather_load(vh_ctx *ctx, const vh_gather_desc *desc);

/* The 3D skybox behind the gather's sky rays: the map's sky cameras
 * (SkyCameras, SourceSharp.MapTools/Rad/Light/SkyCameras.cs) and the compiled
 * BSP that CompiledBspTree.LeafFromPoint (Rad/Light/CompiledBspTree.cs)
 * descends.  With a camera loaded every sky test (TestLineDoesHitSky,
 * Rad/Light/DirectLightGatherer.cs:510-714: each sun ray, and each sky-ambient
 * direction, which then traces every real lane) continues into the skyboxes:
 * a lane whose start lies in a leaf with leaf_recurses set, in a 4-sample group
 * whose first rays did not all find something other than sky, casts one more
 * ray per camera, from origin + start * world_to_sky along its first ray's
 * normalised direction, MaxTraceLength long; a hit short of the end that is not
 * sky adds 1 to its occlusion, which is clamped to [0, 1].  Each lane looks up
 * its own start (the Correct side of StockQuirk.SkyboxRecursionFromLaneZero).
 * Nodes (dnodes): a plane index and two children, >= 0 a node and < 0 the leaf
 * -1 - child; a child past the last node is leaf 0.  Planes (dplanes): normal,
 * dist and type; a type below 3 tests the one coordinate it names (0 x, 1 y,
 * any other z), else the dot product; a point on the plane is in front.
 * leaf_recurses: per leaf 1 when the leaf's area is one of the map's areas and
 * no camera stands in it.  The skybox stays until replaced; ncameras 0 (or a
 * NULL desc) removes it.  VH_EINVAL: more than VH_SKYBOX_MAX_CAMERAS cameras,
 * a non-finite camera, a plane index out of range, or nodes that reach a node
 * twice from node 0. */
#define VH_SKYBOX_MAX_CAMERAS 64
typedef struct vh_skybox_desc {
    int32_t ncameras;
    const float *camera_origin;       /* 3 per camera */
    const float *camera_world_to_sky; /* 1 / scale per camera */
    int32_t nnodes;
    const int32_t *node_plane;
    const int32_t *node_children; /* 2 per node */
    int32_t nplanes;
    const float *plane_normal; /* 3 per plane */
    const float *plane_dist;
    const int32_t *plane_type;
    int32_t nleaves;
    const uint8_t *leaf_recurses;
} vh_skybox_desc;
int vh_gather_skybox(vh_ctx *ctx, const vh_skybox_desc *desc);

#define VH_SCHEDULE_QUEUE 0    /* one persistent launch; waves take 64-sample chunks from a queue */
#define VH_SCHEDULE_PER_FACE 1 /* one launch per face (the unbatched form) */
#define VH_SCHEDULE_GRID 2     /* one launch, one wave per chunk */

/* Lights every sample against the loaded scene: FaceLightJob.EmitGroup /
 * ResolveGroup over 4-sample groups (FaceLightJob.cs:329-463) with
 * DirectLightGatherer's standard, sky and sky-ambient gathers
 * (DirectLightGatherer.cs:441-1306) and LightingValue.AddLight, visibility
 * traced inline.  Results stay on the device; vh_gather_read copies them.
 * waves_per_cu sizes the persistent launch (0 = a default). */
int vh_gather_run(vh_ctx *ctx, int32_t schedule, int32_t waves_per_cu);
/* vh_gather_run with the queue's grain: each persistent wave takes
 * chunks_per_ticket consecutive 64-sample chunks per ticket (1-64; 1 is
 * vh_gather_run).  Larger grains take fewer atomics and keep a wave on one
 * face longer; smaller ones balance the queue's tail.  Ignored by the other
 * schedules.  The autotuner picks it with waves_per_cu. */
int vh_gather_run_shaped(vh_ctx *ctx, int32_t schedule, int32_t waves_per_cu, int32_t chunks_per_ticket);

/* Supersampling, FaceLightJob's passes after the gather (StyleSupersample,
 * EmitSupersamplePoint / ResolveSupersampleItem,
 * SourceSharp.MapTools/Rad/Light/FaceLightJob.cs:698-1207), over the last
 * vh_gather_run's LDR (which = 0) or HDR (1) result, refined in place:
 *   - per sample and normal, intensity = (float)pow((R + G + B) / 256.0, 1/2.2);
 *   - per pass (1 .. extra_passes, while the last one selected any), every
 *     sample not yet supersampled whose largest intensity step to its eight
 *     neighbouring luxels (every normal) is not below 0.0625 is selected;
 *   - a selected sample is re-lit at 4 x 4 points (every light but sky
 *     ambient) and 2 x 2 points (sky ambient only) spread over its luxel
 *     around WorldToLuxel(sample), taking its cluster, every light's record
 *     added; its light becomes direct / 16 + ambient / 4.
 * faces: the gather's faces' luxel frames (vh_final_face, only the frame and
 * width x height are read); sample_st: per sample its luxel (s, t), or NULL
 * when sample k of a face sits on luxel (k % width, k / width).
 * unset_is_garbage 1 reads a neighbouring luxel without a sample as infinity
 * (stock's uninitialised stack); 0 leaves it out (the Correct side).  Samples
 * carry no partial-luxel winding: every point counts.  pass_counts (NULL or
 * extra_passes entries) receives the number of samples each pass re-lit. */
#define VH_SUPERSAMPLE_MAX_PASSES 64
typedef struct vh_supersample_desc {
    const struct vh_final_face *faces;
    const int32_t *sample_st;
    int32_t which;
    int32_t extra_passes; /* DirectLightingSettings.ExtraPasses: 4 */
    int32_t unset_is_garbage;
} vh_supersample_desc;
int vh_gather_supersample(vh_ctx *ctx, const vh_supersample_desc *desc, int32_t *pass_counts);

/* Copies the last run's results for style slot 0 (light style 0): ldr/hdr 4
 * floats (RGB, direct sun amount) per (normal, sample), normal-major over 4
 * normal slots; ao one float per sample.  Any pointer may be NULL. */
int vh_gather_read(vh_ctx *ctx, float *ldr, float *hdr, float *ao);

/* Light styles (FaceLight.Styles / FindOrAllocateStyle,
 * SourceSharp.MapTools/Rad/Light/FaceLight.cs:98-222; FaceLightJob.cs:185,435):
 * each face has up to MaxLightmaps = 4 style slots; slot 0 is style 0, and
 * the other styles of the lights that reach the face take slots 1-3 in the
 * order they first touch it (group by group, light by light); a fifth style's
 * light is dropped, as the reference's "Too many light styles" warning does.
 * A light's style is vh_light.style (0-254).  vh_gather_styles copies the
 * last run's 4 style numbers per face (255 for an empty slot), the number of
 * slots any face uses (slots, 1-4) and the (face, style) pairs dropped;
 * vh_gather_read_slot copies one slot's light as vh_gather_read does slot 0
 * (zeros for a slot no face uses, or where a face has no style in it).
 * Supersampling (vh_gather_supersample) runs its passes per slot, each
 * resampling only its own style's lights; vh_patch_lights reads slot 0. */
int vh_gather_styles(vh_ctx *ctx, uint8_t *face_styles, int32_t *slots, int64_t *dropped);
int vh_gather_read_slot(vh_ctx *ctx, int32_t slot, float *ldr, float *hdr);

/* StockLightColor.Encode (SourceSharp.MapTools/Rad/Ambient/StockLightColor.cs:136-185)
 * into ColorRGBExp32 (4 bytes: R, G, B, signed exponent) for n colours read
 * `stride` floats apart (RGB first). */
int vh_encode_rgbe(vh_ctx *ctx, const float *rgb, int64_t n, int32_t stride, uint8_t *out);

/* The lighting lump of the last gather: for each face f (in order), style 0,
 * each normal b, each sample j, the RGBE of ldr (which = 0) or hdr (1), at
 * byte offset face_base(f) + (b * numSamples(f) + j) * 4, face_base the
 * running sum of 4 * normals * samples (FinalLightFace.cs:270 with one style).
 * out needs 4 * sum(normals * samples) bytes. */
int vh_lightmap_encode(vh_ctx *ctx, int32_t which, uint8_t *out, int64_t capacity);  /* style 0 */

/* PatchLighting.BuildPatchLights with AddSampleToPatch
 * (SourceSharp.MapTools/Rad/Light/PatchLighting.cs:65-249): the last gather's
 * flat style-0 light (which = 0 LDR, 1 HDR) handed to the loaded patches.
 * The gather's faces must be the patch system's faces, in the same order;
 * sample_area gives each gather sample's area (nsamples floats).  Each
 * non-sky leaf patch sums, in sample order, area and area * light over the
 * samples of its face whose average light is at least 1 and whose position
 * grown by sqrt(area)/2 overlaps the patch winding's bounds, and takes
 * light sum * (float)(1 / area sum); parents take their children's
 * area-weighted blend.  The result stays on the device as the next
 * vh_bounce's emission when that call passes emit0 = NULL; direct (3 per
 * patch, may be NULL) receives a copy. */
int vh_patch_lights(vh_ctx *ctx, int32_t which, const float *sample_area, float *direct);

/* ---- The final light: FinalLighting.RunAsync / FinalLightFace.Run
 * (SourceSharp.MapTools/Rad/Final/FinalLightFace.cs:93-299) ----
 *
 * One light style, faces that are not displacements, macro textures the
 * identity, the Correct sides of SampleRadialEdgeOffByOne and
 * PatchRadialNeighbourBumpFromSelf. */

/* FaceLightInfo's luxel frame (SourceSharp.MapTools/Rad/Light/FaceLightInfo.cs:264-290)
 * and the face's final-light inputs. */
typedef struct vh_final_face {
    float origin[3];       /* LuxelOrigin */
    float w2l_s[3], w2l_t[3]; /* WorldToLuxelS / T */
    float l2w_s[3], l2w_t[3]; /* LuxelToWorldS / T */
    int32_t min_s, min_t;  /* LuxelMinS / T */
    int32_t width, height; /* the luxel grid (LightmapTextureSizeInLuxels + 1 each) */
    float minlight;        /* the face entity's _minlight * 128 */
    int32_t lit;           /* 0: TEX_SPECIAL or no light style; the face gets no lump bytes */
} vh_final_face;

/* The faces are the loaded gather's, in its order; a face is bumped when its
 * gather normal count is 4.  luxel_pos: width * height world points per face
 * (FaceLight.Luxels, face after face; luxel j of a face is its lump entry j),
 * or NULL for BuildFaceLuxels' points, LuxelToWorld(s, t) of every grid point
 * s + t * width (Rad/Light/FaceSampleBuilder.cs:242-259), made on the device.
 * sample_bounds: per gather sample, min s, min t, max s, max t in its own
 * face's luxel space (LightSample.MinS ... MaxT).  Neighbours
 * (FaceNeighbours.Neighbours) as CSR: nbr_ptr (nfaces + 1) into nbr.  The
 * patch radial's patches (npatch may be 0): origin (3), face, child1 (-1 for
 * a leaf) and winding points (winding_ptr, npatch + 1 offsets, into winding,
 * 3 floats a point); each face walks its leaf patches highest index first,
 * the order PatchSubdivider.RebuildFaceLists leaves FacePatches in
 * (Rad/Light/PatchSubdivider.cs:408-423).  Everything is copied. */
typedef struct vh_final_desc {
    int32_t nfaces;
    const vh_final_face *faces;
    const float *luxel_pos;
    const float *sample_bounds;
    const int32_t *nbr_ptr, *nbr;
    int32_t npatch;
    const float *patch_origin;
    const int32_t *patch_face, *patch_child1;
    const int32_t *winding_ptr;
    const float *winding;
} vh_final_desc;

/* Uploads the final-light inputs; needs the gather loaded (vh_gather_load)
 * with the same faces. */
int vh_final_load(vh_ctx *ctx, const vh_final_desc *desc);

/* LightmapOffsets.Compute for one style (Rad/Light/LightmapOffsets.cs): each
 * lit face reserves 4 bytes for its average colour, then 4 bytes per luxel per
 * normal.  offsets (nfaces, may be NULL) receives each face's light offset,
 * just past its average word (-1 for a face that is not lit); bytes (may be
 * NULL) the lump's size. */
int vh_final_layout(vh_ctx *ctx, int32_t *offsets, int64_t *bytes);

#define VH_FINAL_NO_BOUNCE 0     /* no patch radial (Bounces = 0) */
#define VH_FINAL_BOUNCE_DEVICE 1 /* patch lights = the last vh_bounce's TotalLight, flat and
                                    bump directions (a bumped face may own leaf patches
                                    only when the patches were loaded with bump flags) */
#define VH_FINAL_BOUNCE_HOST 2   /* patch lights from the patch_light argument */

typedef struct vh_final_opts {
    int32_t which;      /* 0 = the last gather's LDR light, 1 = HDR */
    int32_t bounce;     /* VH_FINAL_* */
    int32_t red_errors; /* -rederrors: a luxel nothing reached is 2550 red, not black */
} vh_final_opts;

/* FinalLightFace.Run for every lit face: LuxelRadial.AddDirect of the face's
 * samples, then each lit neighbour's re-projected by its bounds' corners
 * (BuildLuxelRadial, FinalLightFace.cs:301-389); with a bounce,
 * LuxelRadial.AddBounced of the face's and its neighbours' leaf patches
 * (BuildPatchRadial, :404-416, 489-536); per luxel Sample of both grids
 * (LuxelRadial.cs:314-359) added, a face without samples painted
 * (255, 0, 0), max(value, minlight), the luxels' ColorRGBExp32 at
 * offset + (b * numLuxels + j) * 4 and the median of the flat values of the
 * luxels whose flat sample had weight (element count / 2 of the sorted
 * values, per channel) at offset - 4.  patch_light: TotalLight per normal,
 * 4 x 3 floats per patch, for VH_FINAL_BOUNCE_HOST (else ignored).  lump
 * receives vh_final_layout's bytes (capacity at least that). */
int vh_final_light(vh_ctx *ctx, const vh_final_opts *opts, const float *patch_light, uint8_t *lump,
                   int64_t capacity);

/* ---- The leaf-ambient ray fan: AmbientSampler.ComputeRayCube
 * (SourceSharp.MapTools/Rad/Ambient/AmbientSampler.cs), one light style ----
 *
 * From each sample point, one ray along each of the 162 vertex normals
 * (Rad/Ambient/VertexNormals.cs) out to AmbientCube.RayLength; the closest
 * hit over the loaded scene (the scene's caller ids carry the face index in
 * their low 24 bits); the ray's light by RayAmbientLighting.Accumulate
 * (Rad/Ambient/RayAmbientLighting.cs:40-110, 160-300): the hit face's average
 * colour and the luxel under the hit, blended by the cone's radius at the hit
 * (RemapValClamped 20..40) and tinted by the face's reflectivity, or the
 * sky-ambient colour for a sky face; then AmbientCube.Project
 * (Rad/Ambient/AmbientCube.cs) onto +x, -x, +y, -y, +z, -z, and the flagged
 * world lights on top (AmbientSampler.ComputeCube).  The reference walks the
 * BSP for the hit (AmbientRayTracer); this walks the BVH. */

/* One face as AmbientScene shades it (its TraceSurface and FaceShade). */
typedef struct vh_ambient_face {
    float s[4], t[4];      /* lightmap vectors: luxel s = dot(p, s[0..2]) + s[3], likewise t */
    float min_s, min_t;    /* the lightmap's mins in luxels */
    int32_t smax, tmax;    /* the luxel grid (LightmapTextureSizeInLuxels + 1 each) */
    int32_t light_ofs;     /* the lump offset of luxel 0 of style 0; -1 for no lightmap */
    int32_t sky;           /* SURF_SKY */
    float reflectivity[3];
} vh_ambient_face;

/* One world light (DWorldLight) as the cubes read it. */
typedef struct vh_ambient_light {
    float origin[3], normal[3], intensity[3];
    float radius;          /* 0 = unlimited */
    int32_t type, style;   /* VH_EMIT_*, the light style */
} vh_ambient_light;

typedef struct vh_ambient_desc {
    int32_t nfaces;
    const vh_ambient_face *faces;
    int32_t has_sky_ambient; /* a sky-ambient light exists (RayAmbientLighting.FindSkyAmbient) */
    float sky_ambient[3];    /* its intensity */
    const uint8_t *lump;     /* the lighting lump; NULL = the last vh_final_light's, on the device */
    int64_t lump_bytes;
    int32_t nlights;         /* the pass's world lights (may be 0) */
    const vh_ambient_light *lights;
} vh_ambient_desc;

/* Uploads the faces (copied) and the lump, and decodes each face's
 * AverageTinted (AmbientScene.BuildShades).  The world lights are classified
 * as LeafAmbientSurfaceLights.Classify does (Rad/Ambient/LeafAmbientSurfaceLights.cs,
 * Correct side): surface lights of style 0 whose brightest channel times
 * 1/512^2 is under 0.005 go into the cubes, each added after the fan
 * (AmbientSampler.AddSurfaceLights, AmbientCube.AddEmitSurfaceLights) when the
 * segment from the point to the light is not blocked. */
int vh_ambient_load(vh_ctx *ctx, const vh_ambient_desc *desc);
/* The ambient cube (6 x 3 floats, box directions in the order above) of each
 * of n points (3 floats each), against the loaded scene. */
int vh_ambient_cubes(vh_ctx *ctx, const float *points, int64_t n, float *cubes);
/* The fan's 162 directions (486 floats), as the kernels use them. */
int vh_ambient_directions(float *out);

/* ---- The warm-start network (no counterpart in the reference) ----
 *
 * A per-patch MLP, 25 -> 128 -> 128 -> 128 -> 3 with GELU (tanh form) on
 * the hidden layers, that predicts a patch's converged bounced light; the
 * bounce then iterates from the prediction (VH_BOUNCE_FROM_PREDICTION) to the
 * same residual, so the network changes where the iteration starts, never
 * what it converges to.  Inputs of patch p: 2 (origin - lo) / (hi - lo) - 1
 * over the bounds of the loaded patch origins (3), the plane normal (3),
 * log(max(direct light, 0) + floor) (3), and sin then cos of
 * 2 pi (F origin) / length for the 8 rows of the frequency matrix F (16).
 * With output y the prediction is max(gain * direct light + scale * y, 0):
 * the network learns, in linear light, the correction to a start
 * proportional to the patch's own direct light.  Parameters are FP32, packed per
 * layer as W [in][out] row-major then b: 36,739 floats.  GEMMs run on BF16
 * MFMA with FP32 accumulation; Adam keeps FP32 moments. */
#define VH_MLP_PARAMS 36739
#define VH_MLP_INPUTS 25

typedef struct vh_mlp_opts {
    float length; /* the Fourier features' length scale (> 0) */
    float floor;  /* the direct-light input is log(max(x, 0) + floor) (> 0) */
    float gain;   /* prediction = max(gain * direct + scale * y, 0) */
    float scale;  /* (> 0) */
} vh_mlp_opts;
/* Sets the parameters (VH_MLP_PARAMS floats), the 8 x 3 frequency matrix and
 * the options, and clears the optimiser state; a data set already built
 * (vh_mlp_dataset) is kept. */
int vh_mlp_init(vh_ctx *ctx, const float *params, const float *freq, const vh_mlp_opts *opts);
/* Reads the FP32 parameters (VH_MLP_PARAMS floats). */
int vh_mlp_params(vh_ctx *ctx, float *out);
/* The raw network output y for n host rows of VH_MLP_INPUTS. */
int vh_mlp_forward(vh_ctx *ctx, const float *x, int64_t n, float *y);
/* Mean squared error over n rows x 3 outputs against t (n x 3) and its
 * gradient (VH_MLP_PARAMS floats); n a positive multiple of 32. */
int vh_mlp_grad(vh_ctx *ctx, const float *x, const float *t, int64_t n, float *grad, float *loss);
/* One Adam step with the given gradient (VH_MLP_PARAMS floats). */
int vh_mlp_adam(vh_ctx *ctx, const float *grad, float lr, float beta1, float beta2, float eps);
/* The inputs of every loaded patch (npatch x VH_MLP_INPUTS) from emit0
 * (3 per patch; NULL = the last vh_patch_lights result). */
int vh_mlp_features(vh_ctx *ctx, const float *emit0, float *x);

#define VH_MLP_DATA_CLEAR 0
#define VH_MLP_DATA_ADD 1 /* the last vh_bounce's leaf, non-sky patches: inputs from its
                             direct light; targets (bounced - gain * direct) / scale,
                             formed per batch from the options of vh_mlp_init */
/* The device-resident training set; rows (may be NULL) receives its size. */
int vh_mlp_dataset(vh_ctx *ctx, int32_t op, int64_t *rows);

typedef struct vh_mlp_train_opts {
    int32_t epochs;
    int32_t batch;     /* rows per step, a multiple of 32 */
    int32_t seed;      /* the per-epoch row permutations */
    int32_t use_graph; /* replay one captured training step (hipGraph) */
    float lr, beta1, beta2, eps;
    int32_t fused;     /* the two-launch step (rows, then weights + Adam): the
                        * same parameters, bit for bit, as the per-layer kernels */
} vh_mlp_train_opts;
/* Trains on the data set: per epoch a fresh permutation of the rows, whole
 * batches only; epoch_loss (epochs floats, may be NULL) receives each
 * epoch's mean batch loss. */
int vh_mlp_train(vh_ctx *ctx, const vh_mlp_train_opts *opts, float *epoch_loss);
/* Predicts the bounced light of the loaded patches (leaves that are not sky;
 * parents pulled up) from emit0 (3 per patch; NULL = the last
 * vh_patch_lights result) and keeps it on the device for the next
 * VH_BOUNCE_FROM_PREDICTION bounce; start (3 per patch, may be NULL)
 * receives a copy. */
int vh_mlp_predict(vh_ctx *ctx, const float *emit0, float *start);

/* Kernel timing with HIP events recorded on the launching stream around each
 * launch of the named kernel (no counterpart in the reference: the bench and
 * the profiler read it).  Enable, run, then read: vh_timing_read waits for the
 * recorded events and returns the summed milliseconds and the launch count
 * since the last read or enable. */
#define VH_KERNEL_POINT_LIGHTS 0
#define VH_KERNEL_VISIBILITY 1
#define VH_KERNEL_CLOSEST 2
#define VH_KERNEL_GATHER 3
#define VH_KERNEL_TRANSFERS 4
#define VH_KERNEL_FACELIGHT 5
#define VH_KERNEL_MLP 6 /* the warm-start network's GEMMs (inference) or training steps */
#define VH_KERNEL_FINAL 7 /* vh_final_light's launches, bracketed together */
#define VH_KERNEL_AMBIENT 8 /* the leaf-ambient fan */
#define VH_KERNEL_SUPERSAMPLE 9 /* vh_gather_supersample's launches */
#define VH_KERNEL_LIGHT_BUFFER 10 /* a light-buffer build: cell counts, scan, lists, sort
                                     (with the host's read of the list size between them) */
#define VH_KERNEL_COUNT 11
int vh_timing_enable(vh_ctx *ctx, int32_t on);
int vh_timing_read(vh_ctx *ctx, int32_t kernel, double *total_ms, int64_t *launches);

/* Work counters (no counterpart in the reference): while enabled, traversal
 * launches also count BVH node visits (two child boxes each) and triangle
 * tests; the roofline's algorithmic FLOPs are computed from them. */
int vh_count_enable(vh_ctx *ctx, int32_t on);
int vh_count_read(vh_ctx *ctx, int64_t *node_visits, int64_t *triangle_tests);
/* The lane slots the point-light packet walks occupied while counting: 64 per
 * node visit or triangle fetch of a wave, whatever lanes took part, so
 * node_visits / node_slots is the walks' packet efficiency (no counterpart in
 * the reference).  A light-buffer trace counts instead: node_visits = list
 * entries read, triangle_tests = exact tests, node_slots = 64 per list step of
 * a wave (as many steps as its longest lane), triangle_slots = 0. */
int vh_count_read_slots(vh_ctx *ctx, int64_t *node_slots, int64_t *triangle_slots);

#ifdef __cplusplus
}
#endif
#endif
