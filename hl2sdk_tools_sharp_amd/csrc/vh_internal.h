// Internal declarations shared by the libvradhip.so sources.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vradhip.h"

namespace vh {

// Thread-local last error (vh_last_error).
void set_error(const std::string &msg);
void clear_error();

// A device allocation that only grows; reused across calls so a batch does
// not pay hipMalloc each time.
// Device memory of one context.  Blocks come from, and go back to, a
// process-wide cache per device (vh_context.cpp), so a context opened for
// each lighting pass does not pay hipMalloc and hipFree for its buffers
// again.  release() keeps hipFree's guarantee that the device is done with
// the block: it synchronizes the device first, unless the caller has
// already waited for the context's stream (ReleaseIdle).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int dev = -1;
    int ensure(size_t need, const char *what);
    void release();
};

// While one of these is alive on a thread, release() skips the device
// synchronization: the context's stream has been waited for.
struct ReleaseIdle {
    ReleaseIdle();
    ~ReleaseIdle();
};

// Scene layout in HBM.
//
// BVH node, 64 bytes = 4 x float4, both children's boxes in the parent so one
// node visit tests two boxes with one 64-byte load:
//   q0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   q1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   q2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   q3 = (ref0, ref1, -, -) as int bits
// A child ref >= 0 is an internal node index; ref < 0 is a leaf whose
// triangles start at ~ref and run to the one carrying kTriLast; kEmptyRef is
// a child with nothing in it (its box is empty, lo > hi).
//
// Triangle, 48 bytes = 3 x float4, the reference's intersection format
// (Tracing/KdTree.cs:121-167) repacked for aligned loads:
//   t0 = (nx, ny, nz, d)
//   t1 = (e0, e1, e2, e3)
//   t2 = (e4, e5, id bits, meta bits)
//   meta = cs0 | cs1 << 2 | last-in-leaf << 4 | original index << 5
constexpr int32_t kEmptyRef = (int32_t)0x80000000;
constexpr uint32_t kTriLast = 16u;
constexpr int kOrigShift = 5;
constexpr int kMaxDepth = 32;      // stack entries per lane; the builder caps depth here
constexpr int kMaxOrigIndex = 1 << 27;

struct HostScene {
    std::vector<float> nodes;  // 16 floats per node
    std::vector<float> tris;   // 12 floats per triangle
    int32_t ntri = 0;
    int32_t nnodes = 0;
    int32_t depth = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    float pad = 0;
};

// Builds the intersection-format triangles and the BVH (vh_bvh.cpp).
int build_scene(const float *v9, const int32_t *ids, const uint8_t *flags, int32_t n,
                HostScene &out);

// The same build on the GPU, into the context's scene buffers (vh_build.hip).
int build_scene_device(vh_ctx *c, const float *v9, const int32_t *ids, int32_t n);

}  // namespace vh

struct vh_ctx {
    int device = 0;
    int cu_count = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    std::mutex mu;
    // bounded waits: a blocking call gives up after timeout_ms (VH_ETIMEOUT)
    // and at a cancel request (VH_ECANCELED); the work goes on on the device
    std::atomic<int64_t> timeout_ms{120000};
    std::atomic<int> cancel{0};
    std::string identity;
    // scene
    vh::DevBuf nodes, tris;
    int32_t nnodes = 0, ntri = 0, depth = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool have_scene = false;
    uint64_t scene_gen = 0;  // changes with every scene load
    // the caller's triangles (9 floats each, in the caller's order), kept by
    // vh_scene_load for the light buffers' rasterisation
    vh::DevBuf verts;
    bool have_verts = false;
    // point-light method (VH_POINT_LIGHTS_*) and the light buffers' cells per
    // cube-face edge (0: the default)
    int32_t pl_method = 0, lb_cells = 0;
    // light buffers (vh_lightbuf.hip): per-triangle boxes and vertices (leaf
    // order) of the scene generation lb_boxes_gen, cell offsets and counts, the lists as
    // (depth, triangle) and as triangle records, and a pinned word for the
    // list size
    vh::DevBuf lb_boxes, lb_tverts, lb_off, lb_cur, lb_ent, lb_rec;
    uint64_t lb_boxes_gen = ~0ull;
    uint32_t *lb_pinned = nullptr;
    // the 3D skybox behind the gather's sky rays (vh_gather_skybox): cameras
    // as (origin, world_to_sky), per BSP node (children, plane type) and its
    // plane (normal, dist), per leaf whether a start there recurses
    vh::DevBuf sb_cam, sb_nodes, sb_planes, sb_recurse;
    int32_t sb_ncam = 0, sb_nnodes = 0, sb_nleaves = 0;
    // per-call staging
    vh::DevBuf rays_buf, out_buf, out2_buf, pts_buf, lights_buf;
    vh::DevBuf keys_a, keys_b, perm_a, perm_b, sort_tmp;
    vh_stats stats{};
    // kernel timing (vh_timing_*): event pairs per kernel, resolved on read
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[VH_KERNEL_COUNT];
    std::vector<hipEvent_t> event_pool;
    // work counters (vh_count_*): [node visits, triangle tests, node slots,
    // triangle slots]
    bool counting = false;
    vh::DevBuf counters;
    // patches, transfers and bounce state (vh_bounce.hip)
    struct BounceData *bounce = nullptr;
    // faces, samples and lights of the direct-light gather (vh_gather.hip)
    struct GatherData *gather = nullptr;
    // the warm-start network: weights, optimiser state, data set (vh_mlp.hip)
    struct MlpData *mlp = nullptr;
    // the luxel radial filter and the lump (vh_final.hip)
    struct FinalData *final_ = nullptr;
    // the leaf-ambient fan (vh_ambient.hip)
    struct AmbientData *ambient = nullptr;
};

namespace vh {
// Brackets one kernel launch with events when timing is on.
struct KernelTimer {
    vh_ctx *c;
    int kernel;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(vh_ctx *ctx, int k, hipStream_t st);
    void done();
};
}  // namespace vh

namespace vh {
// Kernel launchers (vh_trace.hip).
int launch_visibility(vh_ctx *c, const float *d_rays, int32_t rw, float shared_reach, int64_t n,
                      float tmin, int32_t skip_id, int32_t sky_passes, uint64_t *d_bits,
                      hipStream_t s);
int launch_closest(vh_ctx *c, const float *d_rays, int32_t rw, float shared_reach, int64_t n,
                   float tmin, int32_t skip_id, int32_t *d_surface, float *d_fraction,
                   hipStream_t s);
int launch_point_lights(vh_ctx *c, const float *d_points, int64_t npoints, const float *d_lights,
                        int32_t nlights, float tmin, uint64_t *d_bits, hipStream_t s);
// The light-buffer form of the point-light trace (vh_lightbuf.hip), over
// points already in spatial order; and its default cells per cube-face edge.
constexpr int kLightBufferCells = 64;
constexpr int kMaxLightBufferCells = 120;  // a face's counters fit one workgroup's 64 KiB of LDS
int lb_resolution(vh_ctx *c);
int launch_point_lights_lb(vh_ctx *c, const float *d_points, const int32_t *order, int64_t npoints,
                           const float *d_lights, int32_t nlights, float tmin, uint64_t *d_bits,
                           hipStream_t s);
// The no-allocation point-light path (vh_trace.hip): noalloc_init finds this
// device's copy of the fixed working memory (once per device, with the device
// current); the launch uses it and allocates nothing.
int noalloc_init(int device);
int launch_point_lights_noalloc(vh_ctx *c, const float *d_points, int64_t npoints,
                                const float *d_lights, int32_t nlights, float tmin,
                                uint64_t *d_bits, hipStream_t s);
// Frees the patch/transfer/bounce buffers (vh_bounce.hip).
void bounce_release(vh_ctx *c);
// Frees the gather's buffers (vh_gather.hip).
void gather_release(vh_ctx *c);
// The loaded gather's faces and results, for the patch-light hand-off.
struct GatherView {
    int32_t nfaces = 0, nsamples = 0;
    const int32_t *face_ptr = nullptr;  // device, nfaces + 1
    const int32_t *face_nc = nullptr;   // device, normals per face (1, or 4 when bumped)
    const std::vector<int32_t> *h_face_ptr = nullptr;  // host copies of the two
    const std::vector<int32_t> *h_normals = nullptr;
    const float *pos = nullptr;         // device, 3 per sample
    const float4 *ldr = nullptr;        // device, [style slot][normal][sample]; normal 0 flat
    const float4 *hdr = nullptr;
    bool have_results = false;
    int32_t nslots = 1;                  // style slots in ldr / hdr
    const uint8_t *face_styles = nullptr;  // device, 4 style numbers per face (255 empty)
    const std::vector<uint8_t> *h_face_styles = nullptr;
    uint64_t style_generation = 0;       // changes when the face styles may have
};
bool gather_view(vh_ctx *c, GatherView *v);
// The loaded patch system and bounce state, for the warm-start network (vh_bounce.hip).
struct BounceView {
    int32_t npatch = 0, nleaves = 0;
    const float4 *origin_area = nullptr;  // origin xyz, area
    const float4 *normal_pd = nullptr;    // normal xyz, plane distance
    const float4 *e0 = nullptr;           // the last vh_bounce's direct light, or null
    const float4 *direct = nullptr;       // the last vh_patch_lights result, or null
    const float4 *total = nullptr;        // the last vh_bounce's bounced light, or null
    const float4 *total_b = nullptr;      // its bump directions (3 per patch), or null
    const int32_t *leaves = nullptr;      // nleaves leaf patches, ascending
    const uint8_t *sky = nullptr;         // per patch
    const int32_t *cluster = nullptr;     // per patch
    uint64_t generation = 0;              // changes with every vh_patches_load
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // bounds of the patch origins
};
bool bounce_view(vh_ctx *c, BounceView *v);
// The device buffer (npatch float4) a predicted start is written into; the
// prediction counts as present only after bounce_pred_ready.
float4 *bounce_pred_buffer(vh_ctx *c);
void bounce_pred_ready(vh_ctx *c);
// Parents from their children, area weighted, deepest level first.
void bounce_pull_up(vh_ctx *c, float4 *v, hipStream_t s);
// Frees the warm-start network's buffers (vh_mlp.hip).
void mlp_release(vh_ctx *c);
// Frees the final-light buffers (vh_final.hip).
void final_release(vh_ctx *c);
// The last vh_final_light lump, on the device (vh_final.hip).
struct FinalLump {
    const void *lump = nullptr;
    int64_t bytes = 0;
};
bool final_lump(vh_ctx *c, FinalLump *v);
// Frees the leaf-ambient buffers (vh_ambient.hip).
void ambient_release(vh_ctx *c);
}  // namespace vh

namespace vh {
// Waits for the work enqueued on `s` so far, bounded by the context's
// deadline and cancel flag (VH_OK, VH_ETIMEOUT, VH_ECANCELED or VH_EDEVICE).
// On a timeout or cancel the work keeps running on the device; it writes
// only the context's own buffers, and later calls queue behind it.
int wait_stream(vh_ctx *c, hipStream_t s);
// VH_ECANCELED when a cancel is pending (checked at a call's start).
int check_cancel(vh_ctx *c);
}  // namespace vh

#define VH_WAIT(c, s)                          \
    do {                                       \
        int rc_w_ = vh::wait_stream((c), (s)); \
        if (rc_w_ != VH_OK) return rc_w_;      \
    } while (0)

#define VH_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            vh::set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
            return VH_EDEVICE;                                                           \
        }                                                                                \
    } while (0)
