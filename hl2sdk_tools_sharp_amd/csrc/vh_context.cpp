// libvradhip C ABI: context lifetime, errors, scene load and the trace entry
// points.  Kernels live in vh_trace.hip.
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>

#include "vh_internal.h"

namespace vh {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }
void clear_error() { g_error.clear(); }

namespace {

// The process-wide cache of device blocks, per device.  A request takes the
// smallest cached block that holds it without wasting more than half of it
// (plus 2 MiB); blocks beyond kPoolCap bytes of cache are freed instead.
// Nothing is freed at exit: the driver reclaims it, and the HIP runtime may
// already be gone when static destructors run.
struct Block {
    void *ptr;
    size_t bytes;
};
std::mutex g_pool_mu;
std::vector<std::vector<Block>> g_pool;
size_t g_pool_bytes = 0;
constexpr size_t kPoolCap = (size_t)48 << 30;
thread_local int g_release_idle = 0;

size_t alloc_size(size_t need)
{
    if (need < 256) return 256;
    const size_t g = need <= ((size_t)1 << 20) ? 4096 : ((size_t)2 << 20);
    return (need + g - 1) / g * g;
}

void *pool_take(int dev, size_t need, size_t &got)
{
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (dev < 0 || (size_t)dev >= g_pool.size()) return nullptr;
    std::vector<Block> &v = g_pool[dev];
    size_t best = v.size();
    for (size_t i = 0; i < v.size(); i++)
        if (v[i].bytes >= need && v[i].bytes <= need + need / 2 + ((size_t)2 << 20) &&
            (best == v.size() || v[i].bytes < v[best].bytes))
            best = i;
    if (best == v.size()) return nullptr;
    Block b = v[best];
    v[best] = v.back();
    v.pop_back();
    g_pool_bytes -= b.bytes;
    got = b.bytes;
    return b.ptr;
}

void pool_give(int dev, void *ptr, size_t bytes)
{
    if (!g_release_idle) (void)hipDeviceSynchronize();
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (dev >= 0 && g_pool_bytes + bytes <= kPoolCap) {
            if ((size_t)dev >= g_pool.size()) g_pool.resize((size_t)dev + 1);
            g_pool[dev].push_back(Block{ptr, bytes});
            g_pool_bytes += bytes;
            return;
        }
    }
    (void)hipFree(ptr);
}

// Frees every cached block of `dev` (before retrying a failed allocation).
void pool_flush(int dev)
{
    std::vector<Block> v;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (dev < 0 || (size_t)dev >= g_pool.size()) return;
        v.swap(g_pool[dev]);
        for (const Block &b : v) g_pool_bytes -= b.bytes;
    }
    for (const Block &b : v) (void)hipFree(b.ptr);
}

}  // namespace

ReleaseIdle::ReleaseIdle() { g_release_idle++; }
ReleaseIdle::~ReleaseIdle() { g_release_idle--; }

int DevBuf::ensure(size_t need, const char *what)
{
    if (need <= bytes && ptr) return VH_OK;
    if (ptr) release();
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) d = 0;
    const size_t want = alloc_size(need);
    size_t got = 0;
    void *p = pool_take(d, want, got);
    if (!p) {
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            pool_flush(d);
            e = hipMalloc(&p, want);
        }
        if (e != hipSuccess) {
            ptr = nullptr;
            set_error(std::string("device allocation of ") + std::to_string(want) + " bytes for " +
                      what + " failed: " + hipGetErrorString(e));
            return VH_EDEVICE;
        }
        got = want;
    }
    ptr = p;
    bytes = got;
    dev = d;
    return VH_OK;
}

int check_cancel(vh_ctx *c)
{
    if (c->cancel.load(std::memory_order_acquire)) {
        set_error("canceled: vh_cancel was called on this context (vh_cancel_clear resets it)");
        return VH_ECANCELED;
    }
    return VH_OK;
}

int wait_stream(vh_ctx *c, hipStream_t s)
{
    hipEvent_t ev = nullptr;
    VH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) {
        hipEventDestroy(ev);
        set_error(std::string("hipEventRecord: ") + hipGetErrorString(e));
        return VH_EDEVICE;
    }
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    const int64_t limit = c->timeout_ms.load(std::memory_order_relaxed);
    int rc = VH_OK;
    for (int spin = 0;; spin++) {
        e = hipEventQuery(ev);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) {
            set_error(std::string("waiting on the device: ") + hipGetErrorString(e));
            rc = VH_EDEVICE;
            break;
        }
        if ((rc = check_cancel(c)) != VH_OK) break;
        const int64_t ms =
            std::chrono::duration_cast<std::chrono::milliseconds>(clock::now() - t0).count();
        if (limit >= 0 && ms >= limit) {
            set_error("timed out after " + std::to_string(ms) + " ms waiting on the device (limit " +
                      std::to_string(limit) + " ms, vh_set_timeout); the work goes on and later "
                      "calls on this context queue behind it");
            rc = VH_ETIMEOUT;
            break;
        }
        // poll briskly for short kernels, then back off
        if (spin > 2000) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    hipEventDestroy(ev);
    return rc;
}

void DevBuf::release()
{
    if (ptr) pool_give(dev, ptr, bytes);
    ptr = nullptr;
    bytes = 0;
    dev = -1;
}

KernelTimer::KernelTimer(vh_ctx *ctx, int k, hipStream_t st) : c(ctx), kernel(k), s(st)
{
    if (!c->timing) return;
    for (hipEvent_t *e : {&a, &b}) {
        if (!c->event_pool.empty()) {
            *e = c->event_pool.back();
            c->event_pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            *e = nullptr;
        }
    }
    if (a) hipEventRecord(a, s);
}

void KernelTimer::done()
{
    if (!c->timing || !a || !b) return;
    hipEventRecord(b, s);
    c->pending[kernel].emplace_back(a, b);
}

namespace {

// Ray slab for the host-pointer entry points: rays are copied, traced and
// read back this many at a time (a multiple of 64 so a slab owns whole
// words, SourceSharp.MapTools/Rad/Bounce/VisMatrix.cs:105-121).
constexpr int64_t kSlabRays = (int64_t)1 << 22;

struct Guard {
    vh_ctx *c;
    std::lock_guard<std::mutex> lk;
    explicit Guard(vh_ctx *ctx) : c(ctx), lk(ctx->mu) {}
};

int check_rays_args(const vh_ctx *c, int32_t rw, int64_t n, float tmin)
{
    if (!c->have_scene) {
        set_error("no scene loaded: call vh_scene_load first");
        return VH_ESTATE;
    }
    if (rw != 6 && rw != 7) {
        set_error("record_words must be 6 or 7, got " + std::to_string(rw));
        return VH_EINVAL;
    }
    if (n < 0) {
        set_error("negative ray count");
        return VH_EINVAL;
    }
    if (std::isnan(tmin)) {
        set_error("a NaN minimum distance cannot be traced; trace this batch on the CPU tracer");
        return VH_ENOTSUP;
    }
    return VH_OK;
}

hipStream_t stream_of(vh_ctx *c, void *s) { return s ? (hipStream_t)s : c->stream; }

int upload_scene(vh_ctx *c, const HostScene &hs)
{
    int rc;
    if ((rc = c->nodes.ensure(hs.nodes.size() * 4, "BVH nodes"))) return rc;
    if ((rc = c->tris.ensure(hs.tris.size() * 4, "triangles"))) return rc;
    VH_HIP(hipMemcpyAsync(c->nodes.ptr, hs.nodes.data(), hs.nodes.size() * 4,
                          hipMemcpyHostToDevice, c->stream));
    VH_HIP(hipMemcpyAsync(c->tris.ptr, hs.tris.data(), hs.tris.size() * 4, hipMemcpyHostToDevice,
                          c->stream));
    VH_HIP(hipStreamSynchronize(c->stream));
    c->nnodes = hs.nnodes;
    c->ntri = hs.ntri;
    c->depth = hs.depth;
    for (int k = 0; k < 3; k++) {
        c->lo[k] = hs.lo[k];
        c->hi[k] = hs.hi[k];
    }
    c->have_scene = true;
    c->have_verts = false;
    c->scene_gen++;
    return VH_OK;
}

int trace_visibility_host(vh_ctx *c, const float *rays, int32_t rw, float shared_reach, int64_t n,
                          float tmin, int32_t skip_id, int32_t sky, uint64_t *bits)
{
    int rc;
    int64_t slab = n < kSlabRays ? n : kSlabRays;
    if ((rc = c->rays_buf.ensure((size_t)slab * rw * 4, "ray slab"))) return rc;
    if ((rc = c->out_buf.ensure((size_t)((slab + 63) / 64) * 8, "hit bits"))) return rc;
    for (int64_t s0 = 0; s0 < n; s0 += slab) {
        int64_t m = n - s0 < slab ? n - s0 : slab;
        VH_HIP(hipMemcpyAsync(c->rays_buf.ptr, rays + s0 * rw, (size_t)m * rw * 4,
                              hipMemcpyHostToDevice, c->stream));
        if ((rc = launch_visibility(c, (const float *)c->rays_buf.ptr, rw, shared_reach, m, tmin,
                                    skip_id, sky, (uint64_t *)c->out_buf.ptr, c->stream)))
            return rc;
        // the slab's bits reach the caller only once its kernel is done (a
        // timeout or cancel between slabs leaves the caller's memory alone)
        VH_WAIT(c, c->stream);
        VH_HIP(hipMemcpy(bits + s0 / 64, c->out_buf.ptr, (size_t)((m + 63) / 64) * 8,
                         hipMemcpyDeviceToHost));
    }
    return VH_OK;
}

int trace_closest_host(vh_ctx *c, const float *rays, int32_t rw, float shared_reach, int64_t n,
                       float tmin, int32_t skip_id, int32_t *surface, float *fraction)
{
    int rc;
    int64_t slab = n < kSlabRays ? n : kSlabRays;
    if ((rc = c->rays_buf.ensure((size_t)slab * rw * 4, "ray slab"))) return rc;
    if ((rc = c->out_buf.ensure((size_t)slab * 4, "hit surfaces"))) return rc;
    if ((rc = c->out2_buf.ensure((size_t)slab * 4, "hit fractions"))) return rc;
    for (int64_t s0 = 0; s0 < n; s0 += slab) {
        int64_t m = n - s0 < slab ? n - s0 : slab;
        VH_HIP(hipMemcpyAsync(c->rays_buf.ptr, rays + s0 * rw, (size_t)m * rw * 4,
                              hipMemcpyHostToDevice, c->stream));
        if ((rc = launch_closest(c, (const float *)c->rays_buf.ptr, rw, shared_reach, m, tmin,
                                 skip_id, (int32_t *)c->out_buf.ptr, (float *)c->out2_buf.ptr,
                                 c->stream)))
            return rc;
        VH_WAIT(c, c->stream);
        VH_HIP(hipMemcpy(surface + s0, c->out_buf.ptr, (size_t)m * 4, hipMemcpyDeviceToHost));
        VH_HIP(hipMemcpy(fraction + s0, c->out2_buf.ptr, (size_t)m * 4, hipMemcpyDeviceToHost));
    }
    return VH_OK;
}

// Known-answer self-test at open, the counterpart of VulkanDevice's
// two-triangle test (SourceSharp.MapTools.Gpu/Interop/VulkanDevice.cs:2100): a
// unit square of two triangles at z = 0 and four rays whose answers are fixed.
int self_test(vh_ctx *c)
{
    const float v9[18] = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 0, 0, 1, 0};
    const int32_t ids[2] = {7, 8};
    HostScene hs;
    int rc = build_scene(v9, ids, nullptr, 2, hs);
    if (rc) return rc;
    if ((rc = upload_scene(c, hs))) return rc;
    const float rays[4 * 7] = {
        0.75f, 0.25f, 1.0f, 0.0f, 0.0f, -2.0f, 1.0f,  // hits at t = 0.5 (triangle 7)
        2.0f,  2.0f,  1.0f, 0.0f, 0.0f, -2.0f, 1.0f,  // misses the square
        0.5f,  0.5f,  1.0f, 0.0f, 0.0f, -0.5f, 1.0f,  // ends above the square
        0.25f, 0.75f, 1.0f, 0.0f, 0.0f, -1.0f, 4.0f,  // hits triangle 8 at t = 1, fraction 0.25
    };
    uint64_t bits = ~0ull;
    int32_t surf[4];
    float frac[4];
    if ((rc = trace_visibility_host(c, rays, 7, 1.0f, 4, 0.0f, -1, 0, &bits))) return rc;
    if ((rc = trace_closest_host(c, rays, 7, 1.0f, 4, 0.0f, -1, surf, frac))) return rc;
    c->have_scene = false;
    bool ok = bits == 0x9ull && surf[0] == 7 && frac[0] == 0.5f && surf[1] == -1 &&
              frac[1] == 1.0f && surf[2] == -1 && surf[3] == 8 && frac[3] == 0.25f;
    if (!ok) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "self-test failed: bits=%llx surf=(%d,%d,%d,%d) frac=(%g,%g,%g,%g)",
                 (unsigned long long)bits, surf[0], surf[1], surf[2], surf[3], frac[0], frac[1],
                 frac[2], frac[3]);
        set_error(buf);
        return VH_DECLINED;
    }
    return VH_OK;
}

void destroy(vh_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    ReleaseIdle idle;  // the stream is drained: the buffers go back to the cache as they are
    bounce_release(c);
    gather_release(c);
    mlp_release(c);
    final_release(c);
    ambient_release(c);
    for (DevBuf *b : {&c->nodes, &c->tris, &c->rays_buf, &c->out_buf, &c->out2_buf, &c->pts_buf,
                      &c->lights_buf, &c->keys_a, &c->keys_b, &c->perm_a, &c->perm_b,
                      &c->sort_tmp, &c->counters, &c->lb_boxes, &c->lb_tverts, &c->lb_off, &c->lb_cur, &c->lb_ent, &c->lb_rec, &c->verts,
                      &c->sb_cam, &c->sb_nodes, &c->sb_planes, &c->sb_recurse})
        b->release();
    if (c->lb_pinned) hipHostFree(c->lb_pinned);
    for (auto &v : c->pending)
        for (auto &pr : v) {
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : c->event_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

const char *vh_last_error(void) { return g_error.c_str(); }

int vh_open(const vh_open_opts *opts, vh_ctx **out)
{
    clear_error();
    if (!out) {
        set_error("out is NULL");
        return VH_EINVAL;
    }
    *out = nullptr;
    vh_open_opts o{};
    o.device = -1;
    o.self_test = 1;
    if (opts) o = *opts;
    if (o.fail_stage == 1) {
        set_error("declined at stage 1 (device query): fault hook");
        return VH_DECLINED;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        set_error(std::string("no HIP device: ") +
                  (e != hipSuccess ? hipGetErrorString(e) : "device count is 0"));
        return VH_DECLINED;
    }
    int dev = o.device;
    if (dev < 0) {
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    }
    if (dev >= count) {
        set_error("device " + std::to_string(dev) + " does not exist (" + std::to_string(count) +
                  " present)");
        return VH_DECLINED;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        set_error("cannot read device properties");
        return VH_DECLINED;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("device is ") + prop.gcnArchName +
                  "; this library is built for gfx950 only");
        return VH_DECLINED;
    }
    vh_ctx *c = new vh_ctx();
    c->device = dev;
    if (hipSetDevice(dev) != hipSuccess) {
        set_error("hipSetDevice failed");
        delete c;
        return VH_DECLINED;
    }
    if (o.fail_stage == 2) {
        set_error("declined at stage 2 (stream): fault hook");
        destroy(c);
        return VH_DECLINED;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("stream creation failed");
        c->stream = nullptr;
        destroy(c);
        return VH_DECLINED;
    }
    c->cu_count = prop.multiProcessorCount;
    // the no-allocation entry point's working memory on this device; a
    // failure here only disables that entry point (it then reports it)
    if (noalloc_init(dev) != VH_OK) clear_error();
    c->identity = std::string("hip-bvh2-exact-1:") + prop.gcnArchName + ":" +
                  std::to_string(prop.multiProcessorCount) + "cu";
    if (o.fail_stage == 3) {
        set_error("declined at stage 3 (self-test): fault hook");
        destroy(c);
        return VH_DECLINED;
    }
    if (o.self_test) {
        int rc = self_test(c);
        if (rc) {
            std::string why = g_error;
            destroy(c);
            set_error(why);
            return VH_DECLINED;
        }
    }
    *out = c;
    return VH_OK;
}

void vh_close(vh_ctx *c) { destroy(c); }

const char *vh_tracer_identity(const vh_ctx *c) { return c ? c->identity.c_str() : ""; }

int vh_supports(const vh_ctx *c, float min_distance, int32_t skip_id, int32_t sky_passes)
{
    (void)c;
    (void)skip_id;
    (void)sky_passes;
    return std::isnan(min_distance) ? 0 : 1;
}

int vh_scene_load(vh_ctx *c, const float *v9, const int32_t *ids, const uint8_t *flags, int32_t n)
{
    return vh_scene_load_with(c, v9, ids, flags, n, VH_BUILD_AUTO);
}

int vh_scene_load_with(vh_ctx *c, const float *v9, const int32_t *ids, const uint8_t *flags, int32_t n,
                       int32_t builder)
{
    clear_error();
    if (!c || !v9) {
        set_error("NULL context or vertices");
        return VH_EINVAL;
    }
    if (builder < VH_BUILD_AUTO || builder > VH_BUILD_DEVICE) {
        set_error("bad builder");
        return VH_EINVAL;
    }
    const bool device = builder == VH_BUILD_DEVICE || (builder == VH_BUILD_AUTO && n >= VH_BUILD_DEVICE_MIN);
    Guard g(c);
    VH_HIP(hipSetDevice(c->device));
    if (!device) {
        HostScene hs;
        int rc = build_scene(v9, ids, flags, n, hs);
        if (rc) return rc;
        if ((rc = upload_scene(c, hs))) return rc;
        if ((rc = c->verts.ensure((size_t)n * 36, "scene vertices"))) return rc;
        VH_HIP(hipMemcpyAsync(c->verts.ptr, v9, (size_t)n * 36, hipMemcpyHostToDevice, c->stream));
        VH_HIP(hipStreamSynchronize(c->stream));
        c->have_verts = true;
        return VH_OK;
    }
    if (n < 16) {
        set_error("the device builder takes at least 16 triangles");
        return VH_EINVAL;
    }
    if (n >= kMaxOrigIndex) {
        set_error("too many triangles for one scene (limit 2^27)");
        return VH_EINVAL;
    }
    (void)flags;  // transparent triangles are opaque here, as in KdRayTracer.TestLines
    c->have_scene = false;
    c->have_verts = false;
    return build_scene_device(c, v9, ids, n);
}

int vh_scene_read(vh_ctx *c, float *nodes, int64_t node_floats, float *tris, int64_t tri_floats)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    if (!c->have_scene) {
        set_error("no scene loaded");
        return VH_ESTATE;
    }
    if ((nodes && node_floats < (int64_t)c->nnodes * 16) || (tris && tri_floats < (int64_t)c->ntri * 12)) {
        set_error("the output arrays are too small for the scene");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    VH_WAIT(c, c->stream);
    if (nodes && c->nnodes > 0)
        VH_HIP(hipMemcpy(nodes, c->nodes.ptr, (size_t)c->nnodes * 64, hipMemcpyDeviceToHost));
    if (tris && c->ntri > 0) VH_HIP(hipMemcpy(tris, c->tris.ptr, (size_t)c->ntri * 48, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_scene_info(const vh_ctx *c, int32_t *triangles, int32_t *nodes, int32_t *depth)
{
    if (!c) return VH_EINVAL;
    if (triangles) *triangles = c->ntri;
    if (nodes) *nodes = c->nnodes;
    if (depth) *depth = c->depth;
    return c->have_scene ? VH_OK : VH_ESTATE;
}

int vh_trace_visibility(vh_ctx *c, const float *rays, int32_t rw, float shared_reach, int64_t n,
                        float tmin, int32_t skip_id, int32_t sky, uint64_t *bits, int64_t bit_words)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    int rc = check_rays_args(c, rw, n, tmin);
    if (rc) return rc;
    int64_t words = (n + 63) / 64;
    if (bit_words < words) {
        set_error(std::to_string(n) + " rays need " + std::to_string(words) +
                  " words of hit bits, and " + std::to_string(bit_words) + " were given");
        return VH_EINVAL;
    }
    if (n == 0) return VH_OK;
    if (!rays || !bits) {
        set_error("NULL rays or bits");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    rc = trace_visibility_host(c, rays, rw, shared_reach, n, tmin, skip_id, sky, bits);
    if (!rc) {
        c->stats.rays_visibility += n;
        c->stats.calls++;
    }
    return rc;
}

int vh_trace_closest(vh_ctx *c, const float *rays, int32_t rw, float shared_reach, int64_t n,
                     float tmin, int32_t skip_id, int32_t *surface, float *fraction,
                     int64_t hit_slots)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    int rc = check_rays_args(c, rw, n, tmin);
    if (rc) return rc;
    if (hit_slots < n) {
        set_error(std::to_string(n) + " rays need " + std::to_string(n) + " hit slots, and " +
                  std::to_string(hit_slots) + " were given");
        return VH_EINVAL;
    }
    if (n == 0) return VH_OK;
    if (!rays || !surface || !fraction) {
        set_error("NULL rays or outputs");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    rc = trace_closest_host(c, rays, rw, shared_reach, n, tmin, skip_id, surface, fraction);
    if (!rc) {
        c->stats.rays_closest += n;
        c->stats.calls++;
    }
    return rc;
}

int vh_trace_point_lights(vh_ctx *c, const float *points, int64_t npoints, const float *lights,
                          int32_t nlights, float tmin, uint64_t *bits, int64_t bit_words)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    int rc = check_rays_args(c, 7, npoints, tmin);
    if (rc) return rc;
    if (nlights < 0) {
        set_error("negative light count");
        return VH_EINVAL;
    }
    int64_t words = (npoints * (int64_t)nlights + 63) / 64;
    if (bit_words < words) {
        set_error(std::to_string(npoints * (int64_t)nlights) + " rays need " +
                  std::to_string(words) + " words of hit bits, and " + std::to_string(bit_words) +
                  " were given");
        return VH_EINVAL;
    }
    if (words == 0) return VH_OK;
    if (!points || !lights || !bits) {
        set_error("NULL points, lights or bits");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    if ((rc = c->pts_buf.ensure((size_t)npoints * 12, "points")) ||
        (rc = c->lights_buf.ensure((size_t)nlights * 12, "lights")) ||
        (rc = c->out_buf.ensure((size_t)words * 8, "hit bits")))
        return rc;
    VH_HIP(hipMemcpyAsync(c->pts_buf.ptr, points, (size_t)npoints * 12, hipMemcpyHostToDevice,
                          c->stream));
    VH_HIP(hipMemcpyAsync(c->lights_buf.ptr, lights, (size_t)nlights * 12, hipMemcpyHostToDevice,
                          c->stream));
    if ((rc = launch_point_lights(c, (const float *)c->pts_buf.ptr, npoints,
                                  (const float *)c->lights_buf.ptr, nlights, tmin,
                                  (uint64_t *)c->out_buf.ptr, c->stream)))
        return rc;
    VH_WAIT(c, c->stream);
    VH_HIP(hipMemcpy(bits, c->out_buf.ptr, (size_t)words * 8, hipMemcpyDeviceToHost));
    c->stats.rays_visibility += npoints * (int64_t)nlights;
    c->stats.calls++;
    return VH_OK;
}

int vh_set_point_light_method(vh_ctx *c, int32_t method, int32_t cells)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    if (method < VH_POINT_LIGHTS_AUTO || method > VH_POINT_LIGHTS_LIGHT_BUFFER) {
        set_error("unknown point-light method " + std::to_string(method));
        return VH_EINVAL;
    }
    if (cells != 0 && (cells < 4 || cells > kMaxLightBufferCells)) {
        set_error("light-buffer cells per face edge must be 4.." + std::to_string(kMaxLightBufferCells) +
                  " (or 0 for the default), not " + std::to_string(cells));
        return VH_EINVAL;
    }
    Guard g(c);
    c->pl_method = method;
    c->lb_cells = cells;
    return VH_OK;
}

int vh_trace_visibility_dev(vh_ctx *c, const float *d_rays, int32_t rw, float shared_reach,
                            int64_t n, float tmin, int32_t skip_id, int32_t sky, uint64_t *d_bits,
                            void *stream)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    int rc = check_rays_args(c, rw, n, tmin);
    if (rc) return rc;
    if (n == 0) return VH_OK;
    VH_HIP(hipSetDevice(c->device));
    rc = launch_visibility(c, d_rays, rw, shared_reach, n, tmin, skip_id, sky, d_bits,
                           stream_of(c, stream));
    if (!rc) {
        c->stats.rays_visibility += n;
        c->stats.calls++;
    }
    return rc;
}

int vh_trace_closest_dev(vh_ctx *c, const float *d_rays, int32_t rw, float shared_reach, int64_t n,
                         float tmin, int32_t skip_id, int32_t *d_surface, float *d_fraction,
                         void *stream)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    int rc = check_rays_args(c, rw, n, tmin);
    if (rc) return rc;
    if (n == 0) return VH_OK;
    VH_HIP(hipSetDevice(c->device));
    rc = launch_closest(c, d_rays, rw, shared_reach, n, tmin, skip_id, d_surface, d_fraction,
                        stream_of(c, stream));
    if (!rc) {
        c->stats.rays_closest += n;
        c->stats.calls++;
    }
    return rc;
}

int vh_trace_point_lights_dev(vh_ctx *c, const float *d_points, int64_t npoints,
                              const float *d_lights, int32_t nlights, float tmin, uint64_t *d_bits,
                              void *stream)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    int rc = check_rays_args(c, 7, npoints, tmin);
    if (rc) return rc;
    if (nlights < 0) {
        set_error("negative light count");
        return VH_EINVAL;
    }
    if (npoints == 0 || nlights == 0) return VH_OK;
    VH_HIP(hipSetDevice(c->device));
    rc = launch_point_lights(c, d_points, npoints, d_lights, nlights, tmin, d_bits,
                             stream_of(c, stream));
    if (!rc) {
        c->stats.rays_visibility += npoints * (int64_t)nlights;
        c->stats.calls++;
    }
    return rc;
}

int vh_trace_point_lights_noalloc(vh_ctx *c, const float *d_points, int64_t npoints,
                                  const float *d_lights, int32_t nlights, float tmin,
                                  uint64_t *d_bits, void *stream)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    int rc = check_rays_args(c, 7, npoints, tmin);
    if (rc) return rc;
    if (nlights < 0) {
        set_error("negative light count");
        return VH_EINVAL;
    }
    if (npoints > 0 && nlights > 0 && (!d_points || !d_lights || !d_bits)) {
        set_error("NULL points, lights or bits");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    rc = launch_point_lights_noalloc(c, d_points, npoints, d_lights, nlights, tmin, d_bits,
                                     stream_of(c, stream));
    if (!rc) {
        c->stats.rays_visibility += npoints * (int64_t)nlights;
        c->stats.calls++;
    }
    return rc;
}

int vh_sync(vh_ctx *c, void *stream)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    VH_HIP(hipSetDevice(c->device));
    return wait_stream(c, stream_of(c, stream));
}

int vh_set_timeout(vh_ctx *c, int64_t ms)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    c->timeout_ms.store(ms < 0 ? -1 : ms, std::memory_order_relaxed);
    return VH_OK;
}

int vh_cancel(vh_ctx *c)
{
    if (!c) return VH_EINVAL;
    c->cancel.store(1, std::memory_order_release);
    return VH_OK;
}

int vh_cancel_clear(vh_ctx *c)
{
    if (!c) return VH_EINVAL;
    c->cancel.store(0, std::memory_order_release);
    return VH_OK;
}

int vh_timing_enable(vh_ctx *c, int32_t on)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    VH_HIP(hipSetDevice(c->device));
    for (auto &v : c->pending) {
        for (auto &pr : v) {
            hipEventSynchronize(pr.second);
            c->event_pool.push_back(pr.first);
            c->event_pool.push_back(pr.second);
        }
        v.clear();
    }
    c->timing = on != 0;
    return VH_OK;
}

int vh_timing_read(vh_ctx *c, int32_t kernel, double *total_ms, int64_t *launches)
{
    clear_error();
    if (!c || kernel < 0 || kernel >= VH_KERNEL_COUNT) {
        set_error("NULL context or bad kernel id");
        return VH_EINVAL;
    }
    Guard g(c);
    VH_HIP(hipSetDevice(c->device));
    double ms = 0;
    int64_t n = 0;
    for (auto &pr : c->pending[kernel]) {
        VH_HIP(hipEventSynchronize(pr.second));
        float e = 0;
        VH_HIP(hipEventElapsedTime(&e, pr.first, pr.second));
        ms += e;
        n++;
        c->event_pool.push_back(pr.first);
        c->event_pool.push_back(pr.second);
    }
    c->pending[kernel].clear();
    if (total_ms) *total_ms = ms;
    if (launches) *launches = n;
    return VH_OK;
}

int vh_count_enable(vh_ctx *c, int32_t on)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    VH_HIP(hipSetDevice(c->device));
    int rc = c->counters.ensure(64, "work counters");
    if (rc) return rc;
    VH_HIP(hipMemsetAsync(c->counters.ptr, 0, 32, c->stream));
    VH_HIP(hipStreamSynchronize(c->stream));
    c->counting = on != 0;
    return VH_OK;
}

int vh_count_read(vh_ctx *c, int64_t *node_visits, int64_t *triangle_tests)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    VH_HIP(hipSetDevice(c->device));
    int64_t v[2] = {0, 0};
    if (c->counters.ptr) {
        VH_HIP(hipDeviceSynchronize());
        VH_HIP(hipMemcpy(v, c->counters.ptr, 16, hipMemcpyDeviceToHost));
    }
    if (node_visits) *node_visits = v[0];
    if (triangle_tests) *triangle_tests = v[1];
    return VH_OK;
}

int vh_count_read_slots(vh_ctx *c, int64_t *node_slots, int64_t *triangle_slots)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    Guard g(c);
    VH_HIP(hipSetDevice(c->device));
    int64_t v[4] = {0, 0, 0, 0};
    if (c->counters.ptr) {
        VH_HIP(hipDeviceSynchronize());
        VH_HIP(hipMemcpy(v, c->counters.ptr, 32, hipMemcpyDeviceToHost));
    }
    if (node_slots) *node_slots = v[2];
    if (triangle_slots) *triangle_slots = v[3];
    return VH_OK;
}

int vh_get_stats(const vh_ctx *c, vh_stats *out)
{
    if (!c || !out) return VH_EINVAL;
    *out = c->stats;
    return VH_OK;
}

}  // extern "C"
