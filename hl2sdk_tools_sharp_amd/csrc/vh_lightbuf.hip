// Light buffers for point-light shadow rays (Haines & Greenberg, "The Light
// Buffer: A Shadow-Testing Accelerator", IEEE CG&A 6(9), 1986), built on the
// GPU per call and exact: a candidate filter in front of the reference's
// triangle test (Tracing/KdRayTracer.cs:1194-1267, tri_test in vh_traverse.h).
//
// For each light, a cube of 6 x N x N direction cells centred on the light.
// Cell c lists every triangle whose padded box has a point whose direction
// from the light falls in c, with the box's smallest depth along the cell
// face's axis, sorted by that depth.  A shadow ray from point p to light L
// (direction d = fl(L - p), reach 1) can only be blocked by a triangle whose
// box holds a point of the segment p + t d, 0 < t < 1; seen from L every such
// point lies in the direction of -d, at a depth below |d_a| (a the cell face's
// axis).  So the ray tests the triangles of the cell of -d, in depth order,
// until one blocks it or the next one starts at or beyond |d_a|: the exact
// test over a superset of the triangles that can block, which is the answer
// of the BVH walk bit for bit (both are the exact test's "any triangle blocks
// strictly short of the end").
//
// Conservativeness.  The boxes are the BVH's leaf boxes, already widened by
// the scene builder's pad (a hit the exact test accepts lies in its leaf's
// widened box).  The segment ends at p + d, not at L: the rounding of
// d = fl(L - p) moves that end by at most 2^-24 (|L| + |p|) per axis, so every
// box is widened by a further delta = 2^-21 (|L| + max|p| + scene extent):
// then the widened box holds a point in the exact direction -d from L at depth
// (1 - t) |d_a| whenever the segment meets the box.  Cell ranges are grown by
// 2^-16 in face coordinates and depths shrunk by 2^-20 against the build's own
// float rounding.
#include <hipcub/hipcub.hpp>

#include "vh_lightbuf.h"
#include "vh_traverse.h"

#pragma clang fp contract(off)

namespace vh {
namespace {

constexpr int kBlock = 64;

// Each triangle's box (leaf order): the widened box of the leaf that holds it,
// from the leaf's parent node.  boxes[2k] = (lo, 0), boxes[2k + 1] = (hi, 0).
__global__ void k_lb_boxes(const float4 *__restrict__ nodes, int32_t nnodes,
                           const float4 *__restrict__ tris, float4 *__restrict__ boxes)
{
    int32_t g = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= nnodes) return;
    const float4 *nd = nodes + (size_t)g * 4;
    const float4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3];
    for (int c = 0; c < 2; c++) {
        const int32_t ref = __float_as_int(c == 0 ? q3.x : q3.y);
        if (ref >= 0 || ref == kEmptyRef) continue;
        const float4 lo = c == 0 ? make_float4(q0.x, q0.z, q2.x, 0.0f) : make_float4(q1.x, q1.z, q2.z, 0.0f);
        const float4 hi = c == 0 ? make_float4(q0.y, q0.w, q2.y, 0.0f) : make_float4(q1.y, q1.w, q2.w, 0.0f);
        for (int32_t k = ~ref;; k++) {
            boxes[(size_t)k * 2] = lo;
            boxes[(size_t)k * 2 + 1] = hi;
            if (__float_as_uint(tris[(size_t)k * 3 + 2].w) & kTriLast) break;
        }
    }
}

// Each triangle's vertices in leaf order (3 float4 a triangle, w unused),
// gathered from the caller's order so the build's loads are coalesced.
__global__ void k_lb_tverts(const float4 *__restrict__ tris, int32_t ntri, const float *__restrict__ verts,
                            float4 *__restrict__ tv)
{
    const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= ntri) return;
    const float *v = verts + (size_t)(__float_as_uint(tris[(size_t)k * 3 + 2].w) >> kOrigShift) * 9;
    for (int q = 0; q < 3; q++) tv[(size_t)k * 3 + q] = make_float4(v[q * 3 + 0], v[q * 3 + 1], v[q * 3 + 2], 0.0f);
}

// The largest finite |coordinate| of the points, as float bits (non-negative
// floats order as their bits).  *out must start at 0.
__global__ void k_lb_pmax(const float *__restrict__ pts, int64_t n, uint32_t *__restrict__ out)
{
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < 3; c++) {
            const float v = fabsf(pts[i * 3 + c]);
            if (v <= 3.4028235e38f) m = fmaxf(m, v);  // NaN and inf rays are never traced
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// A (light, triangle) pair's box relative to the light, widened by delta
// (see the top of this file).  False for a non-finite light or an empty box.
__device__ __forceinline__ bool pair_box(const float4 *__restrict__ boxes, const float *__restrict__ lights,
                                         int32_t l, int32_t k, float pmax, float scene_mag, float lo[3],
                                         float hi[3])
{
    const float Lx = lights[l * 3 + 0], Ly = lights[l * 3 + 1], Lz = lights[l * 3 + 2];
    if (!(fabsf(Lx) <= 3.4028235e38f && fabsf(Ly) <= 3.4028235e38f && fabsf(Lz) <= 3.4028235e38f))
        return false;  // a ray to a non-finite light is never traced
    const float delta = lb_delta(Lx, Ly, Lz, pmax, scene_mag);
    const float4 blo = boxes[(size_t)k * 2], bhi = boxes[(size_t)k * 2 + 1];
    if (!(blo.x <= bhi.x)) return false;  // an empty leaf box holds no triangle
    lo[0] = blo.x - Lx - delta;
    lo[1] = blo.y - Ly - delta;
    lo[2] = blo.z - Lz - delta;
    hi[0] = bhi.x - Lx + delta;
    hi[1] = bhi.y - Ly + delta;
    hi[2] = bhi.z - Lz + delta;
    return true;
}

struct Lists {
    int2 *key;       // (depth bits, triangle) per entry, in fill order
    uint32_t *cell;  // the entry's cell
};

constexpr int kImageThreads = 256;
constexpr int kImageParts = 4;    // workgroups per (light, face) image, each over every fourth triangle
constexpr int kInlineCells = 32;  // rectangles with more cells go to the workgroup's shared list
constexpr int kBigCap = 256;

// kImageParts workgroups per (light, cube face) image, each over its share of
// the triangles (every kImageParts-th) with the image's N x N cell counters in LDS: FILL false counts
// each cell's triangles and stores the counts as count[(image cell) * parts +
// part]; FILL true starts each cell's cursor at that slot's offset (after an
// exclusive scan) and writes the entries, so a cell's list is its parts' runs
// back to back, [off[cell * parts], off[(cell + 1) * parts]).  A triangle seen
// through more than kInlineCells cells of the face (a box near its light) is
// queued and its cells shared out among the workgroup's threads afterwards
// (the few past the queue's capacity are done by their own thread); either
// way each (triangle, cell) is put once, so the two passes agree.  LDS atomics
// only: no two workgroups share a slot.
// The cells of triangle k (leaf order) on face (a, side) of light l
// (lb_cells_*): its widened leaf box, cut to its widened projection when the
// scene's vertices are kept (tv, leaf order; loaded only for a rectangle worth
// cutting).  Returns false when none.
__device__ __forceinline__ bool pair_cells(const float4 *__restrict__ boxes, const float4 *__restrict__ tv,
                                           const float *__restrict__ lights, int32_t l, int32_t k, int a, int side,
                                           int N, float pmax, float scene_mag, float pad, LbCells &pc)
{
    float lo[3], hi[3];
    if (!pair_box(boxes, lights, l, k, pmax, scene_mag, lo, hi) || !lb_cells_box(lo, hi, a, side, N, pc))
        return false;
    if (!tv || !lb_cells_cut(pc)) return true;
    const float Lx = lights[l * 3 + 0], Ly = lights[l * 3 + 1], Lz = lights[l * 3 + 2];
    float v[9];
    for (int q = 0; q < 3; q++) {
        const float4 p = tv[(size_t)k * 3 + q];
        v[q * 3 + 0] = p.x - Lx;
        v[q * 3 + 1] = p.y - Ly;
        v[q * 3 + 2] = p.z - Lz;
    }
    const float r = (pad + 2.0f * lb_delta(Lx, Ly, Lz, pmax, scene_mag)) * 1.0009765625f;
    return lb_cells_raster(v, a, side, N, r, pc);
}

template <bool FILL>
__global__ __launch_bounds__(kImageThreads) void k_lb_image(
    const float4 *__restrict__ boxes, const float4 *__restrict__ tv, int32_t ntri,
    const float *__restrict__ lights, int32_t N, const uint32_t *__restrict__ pmax_bits,
    float scene_mag, float pad, uint32_t *__restrict__ count, const uint32_t *__restrict__ off, Lists ls)
{
    extern __shared__ uint32_t sh[];
    __shared__ uint32_t nbig;
    const int NN = N * N;
    uint32_t *cnt = sh;
    int32_t *bigk = (int32_t *)(sh + NN);
    const int img = blockIdx.x / kImageParts, part = blockIdx.x - img * kImageParts;
    const int l = img / 6, f = img - 6 * (img / 6);
    const int a = f >> 1, side = f & 1;
    const size_t base = (size_t)img * NN;
    for (int i = threadIdx.x; i < NN; i += blockDim.x) cnt[i] = FILL ? off[(base + i) * kImageParts + part] : 0u;
    // every write stays below the image's end, whatever happens
    const uint32_t img_end = FILL ? off[(base + NN) * kImageParts] : 0u;
    if (threadIdx.x == 0) nbig = 0;
    __syncthreads();
    const float pmax = __uint_as_float(*pmax_bits);
    auto put = [&](int c, float dlo, int32_t k) {
        if (!FILL) {
            atomicAdd(&cnt[c], 1u);
        } else {
            const uint32_t pos = atomicAdd(&cnt[c], 1u);
            if (pos < img_end) {
                ls.key[pos] = make_int2(__float_as_int(dlo), k);
                ls.cell[pos] = (uint32_t)(base + c);
            }
        }
    };
    // the parts take the triangles in turn (k = part mod kImageParts): the
    // triangles are in leaf order, so a light's near ones lie together, and
    // dealt out this way they load the image's workgroups evenly
    for (int32_t k = part + kImageParts * (int32_t)threadIdx.x; k < ntri; k += kImageParts * (int32_t)blockDim.x) {
        LbCells pc;
        if (!pair_cells(boxes, tv, lights, l, k, a, side, N, pmax, scene_mag, pad, pc)) continue;
        if ((pc.iu1 - pc.iu0 + 1) * (pc.iw1 - pc.iw0 + 1) > kInlineCells) {
            const uint32_t q = atomicAdd(&nbig, 1u);
            if (q < (uint32_t)kBigCap) {
                bigk[q] = k;
                continue;
            }
        }
        for (int iu = pc.iu0; iu <= pc.iu1; iu++)
            for (int iw = pc.iw0; iw <= pc.iw1; iw++)
                if (lb_cells_has(pc, iu, iw, N)) put(iu * N + iw, pc.box.dlo, k);
    }
    __syncthreads();
    const uint32_t nb = nbig < (uint32_t)kBigCap ? nbig : (uint32_t)kBigCap;
    for (uint32_t q = 0; q < nb; q++) {
        const int32_t k = bigk[q];
        LbCells pc;
        pair_cells(boxes, tv, lights, l, k, a, side, N, pmax, scene_mag, pad, pc);
        const int wu = pc.iw1 - pc.iw0 + 1, cells = (pc.iu1 - pc.iu0 + 1) * wu;
        for (int t = threadIdx.x; t < cells; t += blockDim.x) {
            const int iu = pc.iu0 + t / wu, iw = pc.iw0 + t % wu;
            if (lb_cells_has(pc, iu, iw, N)) put(iu * N + iw, pc.box.dlo, k);
        }
    }
    if (!FILL) {
        __syncthreads();
        for (int i = threadIdx.x; i < NN; i += blockDim.x) count[(base + i) * kImageParts + part] = cnt[i];
    }
}

// The lists in depth order as triangle records: lane j finds its entry's
// rank in its cell by (depth, triangle) -- every key of a cell is distinct, a
// triangle being listed once per cell -- and writes the record at the cell's
// start plus that rank, in three arrays read at different steps of a ray's
// walk down its cell's list:
//   depth[i]                      the cut (every step)
//   plane[i]    = (n, d)          the plane screen (every step)
//   edge[2i]    = (E0, e2), edge[2i + 1] = (E1, e5)   the rest of the test
// E0 and E1 are the edge equations e0 e1 / e3 e4 of the reference's format
// (Tracing/KdTree.cs:121-167) spread over x, y, z, zero on the dropped axis,
// so that the test needs no per-lane axis selects (tri_test_record).
struct Records {
    uint32_t *depth;
    float4 *plane;
    float4 *edge;
};

__device__ __forceinline__ bool key_less(int2 x, int2 y)
{
    return (uint32_t)x.x < (uint32_t)y.x || ((uint32_t)x.x == (uint32_t)y.x && x.y < y.y);
}

__global__ void k_lb_records(const int2 *__restrict__ key, const uint32_t *__restrict__ kcell, uint32_t n,
                             const uint32_t *__restrict__ off, uint32_t ncell, const float4 *__restrict__ tris,
                             int32_t ntri, Records rec)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int2 e = key[j];
    const uint32_t c = kcell[j];
    // every entry was written by the fill when the count and the fill agree,
    // which they do by construction; anything else is left alone
    if (c >= ncell || (uint32_t)e.y >= (uint32_t)ntri) return;
    const uint32_t b = off[(size_t)c * kImageParts], end = off[(size_t)(c + 1) * kImageParts];
    uint32_t rank = 0;
#pragma unroll 4
    for (uint32_t i = b; i < end; i++) rank += key_less(key[i], e) ? 1u : 0u;
    const uint32_t at = b + rank;
    if (at >= end) return;
    const float4 *tp = tris + (size_t)e.y * 3;
    const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
    const uint32_t meta = __float_as_uint(t2.w);
    const int cs0 = meta & 3, cs1 = (meta >> 2) & 3;
    float E0[3] = {0.0f, 0.0f, 0.0f}, E1[3] = {0.0f, 0.0f, 0.0f};
    if (cs0 < 3 && cs1 < 3) {
        E0[cs0] = t1.x;
        E0[cs1] = t1.y;
        E1[cs0] = t1.w;
        E1[cs1] = t2.x;
    }
    rec.depth[at] = (uint32_t)e.x;
    rec.plane[at] = t0;
    rec.edge[(size_t)at * 2] = make_float4(E0[0], E0[1], E0[2], t1.z);
    rec.edge[(size_t)at * 2 + 1] = make_float4(E1[0], E1[1], E1[2], t2.y);
}

// The exact test (tri_test, KdRayTracer.IntersectTriangle) on a record, with
// the hit point on all three axes and the edge equations as 3-D vectors that
// are zero on the dropped axis: b0 = ((E0x hx + E0y hy) + E0z hz) + e2 adds
// the same two products as the reference's (e0 h[cs0] + e1 h[cs1]) + e2 and a
// signed zero, and IEEE addition is commutative, so b0 (and b1) are the
// reference's values up to the sign of a zero sum, which no comparison with
// 1e-10 can tell apart.  (h on the dropped axis is finite whenever t is in
// (0, 1); outside it the answer is false either way.)  A plane screen that
// the test implies comes first: t = num / ddotn in (1e-10, 1) needs num and
// ddotn of one sign (their product, when it underflows, bounds t below
// 1e-10) and |num| <= |ddotn|; only lanes that pass it load the edges, and
// waves whose lanes all fail it skip the divide.
__device__ __forceinline__ bool tri_test_record(const float4 pl, const float4 *__restrict__ edge, const Ray &r,
                                                float tmin)
{
    float ddotn = r.dx * pl.x;
    ddotn = (r.dy * pl.y) + ddotn;
    ddotn = (r.dz * pl.z) + ddotn;
    float odotn = r.ox * pl.x;
    odotn = (r.oy * pl.y) + odotn;
    odotn = (r.oz * pl.z) + odotn;
    const float num = pl.w - odotn;
    // the three screens evaluated together (no short circuit)
    const bool may = ((int)(fabsf(ddotn) > kParallelEps) & (int)(num * ddotn > 0.0f) &
                      (int)(fabsf(num) <= fabsf(ddotn))) != 0;
    bool hit = false;
    if (may) {
        const float4 e0 = edge[0], e1 = edge[1];
        const float t = num / ddotn;
        const float hx = r.ox + (t * r.dx), hy = r.oy + (t * r.dy), hz = r.oz + (t * r.dz);
        float b0 = e0.x * hx;
        b0 = b0 + (e0.y * hy);
        b0 = b0 + (e0.z * hz);
        b0 = b0 + e0.w;
        float b1 = e1.x * hx;
        b1 = b1 + (e1.y * hy);
        b1 = b1 + (e1.z * hz);
        b1 = b1 + e1.w;
        hit = (t > kParallelEps) & (b0 >= kParallelEps) & (b1 >= kParallelEps) & ((b1 + b0) <= 1.0f) &
              (t >= tmin) & (t < 1.0f);
    }
    return hit;
}

__device__ __forceinline__ void wave_add(unsigned long long *counters, int slot, uint32_t v)
{
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&counters[slot], x);
}

// Shadow rays from points (one lane each, in Morton order) to the nlg lights
// of one group, starting at light g0: each ray looks up the cell of its
// direction seen from the light and tests that cell's triangles in depth
// order.  Bit (p * nlights + g0 + l) is set when the ray is blocked.
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_point_lights_lb(
    const uint32_t *__restrict__ rdepth, const float4 *__restrict__ rplane, const float4 *__restrict__ redge,
    const uint32_t *__restrict__ off,
    const float *__restrict__ pts, const int32_t *__restrict__ order, int64_t npoints,
    const float *__restrict__ lights, int32_t g0, int32_t nlg, int32_t nlights, int32_t N,
    float tmin, uint64_t *__restrict__ bits, unsigned long long *__restrict__ counters)
{
    // Workgroups go to the 8 XCDs round robin (workgroup b to XCD b % 8), so
    // XCD x is given the x-th eighth of the points in Morton order: a compact
    // region, seen from each light through a part of its cells, whose lists
    // then fit that XCD's own L2.  The grid is a multiple of 8 blocks.
    const uint32_t per = gridDim.x >> 3;
    const int64_t blk = (int64_t)(blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    const int64_t slot = blk * kBlock + threadIdx.x;
    const bool have = slot < npoints;
    const int64_t p = have ? (order ? order[slot] : slot) : 0;
    const float px = have ? pts[p * 3 + 0] : 0.0f, py = have ? pts[p * 3 + 1] : 0.0f,
                pz = have ? pts[p * 3 + 2] : 0.0f;
    uint32_t nread = 0, ntest = 0, nslot = 0;  // list entries read, exact tests, wave steps
    uint64_t w = 0;
    for (int32_t l = 0; l < nlg; l++) {
        const float *L = lights + (size_t)(g0 + l) * 3;
        Ray r;
        r.ox = px;
        r.oy = py;
        r.oz = pz;
        r.dx = L[0] - px;
        r.dy = L[1] - py;
        r.dz = L[2] - pz;
        r.reach = 1.0f;
        bool blocked = false;
        uint32_t steps = 0;
        int face, iu, iw;
        float depth;
        if (have && ray_finite(r) && lb_cell_of(r.dx, r.dy, r.dz, N, face, iu, iw, depth)) {
            const size_t cell = (((size_t)l * 6 + face) * N + iu) * N + iw;
            const uint32_t end = off[(cell + 1) * kImageParts];
            const uint32_t dbits = __float_as_uint(depth);
            // each step loads the next entry's depth and plane before testing
            // this one, so the loads of consecutive steps overlap; the loop is
            // unrolled twice, so the two register sets take turns (entry j in
            // A while j + 1 loads into B, then the other way) and no step
            // moves one into the other.  Entry j + 1 is loaded only once
            // j < end <= total, so it is at most entry `total`: the plane
            // array is followed by the edges and the depth array by the
            // buffer's tail, and a value past the list is never used.
            uint32_t j = off[cell * kImageParts];
            uint32_t dA = j < end ? rdepth[j] : 0xffffffffu;
            float4 pA = j < end ? rplane[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uint32_t dB;
            float4 pB;
            for (;;) {
                if (j >= end) break;
                dB = rdepth[j + 1];
                pB = rplane[j + 1];
                if (COUNT) ++steps;
                if (dA >= dbits) break;  // this and every later triangle starts at or past p
                if (COUNT) ++ntest;
                if (tri_test_record(pA, redge + (size_t)j * 2, r, tmin)) {
                    blocked = true;
                    break;
                }
                j++;
                if (j >= end) break;
                dA = rdepth[j + 1];
                pA = rplane[j + 1];
                if (COUNT) ++steps;
                if (dB >= dbits) break;
                if (COUNT) ++ntest;
                if (tri_test_record(pB, redge + (size_t)j * 2, r, tmin)) {
                    blocked = true;
                    break;
                }
                j++;
            }
        }
        if (COUNT) {
            // the wave runs as many list steps as its longest lane
            nread += steps;
            uint32_t m = steps;
            for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
            nslot += m;
        }
        w |= (uint64_t)blocked << l;
    }
    if (have) {
        if ((nlights & 63) == 0 && nlg == 64) {
            bits[p * (nlights >> 6) + (g0 >> 6)] = w;
        } else {
            for (int32_t l = 0; l < nlg; l++) {
                if ((w >> l) & 1) {
                    const int64_t rix = p * nlights + g0 + l;
                    atomicOr((unsigned long long *)&bits[rix >> 6], 1ull << (rix & 63));
                }
            }
        }
    }
    if (COUNT) {
        wave_add(counters, 0, nread);
        wave_add(counters, 1, ntest);
        if ((threadIdx.x & 63) == 0) atomicAdd(&counters[2], (unsigned long long)nslot * 64ull);
    }
}

}  // namespace

int lb_resolution(vh_ctx *c) { return c->lb_cells > 0 ? c->lb_cells : kLightBufferCells; }

// The light-buffer form of launch_point_lights: the caller has sorted the
// points (order, device) and cleared the bits when nlights is not a multiple
// of 64.  Per group of up to 64 lights: count, scan, fill and sort the cells,
// then trace.  The entry count is read back once per group to size the list.
int launch_point_lights_lb(vh_ctx *c, const float *d_points, const int32_t *order, int64_t npoints,
                           const float *d_lights, int32_t nlights, float tmin, uint64_t *d_bits,
                           hipStream_t s)
{
    int rc;
    const int N = lb_resolution(c);
    const int32_t ntri = c->ntri;
    if ((rc = c->lb_boxes.ensure((size_t)ntri * 32, "light-buffer boxes"))) return rc;
    if (c->have_verts && (rc = c->lb_tverts.ensure((size_t)ntri * 48, "light-buffer vertices"))) return rc;
    if (c->lb_boxes_gen != c->scene_gen) {
        hipLaunchKernelGGL(k_lb_boxes, dim3((uint32_t)((c->nnodes + 255) / 256)), dim3(256), 0, s,
                           (const float4 *)c->nodes.ptr, c->nnodes, (const float4 *)c->tris.ptr,
                           (float4 *)c->lb_boxes.ptr);
        if (c->have_verts)
            hipLaunchKernelGGL(k_lb_tverts, dim3((uint32_t)((ntri + 255) / 256)), dim3(256), 0, s,
                               (const float4 *)c->tris.ptr, ntri, (const float *)c->verts.ptr,
                               (float4 *)c->lb_tverts.ptr);
        VH_HIP(hipGetLastError());
        c->lb_boxes_gen = c->scene_gen;
    }
    const int32_t nlg_max = nlights < 64 ? nlights : 64;
    const size_t ncell = (size_t)6 * N * N * nlg_max;
    const size_t nslot = ncell * kImageParts;  // (cell, part) slots
    if ((rc = c->lb_off.ensure((nslot + 1) * 4, "light-buffer cells")) ||
        (rc = c->lb_cur.ensure(nslot * 4 + 256, "light-buffer counts")))
        return rc;
    if (!c->lb_pinned) VH_HIP(hipHostMalloc((void **)&c->lb_pinned, 64, hipHostMallocDefault));
    // a word past the counts and the scan's extra element (which each group
    // clears): the points' largest coordinate
    uint32_t *d_pmax = (uint32_t *)((char *)c->lb_cur.ptr + nslot * 4 + 128);
    VH_HIP(hipMemsetAsync(d_pmax, 0, 4, s));
    hipLaunchKernelGGL(k_lb_pmax, dim3(256), dim3(256), 0, s, d_points, npoints, d_pmax);
    VH_HIP(hipGetLastError());
    float scene_mag = 0.0f;
    for (int k = 0; k < 3; k++) scene_mag = fmaxf(scene_mag, fmaxf(fabsf(c->lo[k]), fabsf(c->hi[k])));
    const float pad = lb_pad(scene_mag);
    const float4 *tv = c->have_verts ? (const float4 *)c->lb_tverts.ptr : nullptr;
    size_t scan_bytes = 0;
    VH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t *)c->lb_cur.ptr,
                                            (uint32_t *)c->lb_off.ptr, (int)(nslot + 1), s));
    if ((rc = c->sort_tmp.ensure(scan_bytes, "scan scratch"))) return rc;
    const int64_t blocks = ((npoints + kBlock - 1) / kBlock + 7) / 8 * 8;  // a multiple of 8 (XCD split)
    for (int32_t g0 = 0; g0 < nlights; g0 += 64) {
        const int32_t nlg = nlights - g0 < 64 ? nlights - g0 : 64;
        const size_t nc = (size_t)6 * N * N * nlg;
        const float *gl = d_lights + (size_t)g0 * 3;
        const float4 *boxes = (const float4 *)c->lb_boxes.ptr;
        uint32_t *cur = (uint32_t *)c->lb_cur.ptr;
        const uint32_t *off = (const uint32_t *)c->lb_off.ptr;
        KernelTimer build(c, VH_KERNEL_LIGHT_BUFFER, s);
        const size_t lds = ((size_t)N * N + kBigCap) * 4;
        const size_t ns = nc * kImageParts;
        VH_HIP(hipMemsetAsync(cur + ns, 0, 4, s));
        hipLaunchKernelGGL(k_lb_image<false>, dim3((uint32_t)(nlg * 6 * kImageParts)), dim3(kImageThreads), lds, s, boxes,
                           tv, ntri, gl, N, d_pmax, scene_mag, pad, cur, nullptr, Lists{nullptr, nullptr});
        VH_HIP(hipGetLastError());
        VH_HIP(hipcub::DeviceScan::ExclusiveSum(c->sort_tmp.ptr, scan_bytes, cur, (uint32_t *)c->lb_off.ptr,
                                                (int)(ns + 1), s));
        VH_HIP(hipMemcpyAsync(c->lb_pinned, (uint32_t *)c->lb_off.ptr + ns, 4, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
        const uint32_t total = *(volatile uint32_t *)c->lb_pinned;
        // keys (8 B) and cells (4 B) in fill order; records: depth (4 B), plane (16 B), edges (32 B)
        if ((rc = c->lb_ent.ensure((size_t)total * 12 + 64, "light-buffer lists")) ||
            (rc = c->lb_rec.ensure((size_t)total * 52 + 256, "light-buffer records")))
            return rc;
        Lists ent;
        ent.key = (int2 *)c->lb_ent.ptr;
        ent.cell = (uint32_t *)((char *)c->lb_ent.ptr + (size_t)total * 8);
        Records rec;
        rec.plane = (float4 *)c->lb_rec.ptr;
        rec.edge = (float4 *)((char *)c->lb_rec.ptr + (size_t)total * 16);
        rec.depth = (uint32_t *)((char *)c->lb_rec.ptr + (size_t)total * 48);
        hipLaunchKernelGGL(k_lb_image<true>, dim3((uint32_t)(nlg * 6 * kImageParts)), dim3(kImageThreads), lds, s, boxes,
                           tv, ntri, gl, N, d_pmax, scene_mag, pad, nullptr, off, ent);
        VH_HIP(hipGetLastError());
        if (total > 0)
            hipLaunchKernelGGL(k_lb_records, dim3((total + 255) / 256), dim3(256), 0, s, (const int2 *)ent.key,
                               (const uint32_t *)ent.cell, total, off, (uint32_t)nc, (const float4 *)c->tris.ptr,
                               ntri, rec);
        VH_HIP(hipGetLastError());
        build.done();
        KernelTimer timer(c, VH_KERNEL_POINT_LIGHTS, s);
        if (c->counting)
            hipLaunchKernelGGL(k_point_lights_lb<true>, dim3((uint32_t)blocks), dim3(kBlock), 0, s,
                               (const uint32_t *)rec.depth, (const float4 *)rec.plane, (const float4 *)rec.edge, off,
                               d_points, order, npoints, d_lights, g0, nlg, nlights, N, tmin, d_bits,
                               (unsigned long long *)c->counters.ptr);
        else
            hipLaunchKernelGGL(k_point_lights_lb<false>, dim3((uint32_t)blocks), dim3(kBlock), 0, s,
                               (const uint32_t *)rec.depth, (const float4 *)rec.plane, (const float4 *)rec.edge, off,
                               d_points, order, npoints, d_lights, g0, nlg, nlights, N, tmin, d_bits, nullptr);
        timer.done();
        VH_HIP(hipGetLastError());
    }
    return VH_OK;
}

}  // namespace vh
