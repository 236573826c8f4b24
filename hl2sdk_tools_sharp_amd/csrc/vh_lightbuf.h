// The light buffers' cell geometry (vh_lightbuf.hip), host- and
// device-callable so the host can check it against brute force.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace vh {

constexpr float kCellGrow = 1.52587890625e-5f;  // 2^-16, face-coordinate slack
constexpr float kDepthShrink = 0.99999905f;     // 1 - 2^-20
constexpr float kMinDepth = 1.4551915e-11f;     // 2^-36: shorter directions cannot hit

__host__ __device__ inline int lb_imax(int a, int b) { return a > b ? a : b; }
__host__ __device__ inline int lb_imin(int a, int b) { return a < b ? a : b; }

struct FaceRect {
    int iu0, iu1, iw0, iw1;
    float dlo;
};

// The cells of face (axis a, side s) that a box (relative to the light) can be
// seen in, and its smallest depth there.  Side 0 holds the directions with
// v_a >= max(|v_b|, |v_c|), side 1 those with -v_a >= max(|v_b|, |v_c|);
// b = (a + 1) % 3 and c = (a + 2) % 3 are the face's u and w axes.
__host__ __device__ inline bool lb_face_rect(const float lo[3], const float hi[3], int a, int s, int N,
                                          FaceRect &r)
{
    const int b = a == 2 ? 0 : a + 1, c = a == 0 ? 2 : a - 1;
    float dlo, dhi;
    if (s == 0) {
        if (hi[a] < 0.0f) return false;
        dlo = fmaxf(lo[a], 0.0f);
        dhi = hi[a];
    } else {
        if (lo[a] > 0.0f) return false;
        dlo = fmaxf(-hi[a], 0.0f);
        dhi = -lo[a];
    }
    // a point of the box seen in this face has depth >= its |v_b| and |v_c|
    const float mb = (lo[b] <= 0.0f && hi[b] >= 0.0f) ? 0.0f : fminf(fabsf(lo[b]), fabsf(hi[b]));
    const float mc = (lo[c] <= 0.0f && hi[c] >= 0.0f) ? 0.0f : fminf(fabsf(lo[c]), fabsf(hi[c]));
    dlo = fmaxf(dlo, fmaxf(mb, mc)) * kDepthShrink;
    if (dlo > dhi) return false;
    float umin = -2.0f, umax = 2.0f, wmin = -2.0f, wmax = 2.0f;
    if (dlo > 0.0f) {
        // u = v_b / depth is monotonic in each of v_b and depth over the box,
        // so its extremes are at the corners
        const float u0 = lo[b] / dlo, u1 = hi[b] / dlo, u2 = lo[b] / dhi, u3 = hi[b] / dhi;
        const float w0 = lo[c] / dlo, w1 = hi[c] / dlo, w2 = lo[c] / dhi, w3 = hi[c] / dhi;
        umin = fminf(fminf(u0, u1), fminf(u2, u3));
        umax = fmaxf(fmaxf(u0, u1), fmaxf(u2, u3));
        wmin = fminf(fminf(w0, w1), fminf(w2, w3));
        wmax = fmaxf(fmaxf(w0, w1), fmaxf(w2, w3));
        umin -= kCellGrow * fmaxf(1.0f, fabsf(umin));
        umax += kCellGrow * fmaxf(1.0f, fabsf(umax));
        wmin -= kCellGrow * fmaxf(1.0f, fabsf(wmin));
        wmax += kCellGrow * fmaxf(1.0f, fabsf(wmax));
        if (umin > 1.0f || umax < -1.0f || wmin > 1.0f || wmax < -1.0f) return false;
    }
    const float h = 0.5f * (float)N;
    r.iu0 = lb_imax(0, (int)floorf((fmaxf(umin, -1.0f) + 1.0f) * h));
    r.iu1 = lb_imin(N - 1, (int)floorf((fminf(umax, 1.0f) + 1.0f) * h));
    r.iw0 = lb_imax(0, (int)floorf((fmaxf(wmin, -1.0f) + 1.0f) * h));
    r.iw1 = lb_imin(N - 1, (int)floorf((fminf(wmax, 1.0f) + 1.0f) * h));
    r.dlo = dlo;
    return true;
}

// A triangle's projection onto one cube face, widened for rasterising it into
// the face's cells.  A hit the exact test accepts lies within L-infinity
// distance `pad` of its triangle (the scene builder's widening, far beyond the
// hit point's few ulps of rounding), and the light's rounding adds delta (see
// vh_lightbuf.hip), so every point a ray can be blocked at lies within
// r >= pad + delta of the triangle.  Seen from the light, a point x = y + e
// (y on the triangle, |e| <= r per axis) whose direction is in this face
// (|u(x)|, |w(x)| <= 1) has |u(x) - u(y)| <= 2r / (z(y) - 2r), z the depth
// along the face's axis: so x's face coordinates lie within
// m = 2r / (zmin - 2r) of the projected triangle (whose vertices' depths are
// all above 4r; a triangle that comes closer to the light's face plane keeps
// its box rectangle).  A cell is kept when its square, grown by kCellGrow,
// meets the projected triangle grown by m: the separating-axis test over the
// square's axes (the cell range, from the grown triangle's bounding box) and
// the triangle's three edge normals (edge functions grown by m(|A| + |B|) and
// a rounding slack).
struct TriRaster {
    float A[3], B[3], C[3];  // edge functions, >= 0 inside the grown triangle
    int iu0, iu1, iw0, iw1;  // cells of the grown bounding box on this face
};

__host__ __device__ inline bool lb_tri_raster(const float v[9], int a, int s, float r, int N, TriRaster &t)
{
    const int b = a == 2 ? 0 : a + 1, c = a == 0 ? 2 : a - 1;
    float u[3], w[3], zmin = 3.4028235e38f, big = 1.0f;
    for (int q = 0; q < 3; q++) {
        const float z = s == 0 ? v[q * 3 + a] : -v[q * 3 + a];
        if (!(z > 4.0f * r)) return false;
        zmin = fminf(zmin, z);
        u[q] = v[q * 3 + b] / z;
        w[q] = v[q * 3 + c] / z;
        big = fmaxf(big, fmaxf(fabsf(u[q]), fabsf(w[q])));
    }
    if (!(big < 1024.0f)) return false;  // far outside the face: keep the box rectangle
    const float m = 2.0f * r / (zmin - 2.0f * r) * 1.0009765625f + 1.52587890625e-5f * big;
    const float area = (u[1] - u[0]) * (w[2] - w[0]) - (u[2] - u[0]) * (w[1] - w[0]);
    if (!(fabsf(area) > 1.0e-12f)) return false;  // seen edge on: keep the box rectangle
    for (int e = 0; e < 3; e++) {
        const int i = e, j = e == 2 ? 0 : e + 1;
        float A = w[i] - w[j], B = u[j] - u[i], C = u[i] * w[j] - u[j] * w[i];
        if (area < 0.0f) {
            A = -A;
            B = -B;
            C = -C;
        }
        t.A[e] = A;
        t.B[e] = B;
        t.C[e] = C + m * (fabsf(A) + fabsf(B)) + 3.814697265625e-6f * (big * (fabsf(A) + fabsf(B)) + big * big);
    }
    const float umin = fminf(fminf(u[0], u[1]), u[2]) - m, umax = fmaxf(fmaxf(u[0], u[1]), u[2]) + m;
    const float wmin = fminf(fminf(w[0], w[1]), w[2]) - m, wmax = fmaxf(fmaxf(w[0], w[1]), w[2]) + m;
    if (umin > 1.0f || umax < -1.0f || wmin > 1.0f || wmax < -1.0f) {
        t.iu0 = 1;  // an empty range: the grown triangle misses the face
        t.iu1 = 0;
        t.iw0 = 0;
        t.iw1 = 0;
        return true;
    }
    const float h = 0.5f * (float)N;
    t.iu0 = lb_imax(0, (int)floorf((fmaxf(umin, -1.0f) - kCellGrow + 1.0f) * h));
    t.iu1 = lb_imin(N - 1, (int)floorf((fminf(umax, 1.0f) + kCellGrow + 1.0f) * h));
    t.iw0 = lb_imax(0, (int)floorf((fmaxf(wmin, -1.0f) - kCellGrow + 1.0f) * h));
    t.iw1 = lb_imin(N - 1, (int)floorf((fminf(wmax, 1.0f) + kCellGrow + 1.0f) * h));
    return true;
}

// Whether cell (iu, iw)'s square, grown by kCellGrow, meets the grown triangle
// on every edge normal (the largest edge function over its corners >= 0).
__host__ __device__ inline bool lb_tri_cell(const TriRaster &t, int iu, int iw, int N)
{
    const float step = 2.0f / (float)N;
    const float u0 = -1.0f + step * (float)iu - kCellGrow, u1 = -1.0f + step * (float)(iu + 1) + kCellGrow;
    const float w0 = -1.0f + step * (float)iw - kCellGrow, w1 = -1.0f + step * (float)(iw + 1) + kCellGrow;
    for (int e = 0; e < 3; e++) {
        const float au = t.A[e] > 0.0f ? t.A[e] * u1 : t.A[e] * u0;
        const float bw = t.B[e] > 0.0f ? t.B[e] * w1 : t.B[e] * w0;
        if ((au + bw) + t.C[e] < 0.0f) return false;
    }
    return true;
}

// The widening beyond the scene builders' pad for one light: the rounding of a
// ray's direction fl(L - p) moves its end by at most 2^-24 (|L| + |p|) per
// axis, and this bounds it (|p| <= pmax, the scene within scene_mag) with room.
__host__ __device__ inline float lb_delta(float Lx, float Ly, float Lz, float pmax, float scene_mag)
{
    const float lmag = fmaxf(fmaxf(fabsf(Lx), fabsf(Ly)), fabsf(Lz));
    return 4.76837158203125e-7f * (lmag + pmax + scene_mag) + 9.3132257e-10f;  // 2^-21 (...) + 2^-30
}

// The scene builders' widening (vh_bvh.cpp, vh_build.hip): 2^-16 of the
// largest vertex coordinate, which the scene's bounds hold.
__host__ __device__ inline float lb_pad(float scene_mag) { return scene_mag * (1.0f / 65536.0f) + 1.0e-6f; }

// The cells of one triangle on one face of one light, in two steps.
// lb_cells_box: lo/hi, the triangle's (leaf) box relative to the light, already
// widened by pad + delta, gives the box rectangle; false when it has none.
// When lb_cells_cut says a rectangle has more than two cells, lb_cells_raster
// cuts it to the cells the triangle's widened projection meets: v, its
// vertices relative to the light; r, the widening pad + 2 delta (delta again
// for the rounding of v).  False when no cell is left.
struct LbCells {
    FaceRect box;
    TriRaster tri;
    bool rast;
    int iu0, iu1, iw0, iw1;
};

__host__ __device__ inline bool lb_cells_box(const float lo[3], const float hi[3], int a, int side, int N,
                                             LbCells &pc)
{
    if (!lb_face_rect(lo, hi, a, side, N, pc.box)) return false;
    pc.iu0 = pc.box.iu0;
    pc.iu1 = pc.box.iu1;
    pc.iw0 = pc.box.iw0;
    pc.iw1 = pc.box.iw1;
    pc.rast = false;
    return true;
}

// A rectangle of one or two cells has little to cut: it is kept.
__host__ __device__ inline bool lb_cells_cut(const LbCells &pc)
{
    return (pc.box.iu1 - pc.box.iu0 + 1) * (pc.box.iw1 - pc.box.iw0 + 1) > 2;
}

__host__ __device__ inline bool lb_cells_raster(const float v[9], int a, int side, int N, float r, LbCells &pc)
{
    pc.rast = lb_tri_raster(v, a, side, r, N, pc.tri);
    if (pc.rast) {
        pc.iu0 = lb_imax(pc.iu0, pc.tri.iu0);
        pc.iu1 = lb_imin(pc.iu1, pc.tri.iu1);
        pc.iw0 = lb_imax(pc.iw0, pc.tri.iw0);
        pc.iw1 = lb_imin(pc.iw1, pc.tri.iw1);
    }
    return pc.iu0 <= pc.iu1 && pc.iw0 <= pc.iw1;
}

// Whether cell (iu, iw), inside the range lb_cells_* left, lists the triangle.
__host__ __device__ inline bool lb_cells_has(const LbCells &pc, int iu, int iw, int N)
{
    return !pc.rast || lb_tri_cell(pc.tri, iu, iw, N);
}

// The cell of a shadow ray's direction d (light minus point) seen from the
// light: -d's face (2 a + side), its cell index on that face and its depth
// |d_a|.  Returns false when d is too short for any triangle to be hit
// (every exact test rejects |d . n| <= 1e-10 for a unit normal n).
__host__ __device__ inline bool lb_cell_of(float dx, float dy, float dz, int N, int &face, int &iu,
                                           int &iw, float &depth)
{
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    const int a = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    depth = a == 0 ? ax : (a == 1 ? ay : az);
    if (!(depth >= kMinDepth)) return false;
    const float va = a == 0 ? dx : (a == 1 ? dy : dz);
    const float vb = -(a == 0 ? dy : (a == 1 ? dz : dx));
    const float vc = -(a == 0 ? dz : (a == 1 ? dx : dy));
    const float inv = 1.0f / depth;
    const float h = 0.5f * (float)N;
    iu = lb_imin(N - 1, lb_imax(0, (int)((vb * inv + 1.0f) * h)));
    iw = lb_imin(N - 1, lb_imax(0, (int)((vc * inv + 1.0f) * h)));
    // -d has v_a = -va: side 0 when -va >= 0
    face = 2 * a + (va > 0.0f ? 1 : 0);
    return true;
}

}  // namespace vh
