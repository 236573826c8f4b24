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
