// Checks the light buffers' cell geometry (hl2sdk_tools_sharp_amd/csrc/vh_lightbuf.h,
// the same host/device functions the build and trace kernels call) against the
// oracle's exact test by brute force: for every shadow ray, the triangles its
// cell lists before its depth must include every triangle that blocks it, so
// "some listed triangle blocks" equals "some triangle blocks".  Test
// infrastructure: the oracle (libvradcpu.so) is the checker.
//
// Scenes: seeded axis-aligned boxes (faces split in two, as the visibility
// config) and a soup of arbitrarily oriented triangles.  Lights: random ones,
// plus lights inside triangle boxes, on their faces, edges and corners, and at
// vertices.  Points: random ones plus points sharing one, two or three
// coordinates with a light and points on a 45-degree face boundary.  Each
// triangle's own box stands in for its leaf's box (a leaf's box holds it, so
// the product's lists hold these).
//
// Prints "rays R blocked B mismatches M entries E" per case; exits 1 on a
// mismatch.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "vh_lightbuf.h"
#include "vradcpu.h"

using namespace vh;

namespace {

struct Scene {
    std::vector<float> v9;
    std::vector<vc_tri> tris;
    float mag = 0.0f;
};

void finish(Scene &s)
{
    const int n = (int)(s.v9.size() / 9);
    std::vector<int32_t> ids(n);
    for (int i = 0; i < n; i++) ids[i] = i;
    s.tris.resize(n);
    vc_tri_convert(s.v9.data(), ids.data(), nullptr, n, s.tris.data());
    for (float x : s.v9) s.mag = std::max(s.mag, std::fabs(x));
}

Scene boxes(std::mt19937 &g, int nbox, float vol)
{
    Scene s;
    std::uniform_real_distribution<float> C(0.0f, vol), E(std::log(8.0f), std::log(300.0f));
    for (int b = 0; b < nbox; b++) {
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            const float c = C(g), e = std::exp(E(g));
            lo[k] = c - 0.5f * e;
            hi[k] = c + 0.5f * e;
        }
        for (int ax = 0; ax < 3; ax++)
            for (int side = 0; side < 2; side++) {
                const int u = (ax + 1) % 3, v = (ax + 2) % 3;
                float q[4][3];
                const int su[4] = {0, 1, 1, 0}, sv[4] = {0, 0, 1, 1};
                for (int j = 0; j < 4; j++) {
                    q[j][ax] = side ? hi[ax] : lo[ax];
                    q[j][u] = su[j] ? hi[u] : lo[u];
                    q[j][v] = sv[j] ? hi[v] : lo[v];
                }
                const int t0[3] = {0, 1, 2}, t1[3] = {0, 2, 3};
                for (const int *t : {t0, t1})
                    for (int j = 0; j < 3; j++)
                        for (int k = 0; k < 3; k++) s.v9.push_back(q[side ? t[j] : t[2 - j]][k]);
            }
    }
    finish(s);
    return s;
}

Scene soup(std::mt19937 &g, int ntri, float vol)
{
    Scene s;
    std::uniform_real_distribution<float> C(0.0f, vol), D(-50.0f, 50.0f);
    for (int i = 0; i < ntri; i++) {
        const float c[3] = {C(g), C(g), C(g)};
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) s.v9.push_back(c[k] + D(g));
    }
    finish(s);
    return s;
}

bool blocked_by(const vc_tri *t, int32_t n, const float ray[7], float tmin)
{
    uint64_t w = 0;
    vc_trace_visibility_brute(t, n, ray, 7, 1.0f, 1, tmin, -1, 0, &w);
    return (w & 1) != 0;
}

int run(const char *name, const Scene &s, const std::vector<float> &L, const std::vector<float> &P, int N,
        float tmin)
{
    const int ntri = (int)s.tris.size(), nl = (int)(L.size() / 3), np = (int)(P.size() / 3);
    float pmax = 0.0f;
    for (float x : P) pmax = std::max(pmax, std::fabs(x));
    const float pad = lb_pad(s.mag);
    long long rays = 0, blocked = 0, bad = 0, entries = 0;
    for (int l = 0; l < nl; l++) {
        const float Lx = L[l * 3], Ly = L[l * 3 + 1], Lz = L[l * 3 + 2];
        const float delta = lb_delta(Lx, Ly, Lz, pmax, s.mag);
        const float r = (pad + 2.0f * delta) * 1.0009765625f;
        std::vector<std::vector<std::pair<uint32_t, int>>> cells((size_t)6 * N * N);
        for (int k = 0; k < ntri; k++) {
            const float *v = &s.v9[(size_t)k * 9];
            float lo[3], hi[3], vr[9];
            const float Lc[3] = {Lx, Ly, Lz};
            for (int c = 0; c < 3; c++) {
                lo[c] = std::min(v[c], std::min(v[3 + c], v[6 + c])) - pad - Lc[c] - delta;
                hi[c] = std::max(v[c], std::max(v[3 + c], v[6 + c])) + pad - Lc[c] + delta;
                for (int q = 0; q < 3; q++) vr[q * 3 + c] = v[q * 3 + c] - Lc[c];
            }
            for (int f = 0; f < 6; f++) {
                LbCells pc;
                if (!lb_cells_box(lo, hi, f >> 1, f & 1, N, pc)) continue;
                if (lb_cells_cut(pc) && !lb_cells_raster(vr, f >> 1, f & 1, N, r, pc)) continue;
                uint32_t dbits;
                std::memcpy(&dbits, &pc.box.dlo, 4);
                for (int iu = pc.iu0; iu <= pc.iu1; iu++)
                    for (int iw = pc.iw0; iw <= pc.iw1; iw++)
                        if (lb_cells_has(pc, iu, iw, N)) {
                            cells[((size_t)f * N + iu) * N + iw].push_back({dbits, k});
                            entries++;
                        }
            }
        }
        for (auto &c : cells) std::sort(c.begin(), c.end());
        for (int i = 0; i < np; i++) {
            const float ray[7] = {P[i * 3], P[i * 3 + 1], P[i * 3 + 2], Lx - P[i * 3], Ly - P[i * 3 + 1],
                                  Lz - P[i * 3 + 2], 1.0f};
            const bool truth = blocked_by(s.tris.data(), ntri, ray, tmin);
            bool lb = false;
            int face, iu, iw;
            float depth;
            if (lb_cell_of(ray[3], ray[4], ray[5], N, face, iu, iw, depth)) {
                uint32_t db;
                std::memcpy(&db, &depth, 4);
                for (const auto &e : cells[((size_t)face * N + iu) * N + iw]) {
                    if (e.first >= db) break;
                    if (blocked_by(&s.tris[e.second], 1, ray, tmin)) {
                        lb = true;
                        break;
                    }
                }
            }
            rays++;
            blocked += truth;
            bad += truth != lb;
        }
    }
    std::printf("%s N=%d tmin=%g: rays %lld blocked %lld mismatches %lld entries %lld\n", name, N, tmin, rays,
                blocked, bad, entries);
    return bad == 0 ? 0 : 1;
}

void adversarial_lights(std::mt19937 &g, const Scene &s, int n, std::vector<float> &L)
{
    std::uniform_int_distribution<int> K(0, (int)s.tris.size() - 1);
    std::uniform_real_distribution<float> F(0.0f, 1.0f);
    for (int i = 0; i < n; i++) {
        const float *v = &s.v9[(size_t)K(g) * 9];
        float lo[3], hi[3], p[3];
        for (int c = 0; c < 3; c++) {
            lo[c] = std::min(v[c], std::min(v[3 + c], v[6 + c]));
            hi[c] = std::max(v[c], std::max(v[3 + c], v[6 + c]));
            p[c] = lo[c] + F(g) * (hi[c] - lo[c]);
        }
        if (i % 4 == 1) p[0] = lo[0];                    // on a face plane of the box
        if (i % 4 == 2) p[0] = lo[0], p[1] = hi[1];      // on an edge
        if (i % 4 == 3) p[0] = v[0], p[1] = v[1], p[2] = v[2];  // at a vertex
        L.insert(L.end(), p, p + 3);
    }
}

void aligned_points(std::mt19937 &g, const std::vector<float> &L, std::vector<float> &P)
{
    std::uniform_int_distribution<int> I(0, (int)(P.size() / 3) - 1);
    const int nl = (int)(L.size() / 3);
    for (int l = 0; l < nl; l += 2) {
        for (int j = 0; j < 8; j++) {
            const int i = I(g);
            float q[3] = {P[i * 3], P[i * 3 + 1], P[i * 3 + 2]};
            const float *c = &L[l * 3];
            if (j < 2) q[0] = c[0];
            else if (j < 4) q[0] = c[0], q[1] = c[1];
            else if (j < 5) q[0] = c[0], q[1] = c[1], q[2] = c[2];
            else q[1] = c[1] + (q[2] - c[2]);  // direction on a 45-degree face boundary
            P.insert(P.end(), q, q + 3);
        }
    }
}

}  // namespace

int main()
{
    int fail = 0;
    std::mt19937 g(2024);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    {
        Scene s = boxes(g, 160, 2000.0f);
        std::vector<float> L, P;
        for (int i = 0; i < 16 * 3; i++) L.push_back(2000.0f * U(g));
        adversarial_lights(g, s, 16, L);
        for (int i = 0; i < 700; i++) {  // points one unit off box faces
            const float *v = &s.v9[(size_t)(i * 7 % (s.v9.size() / 9)) * 9];
            const float a = U(g), b = U(g) * (1.0f - a);
            const vc_tri &t = s.tris[(size_t)(i * 7 % (s.v9.size() / 9))];
            for (int c = 0; c < 3; c++) {
                const float n = c == 0 ? t.nx : (c == 1 ? t.ny : t.nz);
                P.push_back(v[c] + a * (v[3 + c] - v[c]) + b * (v[6 + c] - v[c]) + n);
            }
        }
        aligned_points(g, L, P);
        fail |= run("boxes", s, L, P, 64, 0.0f);
        fail |= run("boxes", s, L, P, 5, 0.25f);
        fail |= run("boxes", s, L, P, 120, 1e-3f);
    }
    {
        Scene s = soup(g, 1500, 1500.0f);
        std::vector<float> L, P;
        for (int i = 0; i < 16 * 3; i++) L.push_back(1500.0f * U(g));
        adversarial_lights(g, s, 16, L);
        for (int i = 0; i < 700 * 3; i++) P.push_back(1500.0f * U(g));
        aligned_points(g, L, P);
        fail |= run("soup", s, L, P, 64, 0.0f);
        fail |= run("soup", s, L, P, 17, 0.25f);
    }
    return fail;
}
