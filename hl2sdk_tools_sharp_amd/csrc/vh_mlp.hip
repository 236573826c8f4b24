// The warm-start network on gfx950: a per-patch MLP (25 -> 128 -> 128 -> 128 -> 3,
// GELU) that predicts a patch's converged bounced light from its position,
// normal, direct light and Fourier features of its position.  It has no
// counterpart in the reference; the bounce (vh_bounce.hip) iterates from its
// prediction to the same residual, so it changes only where the iteration
// starts.
//
// The prediction is max(gain * direct + scale * y, 0): the network's output y
// is the correction to a start proportional to the patch's own direct light
// (the gain is the bounced-to-direct ratio of the training maps), in units of
// `scale`.  The correction is learned in linear light so the squared error
// weighs patches as the bounce's L1 residual does; the data set keeps each
// row's direct and bounced light, and the targets (bounced - gain * direct) /
// scale are formed when a batch is drawn, so gain and scale can change
// without rebuilding it.
//
// Precision: FP32 master weights and Adam state; BF16 copies of the weights,
// BF16 activations and hidden-layer gradients as MFMA operands; FP32
// accumulation (v_mfma_f32_16x16x32_bf16), FP32 biases, pre-activations and
// output.
//
// Every GEMM is one kernel (k_gemm): a wave owns one 16 x 16 tile of C and
// walks K in steps of 32, one MFMA per step.  The 64 lanes hold the operands as
// the instruction wants them: lane l carries row (l & 15) of A and column
// (l & 15) of B at the eight K positions 8 * (l >> 4) .. + 7; its four results
// are rows 4 * (l >> 4) .. + 3 of column (l & 15).  Operands whose K runs along
// memory load as one 16-byte vector; transposed operands are gathered.  The
// epilogue is fused per use: bias + GELU + BF16 store (forward), bias (output
// layer), GELU' x the incoming gradient + BF16 store (backward data), and the
// weight gradient with the bias gradient as the column sums of the B operand
// (backward weights).
//
// Layout: every layer is padded to MFMA shapes (inputs 25 -> 32, outputs
// 3 -> 32); padded weights start at zero, see zero gradients and stay zero.
// A training step (batch gather by an affine permutation of the data set,
// forward, loss, backward, Adam) is captured once into a hipGraph and replayed;
// its counters live in device memory so every replay advances by itself.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "vh_internal.h"

#pragma clang fp contract(off)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace vh {
namespace {

constexpr int kLayers = 4;
constexpr int kIn = 25, kHidden = 128, kOut = 3, kFourier = 8;
constexpr int kInPad = 32, kOutPad = 32;
constexpr int kLight = 6;  // data-set light per row: direct rgb, bounced rgb
constexpr int kMaxSplits = 32;  // row splits of a weight-gradient GEMM
constexpr int kInDim[kLayers] = {kIn, kHidden, kHidden, kHidden};
constexpr int kOutDim[kLayers] = {kHidden, kHidden, kHidden, kOut};
constexpr int kInP[kLayers] = {kInPad, kHidden, kHidden, kHidden};
constexpr int kOutP[kLayers] = {kHidden, kHidden, kHidden, kOutPad};
constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)

struct Layout {
    int64_t w[kLayers], b[kLayers], wt[kLayers];  // offsets into the padded arrays
    int64_t nparam;                               // padded FP32 parameters
    int64_t nweight;                              // padded weights (BF16 copies)
};

Layout layout()
{
    Layout L{};
    int64_t o = 0, ow = 0;
    for (int l = 0; l < kLayers; l++) {
        L.w[l] = o;
        o += (int64_t)kInP[l] * kOutP[l];
        L.b[l] = o;
        o += kOutP[l];
        L.wt[l] = ow;  // BF16 copies: W [in][out] at w-offset order, W^T at wt
        ow += (int64_t)kInP[l] * kOutP[l];
    }
    L.nparam = o;
    L.nweight = ow;
    return L;
}

}  // namespace
}  // namespace vh

struct MlpData {
    vh::Layout L;
    bool ready = false;
    float length = 1024.0f;
    float floor = 1e-7f;  // the direct light enters as log(max(x, 0) + floor)
    float gain = 1.0f;    // prediction = max(gain * direct + scale * y, 0)
    float scale = 1.0f;
    float freq[24];       // the 8 x 3 Fourier frequency matrix, row-major
    vh::DevBuf master, m, v, grad, wb, wtb;
    // data set: BF16 features (rows x 32); FP32 direct then bounced light (rows x 6)
    vh::DevBuf X, T;
    int64_t rows = 0, cap = 0;
    // activations of one pass over `act_rows` rows
    vh::DevBuf h[4], z[3], y, dy, dz[3], xf, tf, loss_part, losses, state, loss1;
    // the weight-gradient GEMMs' operands, K-blocked: element (row b, column c)
    // of an R x C matrix at ((b / 8) * C + c) * 8 + b % 8, so eight consecutive
    // rows of one column are one 16-byte load; hk[l] holds layer l's input,
    // dzk[l] the gradient at its output; dw_part the split-K partial sums
    vh::DevBuf hk[4], dzk[4], dw_part, cs_part;
    int64_t act_rows = 0;
    int64_t step = 0;  // Adam steps taken
    // the prediction's rows (leaf, non-sky patches by cluster) for one patch system
    vh::DevBuf pred_rows, pred_groups;
    int64_t pred_n = 0;
    int32_t pred_ngroups = 0;
    uint64_t pred_generation = 0;
};

namespace vh {
namespace {

struct GemmArgs {
    const __bf16 *a;
    int64_t lda;
    const __bf16 *b;
    int64_t ldb;
    int32_t M, N, K;
    const float *bias;   // EPI_GELU, EPI_OUT
    float *zout;         // EPI_GELU (may be null): pre-activation
    int64_t ldz;
    __bf16 *hout;        // EPI_GELU, EPI_DGELU
    int64_t ldh;
    float *cout;         // EPI_OUT, EPI_STORE
    int64_t ldc;
    const float *zin;    // EPI_DGELU: the pre-activation of this layer's input
    int64_t ldzi;
    float *colsum;       // EPI_STORE (may be null): column sums of B
    __bf16 *hblk;        // EPI_GELU, EPI_DGELU (may be null): hout again, K-blocked (N columns)
};

enum { EPI_GELU = 0, EPI_OUT = 1, EPI_DGELU = 2, EPI_STORE = 3 };

// tanh(u) = 1 - 2 / (exp(2u) + 1), with the hardware exponential and
// reciprocal: absolute error a few 1e-7 (the activations are stored as BF16,
// 2^-8 relative), exactly +-1 where exp(2u) over- or underflows.
__device__ __forceinline__ float fast_tanh(float u)
{
    const float e = __expf(2.0f * u);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ __forceinline__ float gelu(float z)
{
    float u = kGeluC * (z + ((0.044715f * z) * z) * z);
    return (0.5f * z) * (1.0f + fast_tanh(u));
}

__device__ __forceinline__ float gelu_grad(float z)
{
    float u = kGeluC * (z + ((0.044715f * z) * z) * z);
    float t = fast_tanh(u);
    float du = kGeluC * (1.0f + ((3.0f * 0.044715f) * z) * z);
    return (0.5f * (1.0f + t)) + ((0.5f * z) * (1.0f - t * t)) * du;
}

// C (M x N) = A (M x K) . B (K x N); M, N multiples of 16, K of 32.
// A(m, k) = a[m * lda + k] (AT: a[k * lda + m]); B(k, n) = b[n * ldb + k]
// (BT: b[k * ldb + n]).
template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    const int lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tiles_n = g.N >> 4;
    if (tile >= (int64_t)(g.M >> 4) * tiles_n) return;  // whole wave
    const int64_t m0 = (tile / tiles_n) * 16;
    const int n0 = (int)(tile % tiles_n) * 16;
    const int r = lane & 15, kg = (lane >> 4) * 8;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float cs = 0.0f;
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        bf16x8 af, bfr;
        const int64_t k = k0 + kg;
        if (!AT) {
            af = *(const bf16x8 *)(g.a + (m0 + r) * g.lda + k);
        } else {
            for (int j = 0; j < 8; j++) af[j] = g.a[(k + j) * g.lda + m0 + r];
        }
        if (!BT) {
            bfr = *(const bf16x8 *)(g.b + (int64_t)(n0 + r) * g.ldb + k);
        } else {
            for (int j = 0; j < 8; j++) bfr[j] = g.b[(k + j) * g.ldb + n0 + r];
        }
        if (EPI == EPI_STORE && g.colsum) {
            for (int j = 0; j < 8; j++) cs += (float)bfr[j];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr, acc, 0, 0, 0);
    }
    const int col = n0 + r;
    const int64_t row0 = m0 + (lane >> 4) * 4;
    __bf16 hv[4];
    for (int i = 0; i < 4; i++) {
        const int64_t row = row0 + i;
        float v = acc[i];
        if (EPI == EPI_GELU) {
            float zz = v + g.bias[col];
            if (g.zout) g.zout[row * g.ldz + col] = zz;
            hv[i] = (__bf16)gelu(zz);
            g.hout[row * g.ldh + col] = hv[i];
        } else if (EPI == EPI_OUT) {
            g.cout[row * g.ldc + col] = v + g.bias[col];
        } else if (EPI == EPI_DGELU) {
            hv[i] = (__bf16)(v * gelu_grad(g.zin[row * g.ldzi + col]));
            g.hout[row * g.ldh + col] = hv[i];
        } else {
            g.cout[row * g.ldc + col] = v;
        }
    }
    if ((EPI == EPI_GELU || EPI == EPI_DGELU) && g.hblk) {
        // rows row0 .. row0 + 3 of one column: four consecutive slots
        __bf16 *dst = g.hblk + ((row0 >> 3) * g.N + col) * 8 + (row0 & 7);
        *(uint2 *)dst = *(const uint2 *)hv;
    }
    if (EPI == EPI_STORE && g.colsum) {
        cs += __shfl_xor(cs, 16, 64);
        cs += __shfl_xor(cs, 32, 64);
        if (m0 == 0 && lane < 16) g.colsum[col] = cs;
    }
}

// Output-layer gradient and loss: d = y - t over the three outputs,
// dy = 2 d / (3 n) as BF16 (padded columns 0), per-block sums of d^2.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ y, const float *__restrict__ t,
                                              int64_t n, __bf16 *__restrict__ dy, __bf16 *__restrict__ dyk,
                                              double *__restrict__ part)
{
    __shared__ double red[256];
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (row < n) {
        const float scale = 2.0f / (3.0f * (float)n);
        for (int c = 0; c < kOutPad; c++) {
            float d = 0.0f;
            if (c < kOut) {
                d = y[row * kOutPad + c] - t[row * kOut + c];
                acc += (double)d * (double)d;
            }
            const __bf16 g = (__bf16)(d * scale);
            dy[row * kOutPad + c] = g;
            if (dyk) dyk[((row >> 3) * kOutPad + c) * 8 + (row & 7)] = g;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The partial sums in block order: loss = sum / (3 n), at out[*index].
__global__ void k_loss_reduce(const double *__restrict__ part, int32_t nblk, int64_t n,
                              float *__restrict__ out, const int64_t *__restrict__ state)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int32_t b = threadIdx.x; b < nblk; b += 256) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[state ? state[4] : 0] = (float)(red[0] / (3.0 * (double)n));
}

struct AdamArgs {
    float *w, *m, *v;
    const float *g;
    int64_t n;
    float lr, beta1, beta2, eps;
    const int64_t *state;  // state[0]: steps taken before this one
    __bf16 *wb, *wtb;
    int64_t woff[kLayers], boff[kLayers], wtoff[kLayers];
    int32_t inp[kLayers], outp[kLayers];
};

// One Adam step (Kingma and Ba, 2015, Algorithm 1) of an FP32 parameter w
// with moments m, v and gradient gi; bc1, bc2 are the bias corrections
// 1 - beta^t.  Updates w, m, v in place.
__device__ __forceinline__ void adam_math(float &w, float &m, float &v, float gi, const AdamArgs &a, float bc1,
                                          float bc2)
{
    const float mi = a.beta1 * m + (1.0f - a.beta1) * gi;
    const float vi = a.beta2 * v + (1.0f - a.beta2) * (gi * gi);
    m = mi;
    v = vi;
    w = w - a.lr * (mi / bc1) / (sqrtf(vi / bc2) + a.eps);
}

__device__ __forceinline__ void adam_corrections(const AdamArgs &a, float &bc1, float &bc2, int32_t ahead = 0)
{
    const double t = (double)(a.state[0] + ahead + 1);
    bc1 = (float)(1.0 - pow((double)a.beta1, t));
    bc2 = (float)(1.0 - pow((double)a.beta2, t));
}

// Adam per FP32 parameter, then the BF16 copies of the weights in both
// layouts.
__global__ __launch_bounds__(256) void k_adam(AdamArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    float bc1, bc2;
    adam_corrections(a, bc1, bc2);
    float w = a.w[i], m = a.m[i], v = a.v[i];
    adam_math(w, m, v, a.g[i], a, bc1, bc2);
    a.m[i] = m;
    a.v[i] = v;
    a.w[i] = w;
    for (int l = 0; l < kLayers; l++) {
        const int64_t off = i - a.woff[l];
        if (off >= 0 && off < (int64_t)a.inp[l] * a.outp[l]) {
            const int64_t k = off / a.outp[l], o = off % a.outp[l];
            a.wb[a.wtoff[l] + off] = (__bf16)w;
            a.wtb[a.wtoff[l] + o * a.inp[l] + k] = (__bf16)w;
        }
    }
}

// The BF16 weight copies from the FP32 master (after a load).
__global__ __launch_bounds__(256) void k_weights(AdamArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const float w = a.w[i];
    for (int l = 0; l < kLayers; l++) {
        const int64_t off = i - a.woff[l];
        if (off >= 0 && off < (int64_t)a.inp[l] * a.outp[l]) {
            const int64_t k = off / a.outp[l], o = off % a.outp[l];
            a.wb[a.wtoff[l] + off] = (__bf16)w;
            a.wtb[a.wtoff[l] + o * a.inp[l] + k] = (__bf16)w;
        }
    }
}

// state: [0] Adam steps, [1] step within the epoch, [2] a, [3] c, [4] loss slot.
__global__ void k_advance(int64_t *state)
{
    state[0] += 1;
    state[1] += 1;
    state[4] += 1;
}

// One batch of `nb` rows: data row (a * (step * nb + i) + c) mod n, a bijection
// of [0, n) for a coprime to n.
__global__ __launch_bounds__(256) void k_batch(const __bf16 *__restrict__ X, const float *__restrict__ T,
                                               int64_t n, const int64_t *__restrict__ state, int32_t nb,
                                               float gain, float inv_scale, __bf16 *__restrict__ xb,
                                               __bf16 *__restrict__ xk, float *__restrict__ tb)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nb) return;
    const uint64_t a = (uint64_t)state[2], c = (uint64_t)state[3];
    const uint64_t pos = (uint64_t)state[1] * (uint64_t)nb + (uint64_t)i;
    // a, c < n < 2^31 and pos < n: the product fits in 64 bits
    const uint64_t j = (a * pos + c) % (uint64_t)n;
    const uint4 *src = (const uint4 *)(X + j * kInPad);
    uint4 *dst = (uint4 *)(xb + i * kInPad);
    for (int q = 0; q < kInPad * 2 / 16; q++) dst[q] = src[q];
    for (int c = 0; c < kInPad; c++) xk[((i >> 3) * kInPad + c) * 8 + (i & 7)] = X[j * kInPad + c];
    // the target: (bounced - gain * direct) / scale
    for (int k = 0; k < kOut; k++)
        tb[i * kOut + k] = (T[j * kLight + kOut + k] - gain * T[j * kLight + k]) * inv_scale;
}

struct FeatArgs {
    const int32_t *rows;  // patches of each row (null: row = patch)
    int64_t n;
    const float4 *origin_area, *normal_pd, *e0, *total;
    float lo[3], ext[3];
    float freq[24];
    float length, floor;
    __bf16 *x;       // n x 32
    float *xf;       // n x 25 (may be null)
    float *t;        // n x 6: direct then bounced light (may be null)
};

// The 25 inputs of a patch: 2 (p - lo) / ext - 1, n, log(max(e0, 0) + floor),
// and sin, cos of 2 pi (F p) / length for the 8 rows of F; for the data set,
// the patch's direct and bounced light.
__global__ __launch_bounds__(256) void k_features(FeatArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int32_t p = a.rows ? a.rows[i] : (int32_t)i;
    const float4 o = a.origin_area[p], nn = a.normal_pd[p], e = a.e0[p];
    float f[kInPad];
    const float pp[3] = {o.x, o.y, o.z};
    for (int k = 0; k < 3; k++) f[k] = 2.0f * (pp[k] - a.lo[k]) / a.ext[k] - 1.0f;
    f[3] = nn.x;
    f[4] = nn.y;
    f[5] = nn.z;
    f[6] = logf(fmaxf(e.x, 0.0f) + a.floor);
    f[7] = logf(fmaxf(e.y, 0.0f) + a.floor);
    f[8] = logf(fmaxf(e.z, 0.0f) + a.floor);
    for (int q = 0; q < kFourier; q++) {
        const float d = (a.freq[3 * q] * o.x + a.freq[3 * q + 1] * o.y) + a.freq[3 * q + 2] * o.z;
        const float ang = 6.283185307179586f * (d / a.length);
        f[9 + q] = sinf(ang);
        f[9 + kFourier + q] = cosf(ang);
    }
    for (int k = kIn; k < kInPad; k++) f[k] = 0.0f;
    for (int k = 0; k < kInPad; k++) a.x[i * kInPad + k] = (__bf16)f[k];
    if (a.xf)
        for (int k = 0; k < kIn; k++) a.xf[i * kIn + k] = f[k];
    if (a.t) {
        const float4 tt = a.total[p];
        float *r = a.t + i * kLight;
        r[0] = e.x;
        r[1] = e.y;
        r[2] = e.z;
        r[3] = tt.x;
        r[4] = tt.y;
        r[5] = tt.z;
    }
}

// Host rows (n x 25 FP32) to BF16 operands (n x 32), padded rows zero.
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ xf, int64_t n, int64_t npad,
                                              __bf16 *__restrict__ x, __bf16 *__restrict__ xk)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= npad) return;
    for (int k = 0; k < kInPad; k++) {
        const __bf16 v = (__bf16)((i < n && k < kIn) ? xf[i * kIn + k] : 0.0f);
        x[i * kInPad + k] = v;
        if (xk) xk[((i >> 3) * kInPad + k) * 8 + (i & 7)] = v;
    }
}

// Weight gradient, split over K: dW (M x N) = A^T B over rows [k0, k1) of a
// split, A (K x M) and B (K x N) K-blocked.  One wave per (16 x 16 tile,
// split); lane l holds column (l & 15) of A^T and of B at rows 8 (l >> 4) .. + 7
// of each step of 32, one 16-byte load each.  Partial tiles and the partial
// column sums of B (the bias gradient) go to the workspace, summed in split
// order by k_dw_reduce.
__global__ __launch_bounds__(256) void k_gemm_dw(const __bf16 *__restrict__ a, int32_t M,
                                                 const __bf16 *__restrict__ b, int32_t N, int32_t K,
                                                 int32_t kc, int32_t splits, float *__restrict__ part,
                                                 float *__restrict__ cpart)
{
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tiles_n = N >> 4;
    const int64_t tiles = (int64_t)(M >> 4) * tiles_n;
    if (w >= tiles * splits) return;  // whole wave
    const int split = (int)(w / tiles);
    const int64_t tile = w % tiles;
    const int m0 = (int)(tile / tiles_n) * 16, n0 = (int)(tile % tiles_n) * 16;
    const int r = lane & 15, kg = lane >> 4;
    const int32_t ka = split * kc, kb = min(K, ka + kc);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float cs = 0.0f;
    for (int32_t k0 = ka; k0 < kb; k0 += 32) {
        const int64_t blk = (k0 >> 3) + kg;
        const bf16x8 af = *(const bf16x8 *)(a + (blk * M + m0 + r) * 8);
        const bf16x8 bfr = *(const bf16x8 *)(b + (blk * N + n0 + r) * 8);
        for (int j = 0; j < 8; j++) cs += (float)bfr[j];
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr, acc, 0, 0, 0);
    }
    float *out = part + (int64_t)split * M * N;
    const int row0 = m0 + kg * 4;
    for (int i = 0; i < 4; i++) out[(int64_t)(row0 + i) * N + n0 + r] = acc[i];
    cs += __shfl_xor(cs, 16, 64);
    cs += __shfl_xor(cs, 32, 64);
    if (m0 == 0 && lane < 16) cpart[(int64_t)split * N + n0 + r] = cs;
}

// dW and db from the split partials, in split order.
__global__ __launch_bounds__(256) void k_dw_reduce(const float *__restrict__ part,
                                                   const float *__restrict__ cpart, int32_t M, int32_t N,
                                                   int32_t splits, float *__restrict__ dw,
                                                   float *__restrict__ db)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t mn = (int64_t)M * N;
    if (i < mn) {
        float s = 0.0f;
        for (int k = 0; k < splits; k++) s += part[(int64_t)k * mn + i];
        dw[i] = s;
    } else if (i < mn + N) {
        const int64_t c = i - mn;
        float s = 0.0f;
        for (int k = 0; k < splits; k++) s += cpart[(int64_t)k * N + c];
        db[c] = s;
    }
}

__global__ __launch_bounds__(256) void k_zero4(float4 *__restrict__ v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The predicted bounced light of one cluster's rows (a block per cluster):
// q = max(gain * e0 + scale * y, 0), then per channel rescaled so that the
// cluster's sum is the prior's, gain * sum(max(e0, 0)): the network places
// the light within the cluster and the prior sets how much there is.  A
// cluster whose q sums to zero takes the prior itself.  Sums run in a fixed
// order.
__global__ __launch_bounds__(256) void k_cluster_start(const float *__restrict__ y,
                                                       const int32_t *__restrict__ rows,
                                                       const int32_t *__restrict__ groups,
                                                       const float4 *__restrict__ e0, float gain, float scale,
                                                       float4 *__restrict__ start)
{
    __shared__ double red[6][256];
    const int32_t a = groups[blockIdx.x], b = groups[blockIdx.x + 1];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t i = a + (int32_t)threadIdx.x; i < b; i += 256) {
        const float *r = y + (int64_t)i * kOutPad;
        const float4 e = e0[rows[i]];
        const float ev[3] = {fmaxf(e.x, 0.0f), fmaxf(e.y, 0.0f), fmaxf(e.z, 0.0f)};
        for (int k = 0; k < 3; k++) {
            acc[k] += (double)fmaxf(gain * ev[k] + scale * r[k], 0.0f);
            acc[3 + k] += (double)(gain * ev[k]);
        }
    }
    for (int k = 0; k < 6; k++) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    float f[3];
    bool use_prior[3];
    for (int k = 0; k < 3; k++) {
        use_prior[k] = !(red[k][0] > 0.0);
        f[k] = use_prior[k] ? 0.0f : (float)(red[3 + k][0] / red[k][0]);
    }
    for (int32_t i = a + (int32_t)threadIdx.x; i < b; i += 256) {
        const float *r = y + (int64_t)i * kOutPad;
        const int32_t p = rows[i];
        const float4 e = e0[p];
        const float ev[3] = {fmaxf(e.x, 0.0f), fmaxf(e.y, 0.0f), fmaxf(e.z, 0.0f)};
        float o[3];
        for (int k = 0; k < 3; k++) {
            const float q = fmaxf(gain * ev[k] + scale * r[k], 0.0f);
            o[k] = use_prior[k] ? gain * ev[k] : q * f[k];
        }
        start[p] = make_float4(o[0], o[1], o[2], 0.0f);
    }
}

// ---- The fused training step: two launches per step ----------------------
//
// k_step_rows: a block per 16 batch rows (eight waves; 4,096 rows fill the
// 256 CUs).  It draws its rows (the affine permutation of k_batch), runs the
// forward pass through all four layers with the activations and the hidden
// layers' FP32 pre-activations in LDS, forms the loss and the output
// gradient, and runs the backward data pass down to the first hidden layer;
// each layer's input and output gradient go to HBM K-blocked for the weight
// gradients.  Wave w owns column tile w (one 16 x 16 MFMA tile) of every
// 128-wide layer; waves 0 and 1 own the two tiles of the 32-wide output
// layer.  The epilogues (GELU, GELU') are most of the work, four values a
// lane per layer.
//
// k_step_weights: a block per 16 x 16 tile of the four weight gradients.  Its
// waves share the row splits of k_gemm_dw (each split summed by one wave, in
// the same K order); wave 0 sums the splits in split order (as k_dw_reduce),
// so the gradient is bit-identical to the per-layer kernels', applies Adam to
// the tile's weights (and the layer's bias gradient, for the tiles of row 0),
// and writes their BF16 copies.  The last block to finish sums the blocks'
// loss partials and advances the step counters.
//
// Every MFMA sees the operands, the K order and the epilogue arithmetic of the
// per-layer kernels (k_batch, k_gemm, k_loss, k_gemm_dw, k_dw_reduce, k_adam),
// so the two paths train to the same bits; only the loss value, a sum of
// squares reported per step, is summed in another order.

constexpr int kFRows = 16;               // batch rows per k_step_rows block
constexpr int kFWaves = kHidden / 16;    // waves per k_step_rows block: a column tile each
constexpr int kActStride = kHidden + 8;  // LDS row stride of an activation (BF16)
constexpr int kZStride = kHidden + 4;    // LDS row stride of a pre-activation (FP32)

struct StepRowsArgs {
    const __bf16 *X;
    const float *T;
    int64_t n;
    int32_t nb;
    float gain, inv_scale;
    const int64_t *state;
    int32_t ahead;  // this step's place after the counters' step
    const __bf16 *wb, *wtb;
    const float *master;
    int64_t boff[kLayers], wtoff[kLayers];
    __bf16 *hk[kLayers], *dzk[kLayers];
    double *loss_part;  // this step's per-block partials
};

// The B fragments of a wave's NC column tiles (ct0 ..) over all K / 32 steps,
// B(k, n) = b[n * K + k]: loaded ahead of the layer that uses them, so their
// global-memory latency overlaps the layer before.
template <int K, int NC>
struct BFrags {
    bf16x8 f[K / 32][NC];
};

template <int K, int NC>
__device__ __forceinline__ void load_b(BFrags<K, NC> &bf, const __bf16 *__restrict__ b, int ct0)
{
    const int lane = threadIdx.x & 63, r = lane & 15, kg = (lane >> 4) * 8;
#pragma unroll
    for (int s = 0; s < K / 32; s++)
#pragma unroll
        for (int c = 0; c < NC; c++) bf.f[s][c] = *(const bf16x8 *)(b + (int64_t)((ct0 + c) * 16 + r) * K + s * 32 + kg);
}

// 16 x 16 MFMA tiles of C = A . B for the wave's NR x NC tiles (row tiles
// rt0 .. rt0 + NR - 1): A from LDS (row-major, kActStride), K in steps of 32
// from 0 up, one MFMA per tile and step.
template <int NR, int NC, int K>
__device__ __forceinline__ void tiles_mma(const __bf16 *A, const BFrags<K, NC> &bf, int rt0,
                                          f32x4 (&acc)[NR][NC])
{
    const int lane = threadIdx.x & 63, r = lane & 15, kg = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < NR; i++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[i][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < K / 32; s++) {
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const bf16x8 af = *(const bf16x8 *)(A + ((rt0 + i) * 16 + r) * kActStride + s * 32 + kg);
#pragma unroll
            for (int c = 0; c < NC; c++)
                acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf.f[s][c], acc[i][c], 0, 0, 0);
        }
    }
}

// Rows i .. i + 3 (batch rows, i a multiple of 4) of column `col` into the
// K-blocked copy of an ncol-wide matrix: four consecutive slots.
__device__ __forceinline__ void store_blocked(__bf16 *dst, int64_t i, int ncol, int col, const __bf16 (&v)[4])
{
    *(uint2 *)(dst + ((i >> 3) * ncol + col) * 8 + (i & 7)) = *(const uint2 *)v;
}

// Hidden layer L forward for the wave's tile (column tile w): z into zl
// (LDS), GELU(z) into dst and the K-blocked copy.
template <int L>
__device__ __forceinline__ void step_forward(const StepRowsArgs &a, const BFrags<kInP[L], 1> &bf,
                                             const __bf16 *src, __bf16 *dst, float *zl, int64_t r0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, rq = (lane >> 4) * 4;
    f32x4 acc[1][1];
    tiles_mma<1, 1, kInP[L]>(src, bf, 0, acc);
    const int col = wave * 16 + r;
    const float bias = a.master[a.boff[L] + col];
    __bf16 hv[4];
    for (int i = 0; i < 4; i++) {
        const int row = rq + i;
        const float zz = acc[0][0][i] + bias;
        zl[row * kZStride + col] = zz;
        hv[i] = (__bf16)gelu(zz);
        dst[row * kActStride + col] = hv[i];
    }
    store_blocked(a.hk[L + 1], r0 + rq, kHidden, col, hv);
}

// Backward data through layer L: dZ_{L-1} = (dZ_L W_L^T) * GELU'(z_{L-1}) for
// the wave's tile, into dst (unless null) and the K-blocked copy.
template <int L>
__device__ __forceinline__ void step_backward(const StepRowsArgs &a, const BFrags<kOutP[L], 1> &bf,
                                              const __bf16 *src, __bf16 *dst, const float *zl, int64_t r0)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, rq = (lane >> 4) * 4;
    f32x4 acc[1][1];
    tiles_mma<1, 1, kOutP[L]>(src, bf, 0, acc);
    const int col = wave * 16 + r;
    __bf16 hv[4];
    for (int i = 0; i < 4; i++) {
        const int row = rq + i;
        hv[i] = (__bf16)(acc[0][0][i] * gelu_grad(zl[row * kZStride + col]));
        if (dst) dst[row * kActStride + col] = hv[i];
    }
    store_blocked(a.dzk[L - 1], r0 + rq, kHidden, col, hv);
}

__global__ __launch_bounds__(64 * kFWaves) void k_step_rows(StepRowsArgs a)
{
    __shared__ __attribute__((aligned(16))) __bf16 act[2][kFRows * kActStride];
    __shared__ float zs[kLayers - 1][kFRows * kZStride];
    __shared__ float tgt[kFRows * kOut];
    __shared__ double lred[2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, rq = (lane >> 4) * 4;
    const int64_t r0 = (int64_t)blockIdx.x * kFRows;
    const int oc = wave & 1;  // waves 0 and 1: the output layer's column tiles

    // each layer's weight fragments are loaded one layer ahead
    BFrags<kInPad, 1> f0;
    load_b(f0, a.wtb + a.wtoff[0], wave);
    BFrags<kHidden, 1> f1;
    load_b(f1, a.wtb + a.wtoff[1], wave);

    // the batch rows (k_batch): four threads per row, 16 bytes each
    if (tid < kFRows * 4) {
        const int i = tid >> 2, q = tid & 3;
        const uint64_t am = (uint64_t)a.state[2], c = (uint64_t)a.state[3];
        const uint64_t pos = (uint64_t)(a.state[1] + a.ahead) * (uint64_t)a.nb + (uint64_t)(r0 + i);
        const uint64_t j = (am * pos + c) % (uint64_t)a.n;
        const uint4 v = ((const uint4 *)(a.X + j * kInPad))[q];
        *(uint4 *)&act[0][i * kActStride + q * 8] = v;
        const __bf16 *e = (const __bf16 *)&v;
        const int64_t row = r0 + i;
        for (int k = 0; k < 8; k++) a.hk[0][((row >> 3) * kInPad + q * 8 + k) * 8 + (row & 7)] = e[k];
        if (q == 0)
            for (int k = 0; k < kOut; k++)
                tgt[i * kOut + k] = (a.T[j * kLight + kOut + k] - a.gain * a.T[j * kLight + k]) * a.inv_scale;
    }
    __syncthreads();

    // forward through the hidden layers, the pre-activations kept
    step_forward<0>(a, f0, act[0], act[1], zs[0], r0);
    BFrags<kHidden, 1> f2;
    load_b(f2, a.wtb + a.wtoff[2], wave);
    __syncthreads();
    step_forward<1>(a, f1, act[1], act[0], zs[1], r0);
    BFrags<kHidden, 1> f3;
    load_b(f3, a.wtb + a.wtoff[3], oc);
    __syncthreads();
    step_forward<2>(a, f2, act[0], act[1], zs[2], r0);
    BFrags<kOutPad, 1> g3;
    load_b(g3, a.wb + a.wtoff[3], wave);
    __syncthreads();

    // output layer (waves 0 and 1), loss and its gradient (k_loss): dy into
    // act[0]
    if (wave < 2) {
        double lacc = 0.0;
        f32x4 acc[1][1];
        tiles_mma<1, 1, kHidden>(act[1], f3, 0, acc);
        const float scale = 2.0f / (3.0f * (float)a.nb);
        const int col = oc * 16 + r;
        const float bias = a.master[a.boff[kLayers - 1] + col];
        __bf16 gv[4];
        for (int i = 0; i < 4; i++) {
            const int row = rq + i;
            float d = 0.0f;
            if (col < kOut) {
                d = (acc[0][0][i] + bias) - tgt[row * kOut + col];
                lacc += (double)d * (double)d;
            }
            gv[i] = (__bf16)(d * scale);
            act[0][row * kActStride + col] = gv[i];
        }
        store_blocked(a.dzk[kLayers - 1], r0 + rq, kOutPad, col, gv);
        for (int off = 32; off > 0; off >>= 1) lacc += __shfl_xor(lacc, off, 64);
        if (lane == 0) lred[wave] = lacc;
    }
    BFrags<kHidden, 1> g2;
    load_b(g2, a.wb + a.wtoff[2], wave);
    __syncthreads();
    if (tid == 0) a.loss_part[blockIdx.x] = lred[0] + lred[1];

    // backward data down to the first hidden layer
    step_backward<3>(a, g3, act[0], act[1], zs[2], r0);
    BFrags<kHidden, 1> g1;
    load_b(g1, a.wb + a.wtoff[1], wave);
    __syncthreads();
    step_backward<2>(a, g2, act[1], act[0], zs[1], r0);
    __syncthreads();
    step_backward<1>(a, g1, act[0], nullptr, zs[0], r0);
}

struct StepWeightsArgs {
    const __bf16 *hk[kLayers], *dzk[kLayers];
    int32_t K, kc, splits;
    int32_t ahead;  // this step's place after the counters' step
    AdamArgs adam;
};

constexpr int kWWaves = 16;  // waves per k_step_weights block
constexpr int kWTiles = 2 * 16 + 2 * 64;  // 16 x 16 tiles of the four weight gradients

// The tile of block b, XCD-aware: blocks go round-robin over the eight XCDs
// (b % 8), and the 20 tiles an XCD gets share their operand slices -- per
// 128 x 128 layer a 2 x 4 patch of tiles (two A and four B column slices
// through its L2), per 32-wide layer the tiles of one B (or A) slice.
__device__ __forceinline__ void weight_tile(int b, int &l, int &mt, int &nt)
{
    const int x = b & 7, j = b >> 3;
    if (j < 16) {
        l = 1 + (j >> 3);
        const int q = j & 7;
        mt = 2 * (x & 3) + (q & 1);
        nt = 4 * (x >> 2) + (q >> 1);
    } else if (j < 18) {
        l = 0;
        mt = j - 16;
        nt = x;
    } else {
        l = kLayers - 1;
        mt = x;
        nt = j - 18;
    }
}

__global__ __launch_bounds__(64 * kWWaves) void k_step_weights(StepWeightsArgs a)
{
    __shared__ f32x4 part[kMaxSplits][64];
    __shared__ float cpart[kMaxSplits][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int l, mt, nt;
    weight_tile(blockIdx.x, l, mt, nt);
    const int M = kInP[l], N = kOutP[l];
    const int m0 = mt * 16, n0 = nt * 16;
    const int r = lane & 15, kg = lane >> 4;
    const int col = n0 + r;
    const AdamArgs &ad = a.adam;
    // wave 0's parameters and their Adam state, loaded ahead of the sums
    float pw[4], pm[4], pv[4], bw = 0.0f, bm = 0.0f, bv = 0.0f;
    const bool has_bias = m0 == 0 && lane < 16;
    if (wave == 0) {
        for (int i = 0; i < 4; i++) {
            const int64_t p = ad.woff[l] + (int64_t)(m0 + kg * 4 + i) * N + col;
            pw[i] = ad.w[p];
            pm[i] = ad.m[p];
            pv[i] = ad.v[p];
        }
        if (has_bias) {
            bw = ad.w[ad.boff[l] + col];
            bm = ad.m[ad.boff[l] + col];
            bv = ad.v[ad.boff[l] + col];
        }
    }
    // the splits of k_gemm_dw, one wave each, eight steps of 32 rows loaded
    // at a time (a step past the split's end reloads its last step and is
    // skipped)
    const __bf16 *A = a.hk[l], *B = a.dzk[l];
    for (int sp = wave; sp < a.splits; sp += kWWaves) {
        const int32_t ka = sp * a.kc, kb = min(a.K, ka + a.kc);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        float cs = 0.0f;
        for (int32_t kbase = ka; kbase < kb; kbase += 256) {
            bf16x8 af[8], bfr[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int32_t k0 = min(kbase + 32 * q, kb - 32);
                const int64_t blk = (k0 >> 3) + kg;
                af[q] = *(const bf16x8 *)(A + (blk * M + m0 + r) * 8);
                bfr[q] = *(const bf16x8 *)(B + (blk * N + n0 + r) * 8);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (kbase + 32 * q < kb) {
                    for (int j = 0; j < 8; j++) cs += (float)bfr[q][j];
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[q], bfr[q], acc, 0, 0, 0);
                }
            }
        }
        cs += __shfl_xor(cs, 16, 64);
        cs += __shfl_xor(cs, 32, 64);
        part[sp][lane] = acc;
        cpart[sp][lane] = cs;
    }
    __syncthreads();
    if (wave != 0) return;
    // the splits in split order (k_dw_reduce), then Adam (k_adam)
    f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
    float gb = 0.0f;
    for (int sp = 0; sp < a.splits; sp++) {
        const f32x4 p = part[sp][lane];
        for (int i = 0; i < 4; i++) g[i] += p[i];
        gb += cpart[sp][lane];
    }
    float bc1, bc2;
    adam_corrections(ad, bc1, bc2, a.ahead);
    for (int i = 0; i < 4; i++) {
        const int row = m0 + kg * 4 + i;
        const int64_t off = (int64_t)row * N + col;
        adam_math(pw[i], pm[i], pv[i], g[i], ad, bc1, bc2);
        ad.w[ad.woff[l] + off] = pw[i];
        ad.m[ad.woff[l] + off] = pm[i];
        ad.v[ad.woff[l] + off] = pv[i];
        ad.wb[ad.wtoff[l] + off] = (__bf16)pw[i];
        ad.wtb[ad.wtoff[l] + (int64_t)col * M + row] = (__bf16)pw[i];
    }
    if (has_bias) {
        adam_math(bw, bm, bv, gb, ad, bc1, bc2);
        ad.w[ad.boff[l] + col] = bw;
        ad.m[ad.boff[l] + col] = bm;
        ad.v[ad.boff[l] + col] = bv;
    }
}

// After a run of `steps` fused steps: each step's loss from its blocks'
// partials (a wave per step: lane-strided sums, then a fixed shuffle tree),
// then the counters advance past the run.
__global__ __launch_bounds__(256) void k_step_tail(const double *__restrict__ part, int32_t nparts,
                                                   int32_t steps, int32_t nb, float *__restrict__ losses,
                                                   int64_t *__restrict__ state)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t slot = state[4];
    for (int k = wave; k < steps; k += 4) {
        double sum = 0.0;
        for (int b = lane; b < nparts; b += 64) sum += part[(int64_t)k * nparts + b];
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        if (lane == 0) losses[slot + k] = (float)(sum / (3.0 * (double)nb));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        state[0] += steps;
        state[1] += steps;
        state[4] += steps;
    }
}

inline uint32_t blocks(int64_t n, int per = 256) { return (uint32_t)((n + per - 1) / per); }

// A device buffer that lives for one call.
struct LocalBuf : DevBuf {
    LocalBuf() = default;
    LocalBuf(const LocalBuf &) = delete;
    LocalBuf &operator=(const LocalBuf &) = delete;
    ~LocalBuf() { release(); }
};

template <bool AT, bool BT, int EPI>
void gemm(const GemmArgs &g, hipStream_t s)
{
    int64_t tiles = (int64_t)(g.M / 16) * (g.N / 16);
    hipLaunchKernelGGL((k_gemm<AT, BT, EPI>), dim3(blocks(tiles, 4)), dim3(256), 0, s, g);
}

AdamArgs adam_args(MlpData *D)
{
    AdamArgs a{};
    a.w = (float *)D->master.ptr;
    a.m = (float *)D->m.ptr;
    a.v = (float *)D->v.ptr;
    a.g = (const float *)D->grad.ptr;
    a.n = D->L.nparam;
    a.wb = (__bf16 *)D->wb.ptr;
    a.wtb = (__bf16 *)D->wtb.ptr;
    for (int l = 0; l < kLayers; l++) {
        a.woff[l] = D->L.w[l];
        a.boff[l] = D->L.b[l];
        a.wtoff[l] = D->L.wt[l];
        a.inp[l] = kInP[l];
        a.outp[l] = kOutP[l];
    }
    return a;
}

int ensure_acts(MlpData *D, int64_t rows)
{
    if (rows <= D->act_rows) return VH_OK;
    int rc;
    if ((rc = D->h[0].ensure((size_t)rows * kInPad * 2, "network inputs"))) return rc;
    for (int l = 1; l < 4; l++)
        if ((rc = D->h[l].ensure((size_t)rows * kHidden * 2, "activations"))) return rc;
    for (int l = 0; l < 3; l++) {
        if ((rc = D->z[l].ensure((size_t)rows * kHidden * 4, "pre-activations"))) return rc;
        if ((rc = D->dz[l].ensure((size_t)rows * kHidden * 2, "activation gradients"))) return rc;
    }
    for (int l = 0; l < 4; l++)
        if ((rc = D->hk[l].ensure((size_t)rows * kInP[l] * 2, "blocked activations")) ||
            (rc = D->dzk[l].ensure((size_t)rows * kOutP[l] * 2, "blocked gradients")))
            return rc;
    if ((rc = D->dw_part.ensure((size_t)kMaxSplits * kHidden * kHidden * 4, "weight-gradient partials")) ||
        (rc = D->cs_part.ensure((size_t)kMaxSplits * kHidden * 4, "bias-gradient partials")))
        return rc;
    if ((rc = D->y.ensure((size_t)rows * kOutPad * 4, "outputs")) ||
        (rc = D->dy.ensure((size_t)rows * kOutPad * 2, "output gradients")) ||
        (rc = D->tf.ensure((size_t)rows * kOut * 4, "targets")) ||
        (rc = D->loss_part.ensure((size_t)blocks(rows) * 8 + 8, "loss partials")))
        return rc;
    D->act_rows = rows;
    return VH_OK;
}

// Forward over `rows` rows of h[0]; keeps the pre-activations when `keep`.
void forward(MlpData *D, int64_t rows, bool keep, hipStream_t s)
{
    const __bf16 *wtb = (const __bf16 *)D->wtb.ptr;
    const float *mp = (const float *)D->master.ptr;
    for (int l = 0; l < kLayers; l++) {
        GemmArgs g{};
        g.a = (const __bf16 *)D->h[l].ptr;
        g.lda = kInP[l];
        g.b = wtb + D->L.wt[l];
        g.ldb = kInP[l];
        g.M = (int32_t)rows;
        g.N = kOutP[l];
        g.K = kInP[l];
        g.bias = mp + D->L.b[l];
        if (l < kLayers - 1) {
            g.zout = keep ? (float *)D->z[l].ptr : nullptr;
            g.ldz = kHidden;
            g.hout = (__bf16 *)D->h[l + 1].ptr;
            g.ldh = kHidden;
            g.hblk = keep ? (__bf16 *)D->hk[l + 1].ptr : nullptr;
            gemm<false, false, EPI_GELU>(g, s);
        } else {
            g.cout = (float *)D->y.ptr;
            g.ldc = kOutPad;
            gemm<false, false, EPI_OUT>(g, s);
        }
    }
}

// Backward from dy over `rows` rows into the padded FP32 gradient.
void backward(MlpData *D, int64_t rows, hipStream_t s)
{
    const __bf16 *wb = (const __bf16 *)D->wb.ptr;
    float *grad = (float *)D->grad.ptr;
    for (int l = kLayers - 1; l >= 0; l--) {
        const __bf16 *dz = l == kLayers - 1 ? (const __bf16 *)D->dy.ptr : (const __bf16 *)D->dz[l].ptr;
        const int64_t ldd = kOutP[l];
        {
            // dW = H^T dZ and db = column sums of dZ, split over the rows
            const int32_t M = kInP[l], N = kOutP[l], K = (int32_t)rows;
            int32_t splits = K / 256;
            splits = splits < 1 ? 1 : (splits > kMaxSplits ? kMaxSplits : splits);
            const int32_t kc = ((K + splits - 1) / splits + 31) / 32 * 32;
            splits = (K + kc - 1) / kc;
            const int64_t waves = (int64_t)(M / 16) * (N / 16) * splits;
            hipLaunchKernelGGL(k_gemm_dw, dim3(blocks(waves, 4)), dim3(256), 0, s,
                               (const __bf16 *)D->hk[l].ptr, M, (const __bf16 *)D->dzk[l].ptr, N, K, kc,
                               splits, (float *)D->dw_part.ptr, (float *)D->cs_part.ptr);
            hipLaunchKernelGGL(k_dw_reduce, dim3(blocks((int64_t)M * N + N)), dim3(256), 0, s,
                               (const float *)D->dw_part.ptr, (const float *)D->cs_part.ptr, M, N, splits,
                               grad + D->L.w[l], grad + D->L.b[l]);
        }
        if (l == 0) break;
        GemmArgs d{};  // dH = dZ W^T, then dZ of the layer below = dH * GELU'(z)
        d.a = dz;
        d.lda = ldd;
        d.b = wb + D->L.wt[l];
        d.ldb = kOutP[l];
        d.M = (int32_t)rows;
        d.N = kInP[l];
        d.K = kOutP[l];
        d.zin = (const float *)D->z[l - 1].ptr;
        d.ldzi = kHidden;
        d.hout = (__bf16 *)D->dz[l - 1].ptr;
        d.ldh = kHidden;
        d.hblk = (__bf16 *)D->dzk[l - 1].ptr;
        gemm<false, false, EPI_DGELU>(d, s);
    }
}

void loss(MlpData *D, const float *t, int64_t rows, float *out, const int64_t *state, hipStream_t s)
{
    uint32_t nb = blocks(rows);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, (const float *)D->y.ptr, t, rows,
                       (__bf16 *)D->dy.ptr, (__bf16 *)D->dzk[kLayers - 1].ptr, (double *)D->loss_part.ptr);
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(256), 0, s, (const double *)D->loss_part.ptr,
                       (int32_t)nb, rows, out, state);
}

int check_mlp(vh_ctx *c)
{
    if (!c->mlp || !c->mlp->ready) {
        set_error("no network: call vh_mlp_init first");
        return VH_ESTATE;
    }
    return VH_OK;
}

FeatArgs feat_args(MlpData *D, const BounceView &bv)
{
    FeatArgs f{};
    f.origin_area = bv.origin_area;
    f.normal_pd = bv.normal_pd;
    for (int k = 0; k < 3; k++) {
        f.lo[k] = bv.lo[k];
        float e = bv.hi[k] - bv.lo[k];
        f.ext[k] = e > 0.0f ? e : 1.0f;
    }
    for (int k = 0; k < 24; k++) f.freq[k] = D->freq[k];
    f.length = D->length;
    f.floor = D->floor;
    return f;
}

// Leaf patches that are not sky, as rows (host list -> device).
// Leaf patches that are not sky, as rows.  With `groups`, the rows are
// ordered by cluster (then patch) and `groups` receives the clusters' row
// ranges (ngroups + 1 offsets).
int leaf_rows(vh_ctx *c, const BounceView &bv, DevBuf &out, int64_t *n, hipStream_t s,
              DevBuf *groups = nullptr, int32_t *ngroups = nullptr)
{
    std::vector<int32_t> leaves(bv.nleaves), cl;
    std::vector<uint8_t> sky(bv.npatch);
    VH_HIP(hipMemcpyAsync(leaves.data(), bv.leaves, leaves.size() * 4, hipMemcpyDeviceToHost, s));
    VH_HIP(hipMemcpyAsync(sky.data(), bv.sky, sky.size(), hipMemcpyDeviceToHost, s));
    if (groups) {
        cl.resize(bv.npatch);
        VH_HIP(hipMemcpyAsync(cl.data(), bv.cluster, cl.size() * 4, hipMemcpyDeviceToHost, s));
    }
    VH_WAIT(c, s);
    std::vector<int32_t> rows;
    rows.reserve(leaves.size());
    for (int32_t p : leaves)
        if (!sky[p]) rows.push_back(p);
    std::vector<int32_t> gp(1, 0);
    if (groups) {
        // stable by cluster: a counting sort when the cluster range is compact
        int32_t lo = rows.empty() ? 0 : cl[rows[0]], hi = lo;
        for (int32_t p : rows) {
            lo = std::min(lo, cl[p]);
            hi = std::max(hi, cl[p]);
        }
        const int64_t span = (int64_t)hi - lo + 1;
        if (!rows.empty() && span <= 4 * (int64_t)rows.size() + 1024) {
            std::vector<int32_t> start((size_t)span + 1, 0), sorted(rows.size());
            for (int32_t p : rows) start[(size_t)(cl[p] - lo) + 1]++;
            for (int64_t k = 0; k < span; k++) start[k + 1] += start[k];
            for (int32_t p : rows) sorted[start[(size_t)(cl[p] - lo)]++] = p;
            rows.swap(sorted);
        } else {
            std::stable_sort(rows.begin(), rows.end(), [&](int32_t a, int32_t b) { return cl[a] < cl[b]; });
        }
        for (size_t i = 1; i <= rows.size(); i++)
            if (i == rows.size() || cl[rows[i]] != cl[rows[i - 1]]) gp.push_back((int32_t)i);
        if (rows.empty()) gp.assign(1, 0);
        *ngroups = (int32_t)gp.size() - 1;
    }
    *n = (int64_t)rows.size();
    int rc = out.ensure(rows.size() * 4 + 16, "patch rows");
    if (rc) return rc;
    if (!rows.empty()) VH_HIP(hipMemcpyAsync(out.ptr, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, s));
    if (groups) {
        if ((rc = groups->ensure(gp.size() * 4 + 16, "cluster rows"))) return rc;
        VH_HIP(hipMemcpyAsync(groups->ptr, gp.data(), gp.size() * 4, hipMemcpyHostToDevice, s));
    }
    VH_WAIT(c, s);
    return VH_OK;
}

}  // namespace

void mlp_release(vh_ctx *c)
{
    MlpData *D = c->mlp;
    if (!D) return;
    for (DevBuf *b : {&D->master, &D->m, &D->v, &D->grad, &D->wb, &D->wtb, &D->X, &D->T,
                      &D->y, &D->dy, &D->xf, &D->tf, &D->loss_part, &D->losses, &D->state, &D->loss1,
                      &D->pred_rows, &D->pred_groups, &D->dw_part, &D->cs_part})
        b->release();
    for (auto &b : D->h) b.release();
    for (auto &b : D->hk) b.release();
    for (auto &b : D->dzk) b.release();
    for (auto &b : D->z) b.release();
    for (auto &b : D->dz) b.release();
    delete D;
    c->mlp = nullptr;
}

}  // namespace vh

using namespace vh;

namespace {

// Packed (36,739) <-> padded parameter layouts.
void pad_params(const Layout &L, const float *in, std::vector<float> &out)
{
    out.assign((size_t)L.nparam, 0.0f);
    int64_t o = 0;
    for (int l = 0; l < kLayers; l++) {
        for (int k = 0; k < kInDim[l]; k++)
            for (int n = 0; n < kOutDim[l]; n++) out[L.w[l] + (int64_t)k * kOutP[l] + n] = in[o++];
        for (int n = 0; n < kOutDim[l]; n++) out[L.b[l] + n] = in[o++];
    }
}

void unpad_params(const Layout &L, const float *in, float *out)
{
    int64_t o = 0;
    for (int l = 0; l < kLayers; l++) {
        for (int k = 0; k < kInDim[l]; k++)
            for (int n = 0; n < kOutDim[l]; n++) out[o++] = in[L.w[l] + (int64_t)k * kOutP[l] + n];
        for (int n = 0; n < kOutDim[l]; n++) out[o++] = in[L.b[l] + n];
    }
}

}  // namespace

extern "C" {

int vh_mlp_init(vh_ctx *c, const float *params, const float *freq, const vh_mlp_opts *o)
{
    clear_error();
    if (!c || !params || !freq || !o || !(o->length > 0.0f) || !(o->floor > 0.0f) ||
        !(o->scale > 0.0f) || !std::isfinite(o->gain)) {
        set_error("NULL context, parameters, frequencies or options, or a length, floor or scale "
                  "that is not positive");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    VH_HIP(hipSetDevice(c->device));
    // a data set already built stays: new weights can train on it
    DevBuf X, T;
    int64_t rows = 0, cap = 0;
    if (c->mlp) {
        std::swap(X, c->mlp->X);
        std::swap(T, c->mlp->T);
        rows = c->mlp->rows;
        cap = c->mlp->cap;
    }
    mlp_release(c);
    MlpData *D = new MlpData();
    c->mlp = D;
    std::swap(D->X, X);
    std::swap(D->T, T);
    D->rows = rows;
    D->cap = cap;
    D->L = layout();
    D->length = o->length;
    D->floor = o->floor;
    D->gain = o->gain;
    D->scale = o->scale;
    std::memcpy(D->freq, freq, sizeof(D->freq));
    std::vector<float> w;
    pad_params(D->L, params, w);
    const size_t pb = (size_t)D->L.nparam * 4, wb = (size_t)D->L.nweight * 2;
    hipStream_t s = c->stream;
    int rc;
    if ((rc = D->master.ensure(pb, "weights")) || (rc = D->m.ensure(pb, "Adam moments")) ||
        (rc = D->v.ensure(pb, "Adam moments")) || (rc = D->grad.ensure(pb, "gradients")) ||
        (rc = D->wb.ensure(wb, "BF16 weights")) || (rc = D->wtb.ensure(wb, "BF16 weights")) ||
        (rc = D->state.ensure(64, "training counters")) || (rc = D->loss1.ensure(64, "loss")))
        return rc;
    std::vector<float> zero((size_t)D->L.nparam, 0.0f);
    VH_HIP(hipMemcpyAsync(D->master.ptr, w.data(), pb, hipMemcpyHostToDevice, s));
    VH_HIP(hipMemcpyAsync(D->m.ptr, zero.data(), pb, hipMemcpyHostToDevice, s));
    VH_HIP(hipMemcpyAsync(D->v.ptr, zero.data(), pb, hipMemcpyHostToDevice, s));
    VH_HIP(hipMemcpyAsync(D->grad.ptr, zero.data(), pb, hipMemcpyHostToDevice, s));
    AdamArgs a = adam_args(D);
    hipLaunchKernelGGL(k_weights, dim3(blocks(a.n)), dim3(256), 0, s, a);
    VH_HIP(hipGetLastError());
    VH_WAIT(c, s);
    D->ready = true;
    return VH_OK;
}

int vh_mlp_params(vh_ctx *c, float *out)
{
    clear_error();
    if (!c || !out) {
        set_error("NULL context or output");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    VH_HIP(hipSetDevice(c->device));
    MlpData *D = c->mlp;
    std::vector<float> w((size_t)D->L.nparam);
    VH_HIP(hipMemcpyAsync(w.data(), D->master.ptr, w.size() * 4, hipMemcpyDeviceToHost, c->stream));
    VH_WAIT(c, c->stream);
    unpad_params(D->L, w.data(), out);
    return VH_OK;
}

int vh_mlp_forward(vh_ctx *c, const float *x, int64_t n, float *y)
{
    clear_error();
    if (!c || n < 0 || (n > 0 && (!x || !y))) {
        set_error("bad forward arguments");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    if (n == 0) return VH_OK;
    VH_HIP(hipSetDevice(c->device));
    MlpData *D = c->mlp;
    hipStream_t s = c->stream;
    const int64_t npad = (n + 15) / 16 * 16;
    if ((rc = ensure_acts(D, npad)) || (rc = D->xf.ensure((size_t)n * kIn * 4, "inputs"))) return rc;
    VH_HIP(hipMemcpyAsync(D->xf.ptr, x, (size_t)n * kIn * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pack, dim3(blocks(npad)), dim3(256), 0, s, (const float *)D->xf.ptr, n, npad,
                       (__bf16 *)D->h[0].ptr, (__bf16 *)nullptr);
    {
        KernelTimer timer(c, VH_KERNEL_MLP, s);
        forward(D, npad, false, s);
        timer.done();
    }
    VH_HIP(hipGetLastError());
    std::vector<float> buf((size_t)n * kOutPad);
    VH_HIP(hipMemcpyAsync(buf.data(), D->y.ptr, buf.size() * 4, hipMemcpyDeviceToHost, s));
    VH_WAIT(c, s);
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < kOut; k++) y[i * kOut + k] = buf[(size_t)i * kOutPad + k];
    return VH_OK;
}

int vh_mlp_grad(vh_ctx *c, const float *x, const float *t, int64_t n, float *grad, float *loss_out)
{
    clear_error();
    if (!c || !x || !t || !grad || n <= 0 || n % 32 != 0) {
        set_error("bad gradient arguments: n must be a positive multiple of 32");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    VH_HIP(hipSetDevice(c->device));
    MlpData *D = c->mlp;
    hipStream_t s = c->stream;
    if ((rc = ensure_acts(D, n)) || (rc = D->xf.ensure((size_t)n * kIn * 4, "inputs"))) return rc;
    VH_HIP(hipMemcpyAsync(D->xf.ptr, x, (size_t)n * kIn * 4, hipMemcpyHostToDevice, s));
    VH_HIP(hipMemcpyAsync(D->tf.ptr, t, (size_t)n * kOut * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pack, dim3(blocks(n)), dim3(256), 0, s, (const float *)D->xf.ptr, n, n,
                       (__bf16 *)D->h[0].ptr, (__bf16 *)D->hk[0].ptr);
    forward(D, n, true, s);
    loss(D, (const float *)D->tf.ptr, n, (float *)D->loss1.ptr, nullptr, s);
    backward(D, n, s);
    VH_HIP(hipGetLastError());
    std::vector<float> g((size_t)D->L.nparam);
    float l = 0.0f;
    VH_HIP(hipMemcpyAsync(g.data(), D->grad.ptr, g.size() * 4, hipMemcpyDeviceToHost, s));
    VH_HIP(hipMemcpyAsync(&l, D->loss1.ptr, 4, hipMemcpyDeviceToHost, s));
    VH_WAIT(c, s);
    unpad_params(D->L, g.data(), grad);
    if (loss_out) *loss_out = l;
    return VH_OK;
}

int vh_mlp_adam(vh_ctx *c, const float *grad, float lr, float beta1, float beta2, float eps)
{
    clear_error();
    if (!c || !grad) {
        set_error("NULL context or gradient");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    VH_HIP(hipSetDevice(c->device));
    MlpData *D = c->mlp;
    hipStream_t s = c->stream;
    std::vector<float> g;
    pad_params(D->L, grad, g);
    int64_t st[5] = {D->step, 0, 0, 0, 0};
    VH_HIP(hipMemcpyAsync(D->grad.ptr, g.data(), g.size() * 4, hipMemcpyHostToDevice, s));
    VH_HIP(hipMemcpyAsync(D->state.ptr, st, sizeof(st), hipMemcpyHostToDevice, s));
    AdamArgs a = adam_args(D);
    a.lr = lr;
    a.beta1 = beta1;
    a.beta2 = beta2;
    a.eps = eps;
    a.state = (const int64_t *)D->state.ptr;
    hipLaunchKernelGGL(k_adam, dim3(blocks(a.n)), dim3(256), 0, s, a);
    VH_HIP(hipGetLastError());
    VH_WAIT(c, s);
    D->step += 1;
    return VH_OK;
}

int vh_mlp_features(vh_ctx *c, const float *emit0, float *x)
{
    clear_error();
    if (!c || !x) {
        set_error("NULL context or output");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    BounceView bv;
    if (!bounce_view(c, &bv)) {
        set_error("no patches loaded: call vh_patches_load first");
        return VH_ESTATE;
    }
    VH_HIP(hipSetDevice(c->device));
    MlpData *D = c->mlp;
    hipStream_t s = c->stream;
    const int64_t P = bv.npatch;
    LocalBuf e;
    if ((rc = ensure_acts(D, (P + 15) / 16 * 16)) || (rc = D->xf.ensure((size_t)P * kIn * 4, "features")))
        return rc;
    FeatArgs f = feat_args(D, bv);
    if (emit0) {
        std::vector<float> buf((size_t)P * 4, 0.0f);
        for (int64_t p = 0; p < P; p++)
            for (int k = 0; k < 3; k++) buf[4 * p + k] = emit0[3 * p + k];
        if ((rc = e.ensure(buf.size() * 4, "emission"))) return rc;
        VH_HIP(hipMemcpyAsync(e.ptr, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, s));
        f.e0 = (const float4 *)e.ptr;
    } else if (bv.direct) {
        f.e0 = bv.direct;
    } else {
        set_error("NULL emission and no vh_patch_lights result");
        return VH_ESTATE;
    }
    f.n = P;
    f.x = (__bf16 *)D->h[0].ptr;
    f.xf = (float *)D->xf.ptr;
    hipLaunchKernelGGL(k_features, dim3(blocks(P)), dim3(256), 0, s, f);
    VH_HIP(hipGetLastError());
    VH_HIP(hipMemcpyAsync(x, D->xf.ptr, (size_t)P * kIn * 4, hipMemcpyDeviceToHost, s));
    VH_WAIT(c, s);
    return VH_OK;
}

int vh_mlp_dataset(vh_ctx *c, int32_t op, int64_t *rows)
{
    clear_error();
    if (!c || (op != VH_MLP_DATA_CLEAR && op != VH_MLP_DATA_ADD)) {
        set_error("NULL context or bad data-set operation");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    MlpData *D = c->mlp;
    if (op == VH_MLP_DATA_CLEAR) {
        D->rows = 0;
        if (rows) *rows = 0;
        return VH_OK;
    }
    BounceView bv;
    if (!bounce_view(c, &bv) || !bv.e0 || !bv.total) {
        set_error("no finished bounce: call vh_bounce first (its emission and light become the row)");
        return VH_ESTATE;
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    LocalBuf idx;
    int64_t n = 0;
    if ((rc = leaf_rows(c, bv, idx, &n, s))) return rc;
    const int64_t need = D->rows + n;
    if (need >= ((int64_t)1 << 31)) {
        set_error("the data set would pass 2^31 rows");
        return VH_EINVAL;
    }
    if (need > D->cap) {
        // grow by half again, keeping the rows already held
        int64_t cap = std::max<int64_t>(need, D->cap + D->cap / 2);
        LocalBuf X2, T2;
        if ((rc = X2.ensure((size_t)cap * kInPad * 2, "data set")) ||
            (rc = T2.ensure((size_t)cap * kLight * 4, "data set")))
            return rc;
        if (D->rows) {
            VH_HIP(hipMemcpyAsync(X2.ptr, D->X.ptr, (size_t)D->rows * kInPad * 2, hipMemcpyDeviceToDevice, s));
            VH_HIP(hipMemcpyAsync(T2.ptr, D->T.ptr, (size_t)D->rows * kLight * 4, hipMemcpyDeviceToDevice, s));
        }
        VH_WAIT(c, s);
        D->X.release();
        D->T.release();
        std::swap(D->X.ptr, X2.ptr);
        std::swap(D->X.bytes, X2.bytes);
        std::swap(D->T.ptr, T2.ptr);
        std::swap(D->T.bytes, T2.bytes);
        D->cap = cap;
    }
    if (n > 0) {
        FeatArgs f = feat_args(D, bv);
        f.rows = (const int32_t *)idx.ptr;
        f.n = n;
        f.e0 = bv.e0;
        f.total = bv.total;
        f.x = (__bf16 *)D->X.ptr + D->rows * kInPad;
        f.t = (float *)D->T.ptr + D->rows * kLight;
        hipLaunchKernelGGL(k_features, dim3(blocks(n)), dim3(256), 0, s, f);
        VH_HIP(hipGetLastError());
        VH_WAIT(c, s);
    }
    D->rows = need;
    if (rows) *rows = need;
    return VH_OK;
}

int vh_mlp_train(vh_ctx *c, const vh_mlp_train_opts *o, float *epoch_loss)
{
    clear_error();
    if (!c || !o || o->epochs < 1 || o->batch < 32 || o->batch % 32 != 0 || !(o->lr > 0.0f)) {
        set_error("bad training options: epochs >= 1, batch a positive multiple of 32, lr > 0");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    MlpData *D = c->mlp;
    const int64_t nb = o->batch, n = D->rows;
    const int64_t spe = n / nb;
    if (spe < 1) {
        set_error("the data set holds " + std::to_string(n) + " rows, fewer than one batch");
        return VH_ESTATE;
    }
    VH_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int64_t total = spe * o->epochs;
    if ((rc = ensure_acts(D, nb)) || (rc = D->losses.ensure((size_t)total * 4 + 16, "losses"))) return rc;
    AdamArgs aa = adam_args(D);
    aa.lr = o->lr;
    aa.beta1 = o->beta1;
    aa.beta2 = o->beta2;
    aa.eps = o->eps;
    aa.state = (const int64_t *)D->state.ptr;
    int64_t *state = (int64_t *)D->state.ptr;
    const bool fused = o->fused != 0;
    StepRowsArgs ra{};
    StepWeightsArgs wa{};
    constexpr int32_t kRun = 8;  // fused steps per graph: the counters advance once per run
    const int32_t nparts = (int32_t)(nb / kFRows);
    if (fused) {
        if ((rc = D->loss_part.ensure((size_t)kRun * nparts * 8 + 8, "loss partials"))) return rc;
        ra.X = (const __bf16 *)D->X.ptr;
        ra.T = (const float *)D->T.ptr;
        ra.n = n;
        ra.nb = (int32_t)nb;
        ra.gain = D->gain;
        ra.inv_scale = 1.0f / D->scale;
        ra.state = state;
        ra.wb = (const __bf16 *)D->wb.ptr;
        ra.wtb = (const __bf16 *)D->wtb.ptr;
        ra.master = (const float *)D->master.ptr;
        for (int l = 0; l < kLayers; l++) {
            ra.boff[l] = D->L.b[l];
            ra.wtoff[l] = D->L.wt[l];
            ra.hk[l] = (__bf16 *)D->hk[l].ptr;
            ra.dzk[l] = (__bf16 *)D->dzk[l].ptr;
            wa.hk[l] = ra.hk[l];
            wa.dzk[l] = ra.dzk[l];
        }
        // the row splits of the per-layer weight gradient (backward())
        const int32_t K = (int32_t)nb;
        int32_t splits = K / 256;
        splits = splits < 1 ? 1 : (splits > kMaxSplits ? kMaxSplits : splits);
        const int32_t kc = ((K + splits - 1) / splits + 31) / 32 * 32;
        wa.K = K;
        wa.kc = kc;
        wa.splits = (K + kc - 1) / kc;
        wa.adam = aa;
    }
    // `count` fused steps (at most kRun), then the losses and the counters
    auto fused_run = [&](int32_t count) {
        for (int32_t k = 0; k < count; k++) {
            StepRowsArgs r = ra;
            r.ahead = k;
            r.loss_part = (double *)D->loss_part.ptr + (int64_t)k * nparts;
            StepWeightsArgs w = wa;
            w.ahead = k;
            hipLaunchKernelGGL(k_step_rows, dim3((uint32_t)nparts), dim3(64 * kFWaves), 0, s, r);
            hipLaunchKernelGGL(k_step_weights, dim3((uint32_t)kWTiles), dim3(64 * kWWaves), 0, s, w);
        }
        hipLaunchKernelGGL(k_step_tail, dim3(1), dim3(256), 0, s, (const double *)D->loss_part.ptr, nparts,
                           count, (int32_t)nb, (float *)D->losses.ptr, state);
    };
    auto step = [&]() {
        hipLaunchKernelGGL(k_batch, dim3(blocks(nb)), dim3(256), 0, s, (const __bf16 *)D->X.ptr,
                           (const float *)D->T.ptr, n, (const int64_t *)state, (int32_t)nb, D->gain,
                           1.0f / D->scale, (__bf16 *)D->h[0].ptr, (__bf16 *)D->hk[0].ptr,
                           (float *)D->tf.ptr);
        forward(D, nb, true, s);
        loss(D, (const float *)D->tf.ptr, nb, (float *)D->losses.ptr, state, s);
        backward(D, nb, s);
        hipLaunchKernelGGL(k_adam, dim3(blocks(aa.n)), dim3(256), 0, s, aa);
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, s, state);
    };
    // the replayed unit: one step of the per-layer kernels, or a run of kRun
    // fused steps (and one of the epoch's remainder)
    const int32_t unit = fused ? kRun : 1;
    const int32_t rem = (int32_t)(spe % unit);
    hipGraph_t graphs[2] = {nullptr, nullptr};
    hipGraphExec_t execs[2] = {nullptr, nullptr};
    auto launch_unit = [&](int32_t count) {
        if (fused)
            fused_run(count);
        else
            step();
    };
    auto release = [&]() {
        for (int q = 0; q < 2; q++) {
            if (execs[q]) (void)hipGraphExecDestroy(execs[q]);
            if (graphs[q]) (void)hipGraphDestroy(graphs[q]);
        }
    };
    if (o->use_graph) {
        for (int q = 0; q < 2; q++) {
            const int32_t count = q == 0 ? unit : rem;
            if (count == 0) continue;
            VH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch_unit(count);
            hipError_t e = hipStreamEndCapture(s, &graphs[q]);
            if (e == hipSuccess) e = hipGraphInstantiate(&execs[q], graphs[q], nullptr, nullptr, 0);
            if (e != hipSuccess) {
                release();
                set_error(std::string("training-step graph failed: ") + hipGetErrorString(e));
                return VH_EDEVICE;
            }
        }
    }
    std::mt19937_64 rng((uint64_t)o->seed * 0x9E3779B97F4A7C15ull + 1);
    hipError_t failure = hipSuccess;
    std::vector<int64_t> st(6);
    for (int32_t ep = 0; ep < o->epochs && failure == hipSuccess; ep++) {
        // a multiplier coprime to n and an offset: a fresh permutation of the rows
        uint64_t a;
        do {
            a = rng() % (uint64_t)n;
        } while (a == 0 || std::gcd(a, (uint64_t)n) != 1);
        if (n == 1) a = 1;
        st = {D->step + (int64_t)ep * spe, 0, (int64_t)a, (int64_t)(rng() % (uint64_t)n), (int64_t)ep * spe, 0};
        failure = hipMemcpyAsync(state, st.data(), st.size() * 8, hipMemcpyHostToDevice, s);
        if (failure == hipSuccess) failure = hipStreamSynchronize(s);
        KernelTimer timer(c, VH_KERNEL_MLP, s);
        for (int64_t k = 0; k < spe && failure == hipSuccess;) {
            const int32_t count = spe - k >= unit ? unit : rem;
            const int q = count == unit ? 0 : 1;
            if (execs[q]) {
                failure = hipGraphLaunch(execs[q], s);
            } else {
                launch_unit(count);
                failure = hipGetLastError();
            }
            k += count;
        }
        timer.done();
    }
    if (failure == hipSuccess) failure = hipStreamSynchronize(s);
    release();
    if (failure != hipSuccess) {
        set_error(std::string("training failed: ") + hipGetErrorString(failure));
        return VH_EDEVICE;
    }
    D->step += total;
    if (epoch_loss) {
        std::vector<float> l((size_t)total);
        VH_HIP(hipMemcpy(l.data(), D->losses.ptr, l.size() * 4, hipMemcpyDeviceToHost));
        for (int32_t ep = 0; ep < o->epochs; ep++) {
            double sum = 0.0;
            for (int64_t k = 0; k < spe; k++) sum += l[(size_t)(ep * spe + k)];
            epoch_loss[ep] = (float)(sum / (double)spe);
        }
    }
    return VH_OK;
}

int vh_mlp_predict(vh_ctx *c, const float *emit0, float *start)
{
    clear_error();
    if (!c) {
        set_error("NULL context");
        return VH_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = check_mlp(c);
    if (rc) return rc;
    BounceView bv;
    if (!bounce_view(c, &bv)) {
        set_error("no patches loaded: call vh_patches_load first");
        return VH_ESTATE;
    }
    if (!emit0 && !bv.direct) {
        set_error("NULL emission and no vh_patch_lights result");
        return VH_ESTATE;
    }
    VH_HIP(hipSetDevice(c->device));
    MlpData *D = c->mlp;
    hipStream_t s = c->stream;
    const int64_t P = bv.npatch;
    LocalBuf e;
    if (D->pred_generation != bv.generation) {
        // once per patch system: the rows, grouped by cluster
        D->pred_generation = 0;
        if ((rc = leaf_rows(c, bv, D->pred_rows, &D->pred_n, s, &D->pred_groups, &D->pred_ngroups)))
            return rc;
        D->pred_generation = bv.generation;
    }
    const DevBuf &idx = D->pred_rows, &groups = D->pred_groups;
    const int64_t n = D->pred_n;
    const int32_t ngroups = D->pred_ngroups;
    const int64_t npad = (n + 15) / 16 * 16;
    if ((rc = ensure_acts(D, npad > 0 ? npad : 16))) return rc;
    FeatArgs f = feat_args(D, bv);
    if (emit0) {
        std::vector<float> buf((size_t)P * 4, 0.0f);
        for (int64_t p = 0; p < P; p++)
            for (int k = 0; k < 3; k++) buf[4 * p + k] = emit0[3 * p + k];
        if ((rc = e.ensure(buf.size() * 4, "emission"))) return rc;
        VH_HIP(hipMemcpyAsync(e.ptr, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, s));
        VH_WAIT(c, s);
        f.e0 = (const float4 *)e.ptr;
    } else {
        f.e0 = bv.direct;
    }
    float4 *pred = bounce_pred_buffer(c);
    if (!pred) return VH_EDEVICE;
    hipLaunchKernelGGL(k_zero4, dim3(blocks(P)), dim3(256), 0, s, pred, P);
    if (n > 0) {
        // rows past n (padding) are zero inputs; their outputs are not read
        hipLaunchKernelGGL(k_pack, dim3(blocks(npad)), dim3(256), 0, s, (const float *)nullptr, (int64_t)0,
                           npad, (__bf16 *)D->h[0].ptr, (__bf16 *)nullptr);
        f.rows = (const int32_t *)idx.ptr;
        f.n = n;
        f.x = (__bf16 *)D->h[0].ptr;
        hipLaunchKernelGGL(k_features, dim3(blocks(n)), dim3(256), 0, s, f);
        {
            KernelTimer timer(c, VH_KERNEL_MLP, s);
            forward(D, npad, false, s);
            timer.done();
        }
        hipLaunchKernelGGL(k_cluster_start, dim3((uint32_t)ngroups), dim3(256), 0, s, (const float *)D->y.ptr,
                           (const int32_t *)idx.ptr, (const int32_t *)groups.ptr, f.e0, D->gain, D->scale,
                           pred);
    }
    bounce_pull_up(c, pred, s);
    VH_HIP(hipGetLastError());
    if (start) {
        std::vector<float> buf((size_t)P * 4);
        VH_HIP(hipMemcpyAsync(buf.data(), pred, buf.size() * 4, hipMemcpyDeviceToHost, s));
        VH_WAIT(c, s);
        for (int64_t p = 0; p < P; p++)
            for (int k = 0; k < 3; k++) start[3 * p + k] = buf[4 * p + k];
    }
    VH_WAIT(c, s);
    bounce_pred_ready(c);
    return VH_OK;
}

}  // extern "C"
