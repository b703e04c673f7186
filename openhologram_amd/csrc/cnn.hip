// cnn.hip -- the learned correction of a coarse depth-map hologram: a small
// residual CNN (3x3 convolutions) with its forward and backward passes and an
// Adam step, all as HIP kernels on gfx950.  There is no counterpart in the
// reference; the network corrects a 16-layer depth-map field toward the
// 256-layer one (both produced by oph_depthmap).
//
// Network (channels C, blocks NB; real/imaginary parts as 2 channels):
//   h0 = conv_head(x)                       2 -> C
//   h_b = h_{b-1} + conv_b2(relu(conv_b1(h_{b-1})))     b = 1..NB
//   y  = x + conv_tail(h_NB)                C -> 2
// Tensors are NHWC float32 in HBM.  A convolution is an implicit GEMM on
// MFMA: M = pixels, N = output channels, K = 9 taps x input channels; each
// wave produces 32 pixels x 32 channels with v_mfma_f32_32x32x16_bf16,
// operands rounded to BF16 (weights and activations), FP32 accumulation;
// bias, ReLU and the residual add are fused into the epilogue.  The input
// gradient is the same convolution with flipped taps and transposed
// channels; the weight gradient is a GEMM over pixels (M = output channels,
// N = 9 x input channels + 1 bias column) split into pixel chunks whose
// partial sums are reduced in a fixed order (deterministic).
#include <cmath>
#include <cstdio>
#include <vector>

#include "oph_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct oph_cnn {
    int blocks = 8, ch = 32;
    int64_t nparams = 0;
    float* params = nullptr;  // device, fp32
    float* grads = nullptr;
    float* m = nullptr;       // Adam moments
    float* v = nullptr;
    float* wt = nullptr;      // flipped/transposed weights for input gradients
    int64_t step = 0;
    // activation workspace
    float* act = nullptr;
    size_t act_bytes = 0;
    float* part = nullptr;    // weight-gradient partials
    size_t part_bytes = 0;
    double* red = nullptr;    // loss reduction
};

namespace {

constexpr int CMAX = 32;

// parameter offsets: head W[C][9][2], b[C]; block b: W1[C][9][C], b1[C],
// W2[C][9][C], b2[C]; tail W[2][9][C], b[2]
struct Layout {
    int C, NB;
    __host__ __device__ int64_t head_w() const { return 0; }
    __host__ __device__ int64_t head_b() const { return (int64_t)C * 9 * 2; }
    __host__ __device__ int64_t block_base(int b) const {
        return head_b() + C + (int64_t)b * (2 * ((int64_t)C * 9 * C + C));
    }
    __host__ __device__ int64_t w1(int b) const { return block_base(b); }
    __host__ __device__ int64_t b1(int b) const { return w1(b) + (int64_t)C * 9 * C; }
    __host__ __device__ int64_t w2(int b) const { return b1(b) + C; }
    __host__ __device__ int64_t b2(int b) const { return w2(b) + (int64_t)C * 9 * C; }
    __host__ __device__ int64_t tail_w() const { return block_base(NB); }
    __host__ __device__ int64_t tail_b() const { return tail_w() + 2 * 9 * (int64_t)C; }
    __host__ __device__ int64_t total() const { return tail_b() + 2; }
};

// ---- forward convolution (implicit GEMM) -------------------------------------

struct ConvArgs {
    const float* in;      // [B][H][W][CIN]
    const float* w;       // [COUT][9][CIN]
    const float* bias;    // [COUT] or null
    const float* res;     // residual added after the bias (or null), [..][COUT]
    float* out;           // [B][H][W][COUT]
    int B, H, W;
    int relu;
};

// A operand for lane (pixel p, k0..k0+7): taps/channels of the input patch
template <int CIN>
__device__ __forceinline__ bf16x8 load_patch(const ConvArgs& a, int b, int y, int x, int k0) {
    bf16x8 v;
    if constexpr (CIN % 8 == 0) {
        const int tap = k0 / CIN, c = k0 % CIN;
        const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        if (tap < 9 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
            const float4* s =
                reinterpret_cast<const float4*>(a.in + (((size_t)b * a.H + yy) * a.W + xx) * CIN + c);
            const float4 p0 = s[0], p1 = s[1];
            v[0] = (__bf16)p0.x; v[1] = (__bf16)p0.y; v[2] = (__bf16)p0.z; v[3] = (__bf16)p0.w;
            v[4] = (__bf16)p1.x; v[5] = (__bf16)p1.y; v[6] = (__bf16)p1.z; v[7] = (__bf16)p1.w;
        } else {
            v = (bf16x8)(__bf16)0.f;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + e;
            const int tap = k / CIN, c = k % CIN;
            const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
            float f = 0.f;
            if (tap < 9 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                f = a.in[(((size_t)b * a.H + yy) * a.W + xx) * CIN + c];
            v[e] = (__bf16)f;
        }
    }
    return v;
}

// B operand for lane (output channel n, k0..k0+7)
template <int CIN, int COUT>
__device__ __forceinline__ bf16x8 load_weights(const float* w, int n, int k0) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = k0 + e;
        float f = 0.f;
        if (n < COUT && k < 9 * CIN) f = w[(size_t)n * 9 * CIN + k];
        v[e] = (__bf16)f;
    }
    return v;
}

// One wave: 32 consecutive pixels (row-major over B*H*W) x 32 output
// channels.  Lane l: pixel l%32, k half l/32; accumulator register v of lane
// l is pixel 8(v/4) + 4(l/32) + v%4, channel l%32.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv3x3_fwd(ConvArgs a) {
    constexpr int K = 9 * CIN;
    constexpr int KSTEPS = (K + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t npix = (int64_t)a.B * a.H * a.W;
    const int64_t p0 = tile * 32;
    if (p0 >= npix) return;
    const int64_t p = p0 + (lane & 31);
    const bool pok = p < npix;
    const int b = (int)(p / ((int64_t)a.H * a.W));
    const int rem = (int)(p % ((int64_t)a.H * a.W));
    const int y = rem / a.W, x = rem % a.W;
    f32x16 acc = (f32x16)0.f;
    const int n = lane & 31;
#pragma unroll 2
    for (int s = 0; s < KSTEPS; ++s) {
        const int k0 = 16 * s + 8 * (lane >> 5);
        const bf16x8 A = pok ? load_patch<CIN>(a, b, y, x, k0) : (bf16x8)(__bf16)0.f;
        const bf16x8 Bv = load_weights<CIN, COUT>(a.w, n, k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bv, acc, 0, 0, 0);
    }
    if (n >= COUT) return;
    const float bias = a.bias ? a.bias[n] : 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int64_t q = p0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (q >= npix) continue;
        float o = acc[v] + bias;
        if (a.relu) o = fmaxf(o, 0.f);
        if (a.res) o += a.res[q * COUT + n];
        a.out[q * COUT + n] = o;
    }
}

// flipped taps, transposed channels: wt[c_in][8 - tap][c_out] = w[c_out][tap][c_in]
__global__ void conv_flip(const float* w, float* wt, int cin, int cout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cin * cout * 9) return;
    const int co = i / (9 * cin), r = i % (9 * cin), tap = r / cin, ci = r % cin;
    wt[((size_t)ci * 9 + (8 - tap)) * cout + co] = w[i];
}

// relu'(pre) from the post-activation: g *= [act > 0]
__global__ void relu_mask(float* g, const float* act, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(act[i] > 0.f)) g[i] = 0.f;
}

// ---- weight gradient -------------------------------------------------------------
// dW[n][k] = sum_p dY[p][n] X[p + off(tap k)][c k], db[n] = sum_p dY[p][n].
// GEMM M = COUT (padded to 32), N = 9 CIN + 1 (bias column), K = pixels;
// workgroup g sums pixel chunk g into part[g][32][NCOL].

struct DwArgs {
    const float* dy;   // [npix][COUT]
    const float* x;    // [npix][CIN]
    float* part;       // [nchunks][32][ncol]
    int B, H, W;
    int64_t chunk;     // pixels per workgroup (multiple of 16)
};

template <int CIN, int COUT>
__global__ __launch_bounds__(64) void conv3x3_dw(DwArgs a) {
    constexpr int NCOL = 9 * CIN + 1;
    constexpr int NT = (NCOL + 31) / 32;
    const int lane = threadIdx.x;
    const int64_t npix = (int64_t)a.B * a.H * a.W;
    const int64_t pb = (int64_t)blockIdx.x * a.chunk;
    const int64_t pe = min(npix, pb + a.chunk);
    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f32x16)0.f;
    const int row = lane & 31;          // output channel (A row) / column within a tile (B)
    const int kh = 8 * (lane >> 5);     // pixel offset of this lane's 8 values
    for (int64_t q0 = pb; q0 < pe; q0 += 16) {
        bf16x8 A;
        int qb[8], qy[8], qx[8];
        bool qv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int64_t q = q0 + kh + e;
            qv[e] = q < pe;
            const int64_t qq = qv[e] ? q : pb;
            qb[e] = (int)(qq / ((int64_t)a.H * a.W));
            const int rem = (int)(qq % ((int64_t)a.H * a.W));
            qy[e] = rem / a.W;
            qx[e] = rem % a.W;
            float f = 0.f;
            if (qv[e] && row < COUT) f = a.dy[q * COUT + row];
            A[e] = (__bf16)f;
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int col = 32 * t + row;   // (tap, c) or the bias column
            const int tap = col / CIN, c = col % CIN;
            const int dyo = tap / 3 - 1, dxo = tap % 3 - 1;
            bf16x8 Bv;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float f = 0.f;
                if (qv[e] && col < NCOL) {
                    if (col == NCOL - 1) {
                        f = 1.f;
                    } else {
                        const int yy = qy[e] + dyo, xx = qx[e] + dxo;
                        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                            f = a.x[(((size_t)qb[e] * a.H + yy) * a.W + xx) * CIN + c];
                    }
                }
                Bv[e] = (__bf16)f;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bv, acc[t], 0, 0, 0);
        }
    }
    float* dst = a.part + (size_t)blockIdx.x * 32 * NCOL;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int r = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);  // output channel
            const int col = 32 * t + (lane & 31);
            if (col < NCOL) dst[(size_t)r * NCOL + col] = acc[t][v];
        }
}

// sum the chunk partials in chunk order into the gradient (W and b)
__global__ void dw_reduce(const float* part, int nchunks, int cout, int ncol, float* gw, float* gb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cout * ncol) return;
    const int r = i / ncol, col = i % ncol;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += part[((size_t)c * 32 + r) * ncol + col];
    if (col == ncol - 1)
        gb[r] = (float)s;
    else
        gw[(size_t)r * (ncol - 1) + col] = (float)s;
}

// ---- loss, Adam ------------------------------------------------------------------

// dy = 2 (y - t) / n, per-block partial sums of (y - t)^2 (fixed order)
__global__ void mse_grad(const float* y, const float* t, float* dy, int64_t n, double* part) {
    __shared__ double s[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)n;
        e2 = (double)d * d;
    }
    s[threadIdx.x] = e2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void sum_partials(double* part, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        part[n] = s;
    }
}

__global__ void adam_step(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void add_inplace(float* a, const float* b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += b[i];
}

// ---- host helpers ------------------------------------------------------------------

template <int CIN, int COUT>
int conv_fwd(oph_ctx* ctx, const float* in, const float* w, const float* bias, const float* res, float* out,
             int B, int H, int W, int relu) {
    ConvArgs a{in, w, bias, res, out, B, H, W, relu};
    const int64_t npix = (int64_t)B * H * W;
    const int64_t tiles = (npix + 31) / 32;
    OPH_LAUNCH(ctx, "conv3x3_fwd", (conv3x3_fwd<CIN, COUT>), dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, a);
    return OPH_OK;
}

template <int CIN, int COUT>
int conv_dw(oph_ctx* ctx, oph_cnn* net, const float* dy, const float* x, int B, int H, int W, float* gw,
            float* gb) {
    constexpr int NCOL = 9 * CIN + 1;
    const int64_t npix = (int64_t)B * H * W;
    int64_t chunk = 4096;
    int64_t nch = (npix + chunk - 1) / chunk;
    const size_t need = (size_t)nch * 32 * NCOL * sizeof(float);
    if (need > net->part_bytes) return oph_fail(ctx, OPH_ENOMEM, "weight-gradient workspace too small");
    DwArgs a{dy, x, net->part, B, H, W, chunk};
    OPH_LAUNCH(ctx, "conv3x3_dw", (conv3x3_dw<CIN, COUT>), dim3((unsigned)nch), dim3(64), 0, a);
    OPH_LAUNCH(ctx, "dw_reduce", dw_reduce, dim3((COUT * NCOL + 255) / 256), dim3(256), 0, net->part, (int)nch, COUT,
               NCOL, gw, gb);
    return OPH_OK;
}

template <int CIN, int COUT>
int conv_dx(oph_ctx* ctx, oph_cnn* net, const float* dy, const float* w, float* dx, const float* res, int B, int H,
            int W) {
    // dX = conv(dY, flipped/transposed W): input channels COUT, output CIN
    OPH_LAUNCH(ctx, "conv_flip", conv_flip, dim3((CIN * COUT * 9 + 255) / 256), dim3(256), 0, w, net->wt, CIN,
               COUT);
    return conv_fwd<COUT, CIN>(ctx, dy, net->wt, nullptr, res, dx, B, H, W, 0);
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// activation workspace: x copy handled by caller; h_0..h_NB and a_1..a_NB,
// gradients dh, da, and the output
struct Acts {
    float* h[CMAX + 1];
    float* a[CMAX + 1];
    float* y;
    float* g0;
    float* g1;
    float* g2;
};

int ensure_workspace(oph_ctx* ctx, oph_cnn* net, int64_t npix, Acts& A) {
    const size_t per = (size_t)npix * net->ch * sizeof(float);
    const size_t need = per * (2 * net->blocks + 1 + 3) + (size_t)npix * 2 * sizeof(float);
    if (need > net->act_bytes) {
        if (net->act) hipFree(net->act);
        net->act = nullptr;
        net->act_bytes = 0;
        if (hipMalloc(&net->act, need) != hipSuccess) {
            hipGetLastError();
            return oph_fail(ctx, OPH_ENOMEM, "CNN activation workspace (%zu bytes)", need);
        }
        net->act_bytes = need;
    }
    char* p = reinterpret_cast<char*>(net->act);
    for (int b = 0; b <= net->blocks; ++b) { A.h[b] = reinterpret_cast<float*>(p); p += per; }
    for (int b = 1; b <= net->blocks; ++b) { A.a[b] = reinterpret_cast<float*>(p); p += per; }
    A.g0 = reinterpret_cast<float*>(p); p += per;
    A.g1 = reinterpret_cast<float*>(p); p += per;
    A.g2 = reinterpret_cast<float*>(p); p += per;
    A.y = reinterpret_cast<float*>(p);
    const int64_t nch = (npix + 4095) / 4096;
    const size_t pneed = (size_t)nch * 32 * (9 * 32 + 1) * sizeof(float);
    if (pneed > net->part_bytes) {
        if (net->part) hipFree(net->part);
        net->part = nullptr;
        net->part_bytes = 0;
        if (hipMalloc(&net->part, pneed) != hipSuccess) {
            hipGetLastError();
            return oph_fail(ctx, OPH_ENOMEM, "CNN gradient workspace");
        }
        net->part_bytes = pneed;
    }
    return OPH_OK;
}

int forward(oph_ctx* ctx, oph_cnn* net, const float* x, int B, int H, int W, Acts& A) {
    const Layout L{net->ch, net->blocks};
    const float* P = net->params;
    int rc;
    if ((rc = conv_fwd<2, 32>(ctx, x, P + L.head_w(), P + L.head_b(), nullptr, A.h[0], B, H, W, 0))) return rc;
    for (int b = 0; b < net->blocks; ++b) {
        if ((rc = conv_fwd<32, 32>(ctx, A.h[b], P + L.w1(b), P + L.b1(b), nullptr, A.a[b + 1], B, H, W, 1))) return rc;
        if ((rc = conv_fwd<32, 32>(ctx, A.a[b + 1], P + L.w2(b), P + L.b2(b), A.h[b], A.h[b + 1], B, H, W, 0)))
            return rc;
    }
    return conv_fwd<32, 2>(ctx, A.h[net->blocks], P + L.tail_w(), P + L.tail_b(), x, A.y, B, H, W, 0);
}

// gradients of loss = mean (y - t)^2 into net->grads; loss into *loss
int backward(oph_ctx* ctx, oph_cnn* net, const float* x, const float* t, int B, int H, int W, Acts& A,
             double* loss) {
    const Layout L{net->ch, net->blocks};
    const float* P = net->params;
    float* G = net->grads;
    const int64_t npix = (int64_t)B * H * W;
    const int64_t n2 = npix * 2, nc = npix * net->ch;
    const unsigned nb = (unsigned)((n2 + 255) / 256);
    float* dy = A.g2;  // 2-channel output gradient (fits in a C-channel buffer)
    OPH_LAUNCH(ctx, "mse_grad", mse_grad, dim3(nb), dim3(256), 0, A.y, t, dy, n2, net->red);
    OPH_LAUNCH(ctx, "sum_partials", sum_partials, dim3(1), dim3(64), 0, net->red, (int)nb);
    int rc;
    // tail: y = x + conv(h_NB)
    if ((rc = conv_dw<32, 2>(ctx, net, dy, A.h[net->blocks], B, H, W, G + L.tail_w(), G + L.tail_b()))) return rc;
    float* dh = A.g0;
    if ((rc = conv_dx<32, 2>(ctx, net, dy, P + L.tail_w(), dh, nullptr, B, H, W))) return rc;
    for (int b = net->blocks - 1; b >= 0; --b) {
        // h_{b+1} = h_b + conv2(a); a = relu(conv1(h_b))
        float* da = A.g1;
        if ((rc = conv_dw<32, 32>(ctx, net, dh, A.a[b + 1], B, H, W, G + L.w2(b), G + L.b2(b)))) return rc;
        if ((rc = conv_dx<32, 32>(ctx, net, dh, P + L.w2(b), da, nullptr, B, H, W))) return rc;
        OPH_LAUNCH(ctx, "relu_mask", relu_mask, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, da, A.a[b + 1], nc);
        if ((rc = conv_dw<32, 32>(ctx, net, da, A.h[b], B, H, W, G + L.w1(b), G + L.b1(b)))) return rc;
        // dh_b = dh_{b+1} + conv1^T(da)
        if ((rc = conv_dx<32, 32>(ctx, net, da, P + L.w1(b), A.g2, dh, B, H, W))) return rc;
        float* tmp = dh;
        dh = A.g2;
        A.g2 = tmp;
    }
    if ((rc = conv_dw<2, 32>(ctx, net, dh, x, B, H, W, G + L.head_w(), G + L.head_b()))) return rc;
    if (loss) {
        double s = 0.0;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(&s, net->red + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        *loss = s / (double)n2;
    }
    return OPH_OK;
}

}  // namespace

extern "C" {

int oph_cnn_create(oph_ctx* ctx, int blocks, int channels, uint64_t seed, oph_cnn** out) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!out) return oph_fail(ctx, OPH_EINVAL, "null output");
    if (channels != 32 || blocks < 0 || blocks > CMAX)
        return oph_fail(ctx, OPH_ENOTSUP, "channels must be 32 and blocks 0..%d", CMAX);
    oph_cnn* n = new oph_cnn();
    n->blocks = blocks;
    n->ch = channels;
    const Layout L{channels, blocks};
    n->nparams = L.total();
    const size_t pb = (size_t)n->nparams * sizeof(float);
    if (hipMalloc(&n->params, pb) != hipSuccess || hipMalloc(&n->grads, pb) != hipSuccess ||
        hipMalloc(&n->m, pb) != hipSuccess || hipMalloc(&n->v, pb) != hipSuccess ||
        hipMalloc(&n->wt, (size_t)9 * 32 * 32 * sizeof(float)) != hipSuccess ||
        hipMalloc(&n->red, (size_t)(1 << 22) * sizeof(double)) != hipSuccess) {
        hipGetLastError();
        oph_cnn_destroy(ctx, n);
        return oph_fail(ctx, OPH_ENOMEM, "CNN parameter allocation");
    }
    // He-uniform init (seeded SplitMix64); the residual branch's second
    // convolution and the tail start small so the network starts near identity
    std::vector<float> h((size_t)n->nparams, 0.f);
    uint64_t s = seed;
    auto fill = [&](int64_t off, int64_t cnt, int fan_in, float gain) {
        const float lim = gain * sqrtf(6.f / (float)fan_in);
        for (int64_t i = 0; i < cnt; ++i) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            h[off + i] = (float)((2.0 * u - 1.0) * lim);
        }
    };
    fill(L.head_w(), (int64_t)channels * 18, 18, 1.f);
    for (int b = 0; b < blocks; ++b) {
        fill(L.w1(b), (int64_t)channels * 9 * channels, 9 * channels, 1.f);
        fill(L.w2(b), (int64_t)channels * 9 * channels, 9 * channels, 0.1f);
    }
    fill(L.tail_w(), 2 * 9 * (int64_t)channels, 9 * channels, 0.1f);
    if (hipMemcpy(n->params, h.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, pb) != hipSuccess || hipMemset(n->v, 0, pb) != hipSuccess ||
        hipMemset(n->grads, 0, pb) != hipSuccess) {
        hipGetLastError();
        oph_cnn_destroy(ctx, n);
        return oph_fail(ctx, OPH_EDEVICE, "CNN parameter upload");
    }
    *out = n;
    return OPH_OK;
}

void oph_cnn_destroy(oph_ctx* ctx, oph_cnn* n) {
    if (!n) return;
    if (ctx) hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    for (void* p : {(void*)n->params, (void*)n->grads, (void*)n->m, (void*)n->v, (void*)n->wt, (void*)n->act,
                    (void*)n->part, (void*)n->red})
        if (p) hipFree(p);
    delete n;
}

int64_t oph_cnn_num_params(const oph_cnn* n) { return n ? n->nparams : 0; }

int oph_cnn_get_params(oph_ctx* ctx, oph_cnn* n, float* host, int which) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !host) return oph_fail(ctx, OPH_EINVAL, "null argument");
    const float* src = which == 1 ? n->grads : n->params;
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(host, src, (size_t)n->nparams * 4, hipMemcpyDeviceToHost, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

int oph_cnn_set_params(oph_ctx* ctx, oph_cnn* n, const float* host) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !host) return oph_fail(ctx, OPH_EINVAL, "null argument");
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(n->params, host, (size_t)n->nparams * 4, hipMemcpyHostToDevice, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    n->step = 0;
    OPH_HIP_CHECK(ctx, hipMemsetAsync(n->m, 0, (size_t)n->nparams * 4, ctx->stream));
    OPH_HIP_CHECK(ctx, hipMemsetAsync(n->v, 0, (size_t)n->nparams * 4, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

int oph_cnn_forward(oph_ctx* ctx, oph_cnn* n, const float* x, float* y, int B, int H, int W, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !x || !y || B < 1 || H < 1 || W < 1) return oph_fail(ctx, OPH_EINVAL, "bad argument");
    const int64_t npix = (int64_t)B * H * W;
    Acts A;
    if ((rc = ensure_workspace(ctx, n, npix, A))) return rc;
    const size_t bytes = (size_t)npix * 2 * sizeof(float);
    const float* dx = x;
    float* dyv = y;
    if (!(flags & OPH_DEVICE_PTR)) {
        float* s = static_cast<float*>(oph_scratch(ctx, SLOT_IN, 2 * bytes));
        if (!s) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(s, x, bytes, hipMemcpyHostToDevice, ctx->stream));
        dx = s;
        dyv = s + npix * 2;
    }
    if ((rc = forward(ctx, n, dx, B, H, W, A))) return rc;
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(dyv, A.y, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(y, dyv, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}

static int train_or_grad(oph_ctx* ctx, oph_cnn* n, const float* x, const float* t, int B, int H, int W,
                         float lr, double* loss, int flags, bool update) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !x || !t || B < 1 || H < 1 || W < 1) return oph_fail(ctx, OPH_EINVAL, "bad argument");
    const int64_t npix = (int64_t)B * H * W;
    if ((npix * 2 + 255) / 256 >= (1 << 22)) return oph_fail(ctx, OPH_EINVAL, "batch too large");
    Acts A;
    if ((rc = ensure_workspace(ctx, n, npix, A))) return rc;
    const size_t bytes = (size_t)npix * 2 * sizeof(float);
    const float* dx = x;
    const float* dt = t;
    if (!(flags & OPH_DEVICE_PTR)) {
        float* s = static_cast<float*>(oph_scratch(ctx, SLOT_IN, 2 * bytes));
        if (!s) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(s, x, bytes, hipMemcpyHostToDevice, ctx->stream));
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(s + npix * 2, t, bytes, hipMemcpyHostToDevice, ctx->stream));
        dx = s;
        dt = s + npix * 2;
    }
    if ((rc = forward(ctx, n, dx, B, H, W, A))) return rc;
    if ((rc = backward(ctx, n, dx, dt, B, H, W, A, loss))) return rc;
    if (update) {
        n->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
        OPH_LAUNCH(ctx, "adam_step", adam_step, dim3((unsigned)((n->nparams + 255) / 256)), dim3(256), 0, n->params,
                   n->grads, n->m, n->v, n->nparams, lr, b1, b2, 1e-8f, c1, c2);
    }
    return oph_finish(ctx, flags);
}

int oph_cnn_train_step(oph_ctx* ctx, oph_cnn* n, const float* x, const float* t, int B, int H, int W, float lr,
                       double* loss, int flags) {
    return train_or_grad(ctx, n, x, t, B, H, W, lr, loss, flags, true);
}

int oph_cnn_gradients(oph_ctx* ctx, oph_cnn* n, const float* x, const float* t, int B, int H, int W, double* loss,
                      int flags) {
    return train_or_grad(ctx, n, x, t, B, H, W, 0.f, loss, flags, false);
}

// file: "OPHCNN1\0", int32 blocks, int32 channels, int64 nparams, float32 params
int oph_cnn_save(oph_ctx* ctx, oph_cnn* n, const char* path) {
    if (!n || !path) return oph_fail(ctx, OPH_EINVAL, "null argument");
    std::vector<float> h((size_t)n->nparams);
    int rc = oph_cnn_get_params(ctx, n, h.data(), 0);
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return oph_fail(ctx, OPH_EIO, "cannot write %s", path);
    const char magic[8] = {'O', 'P', 'H', 'C', 'N', 'N', '1', 0};
    const int32_t hdr[2] = {n->blocks, n->ch};
    const int64_t np = n->nparams;
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(hdr, 4, 2, f) == 2 && fwrite(&np, 8, 1, f) == 1 &&
              fwrite(h.data(), 4, h.size(), f) == h.size();
    fclose(f);
    return ok ? OPH_OK : oph_fail(ctx, OPH_EIO, "short write to %s", path);
}

int oph_cnn_load(oph_ctx* ctx, oph_cnn* n, const char* path) {
    if (!n || !path) return oph_fail(ctx, OPH_EINVAL, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return oph_fail(ctx, OPH_EIO, "cannot read %s", path);
    char magic[8];
    int32_t hdr[2];
    int64_t np = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "OPHCNN1", 8) == 0 && fread(hdr, 4, 2, f) == 2 &&
              fread(&np, 8, 1, f) == 1 && hdr[0] == n->blocks && hdr[1] == n->ch && np == n->nparams;
    std::vector<float> h(ok ? (size_t)np : 0);
    ok = ok && fread(h.data(), 4, h.size(), f) == h.size();
    fclose(f);
    if (!ok) return oph_fail(ctx, OPH_EIO, "%s is not a matching network file", path);
    return oph_cnn_set_params(ctx, n, h.data());
}

}  // extern "C"
