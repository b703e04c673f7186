// depthmap.hip -- depth-map layer hologram (angular-spectrum method) on gfx950.
//
// Replaces HologramGenerator's CPU/GPU paths (DM/Hologram/src/
// HologramGenerator_CPU.cpp:44-270, HologramGenerator_GPU.cpp:194-368,
// HologramKernel.cu:10-187, HologramGenerator.cpp:391-522).  Per colour:
//   u_l   = img/255 [img > 0] [idx == l]                  (CPU.cpp:57-66,139-141)
//   U     = rho sum_l exp(i k z_l) H_l . fftshift(FFT2(fftshift(u_l)))
//                                                          (CPU.cpp:152-160)
//   H_l   = exp(i k (-z_l) sqrt(1 - (lambda fx)^2 - (lambda fy)^2)) on the
//           reference's grids fx = -1/(2px) + x/Sx, fy = 1/(2py) - 1/Sy - y/Sy
//                                                          (CPU.cpp:196-216)
//   encode: keep centred spectrum rows [cropy1, cropy2], normalised inverse
//           fftwShift, times exp(i 2 pi yy/(4 py)), real part (CPU.cpp:231-270,
//           HologramGenerator.cpp:481-522)
// with the scene (w x h) embedded in an Nx x Ny transform grid (Nx = w,
// Ny = h reproduces the reference, which does not pad).
//
// The two fftshifts around each transform cancel against the shifts of the
// encode: with k in FFT order, U_ref(y(k)) = (-1)^(kx+ky) A(k) where
// A = sum_l m_l H_l(y(k)) FFT2(u_l), and the encoded field is IFFT2(crop(A)).
// So the pipeline works on plain FFTs and evaluates H at the reference's
// (off-by-one in fy) frequency of each FFT-order index:
//   fx = s(kx)/Sx,  |fy| = |s(ky) + 1|/Sy,  s(k) = k < N/2 ? k : k - N.
// The carrier exp(i k z_l) cancels the constant part of H, leaving the phase
//   (z_l lambda / 2)(fx^2 + fy^2) + (z_l/lambda) p(u),
//   p(u) = u^2/8 + u^3/16 + 5u^4/128 + 7u^5/256,  u = lambda^2 (fx^2 + fy^2),
// whose first term is reduced per axis in FP64.
//
// Passes (per colour; the depth index is shared):
//   dm_index  depth index of every pixel by a 256-entry table of the u8 depth
//             (built on the host; the reference makes one pass per level);
//   dm_cols   a workgroup owns 4 scene columns and a range of layers; two
//             real columns are packed into one complex column, transformed
//             along y (length Ny) and separated; only rows ky <= Ny/2 are
//             stored (Hermitian symmetry of real input), with a presence flag
//             per (layer, column group) -- absent groups are never written;
//   dm_rows   a workgroup owns the row pair (ky, Ny-ky): per layer it
//             transforms row ky along x once and obtains row Ny-ky from the
//             same transform by symmetry, multiplies both by H and
//             accumulates over all layers in registers; then the encode crop,
//             the inverse row transforms, and the scene columns are stored.
//             Rows 0 and Ny/2 are real (transforms of real columns at the DC
//             and Nyquist frequencies): workgroup 0 packs them into one
//             complex row and separates the transform, so the pass is Ny/2
//             workgroups (a multiple of the CU count) instead of Ny/2 + 1;
//   dm_final  inverse column transform, 1/(Nx Ny), the shift phase and the
//             real part (or the complex field) in the scene window.
#include <cmath>
#include <vector>

#include "fft_engine.h"
#include "fft_kernels.h"
#include "fft_mixed.h"

namespace {

constexpr int GROUP = 4;        // scene columns per dm_cols workgroup (2 complex lines)
constexpr int MAX_LAYERS = 4096;

struct DmArgs {
    int w, h, Nx, Ny, ox, oy;
    int L, lo, hi;                 // layers 1..L, rendered range [lo, hi]
    int ngroups;
    const unsigned char* img;      // h x w, this colour
    const unsigned short* idx;     // h x w
    float2* I;                     // [L][ngroups][Ny/2 + 1][GROUP]: each column group's
                                   // rows contiguous (I_l(ky, x) at group x / GROUP)
    unsigned char* flags;          // [L][ngroups]
    unsigned char* any;            // [L]
    unsigned* fbits;               // [L][fwpl]: the flags as bits (group g: word g/32, bit g%32)
    int fwpl;                      // words per layer of fbits
    int fbits_lds;                 // dm_rows_h stages fbits in LDS and loads present groups only
    const double* lc;              // per layer: cx, cy, K (3 doubles)
    float ux, uy;                  // lambda^2 / Sx^2, lambda^2 / Sy^2
    int exact;                     // FP64 transfer path
    int horner;                    // equally spaced layers: chunked recurrence in dm_rows_h
    double dcx, dcy, dK;           // layer-step constants (z_{l+1} - z_l replacing z_l)
    double lam2_sx2, lam2_sy2, inv_lambda;
    float2 rho;
    int encode, cy1, cy2;          // crop range (centred rows) when encoding
    float2* T;                     // [Ny][w] after the inverse row transform
    float2* spectrum;              // optional centred spectrum [Ny][Nx]
    double scale;                  // 1/(Nx Ny)
    double* fringe;                // h x w (double) or
    float* fringe32;               // h x w (float) or
    float2* field;                 // h x w complex
};

__device__ __forceinline__ int sgnidx(int k, int n) { return k < n / 2 ? k : k - n; }

// ---- depth index -------------------------------------------------------------

// The depth index depends on the u8 depth value only: the host forms the
// 256-entry table (depth_index_table) and every pixel is one lookup.
__global__ void dm_index(const unsigned char* dep, int n, const double* table, unsigned short* idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    idx[i] = (unsigned short)table[dep[i]];
}

__global__ void dm_zero(unsigned char* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---- column pass ---------------------------------------------------------------

template <int LY>
__global__ __launch_bounds__(2 * (LY / 16)) void dm_cols(DmArgs a, int layers_per_wg,
                                                         const float2* __restrict__ twy) {
    constexpr int TPL = LY / 16;
    constexpr int NT = 2 * TPL;
    constexpr int PL = fft::padded_len(LY);
    extern __shared__ unsigned char smem[];
    float2* lines = reinterpret_cast<float2*>(smem);                 // 2 x PL
    unsigned* pix = reinterpret_cast<unsigned*>(lines + 2 * PL);     // h x GROUP: idx << 8 | img
    unsigned* mask = pix + (size_t)a.h * GROUP;                      // layer bitmask
    // XCD-aware: neighbouring column groups (whose 32-byte row segments share
    // cache lines of I) run on the same XCD at the same time
    const int lin = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int g = lin % gridDim.x;
    const int x0 = g * GROUP;
    const int l_begin = a.lo + lin / gridDim.x * layers_per_wg;
    const int l_end = min(a.hi, l_begin + layers_per_wg - 1);
    const int nwords = (a.L + 32) / 32;
    const int p = threadIdx.x / TPL, t = threadIdx.x % TPL;
    float2* line = lines + p * PL;
    fft::Twiddles<LY> Tw;
    Tw.load(twy, t);
    for (int i = threadIdx.x; i < nwords; i += NT) mask[i] = 0u;
    __syncthreads();
    for (int e = threadIdx.x; e < a.h * GROUP; e += NT) {
        const int c = e % GROUP, r = e / GROUP;
        const int x = x0 + c;
        unsigned v = 0u;
        if (x < a.w) {
            const size_t o = (size_t)r * a.w + x;
            const unsigned im = a.img[o];
            const unsigned id = a.idx[o];
            // alpha = [img > 0] (CPU.cpp:61): a zero pixel contributes nothing
            if (im > 0u && id >= (unsigned)l_begin && id <= (unsigned)l_end) {
                v = (id << 8) | im;
                atomicOr(&mask[id >> 5], 1u << (id & 31));
            }
        }
        pix[e] = v;
    }
    __syncthreads();
    const size_t gstride = (size_t)(a.Ny / 2 + 1) * GROUP;
    const size_t lstride = gstride * a.ngroups;
    for (int l = l_begin; l <= l_end; ++l) {
        if (!((mask[l >> 5] >> (l & 31)) & 1u)) continue;
        // two real columns (x0 + 2p, x0 + 2p + 1) as one complex column
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int i = t + m * TPL;
            const int r = i - a.oy;
            float2 v = make_float2(0.f, 0.f);
            if (r >= 0 && r < a.h) {
                const unsigned wa = pix[r * GROUP + 2 * p], wb = pix[r * GROUP + 2 * p + 1];
                if ((int)(wa >> 8) == l) v.x = (float)(wa & 255u) * (1.f / 255.f);
                if ((int)(wb >> 8) == l) v.y = (float)(wb & 255u) * (1.f / 255.f);
            }
            line[fft::lane_pad<LY>(t, m)] = v;
        }
        __syncthreads();
        fft::line<LY, -1>(line, Tw, t);
        // separate: A(k) = (Z(k) + conj Z(-k))/2, B(k) = (Z(k) - conj Z(-k))/(2i)
        float2* dst = a.I + (size_t)(l - 1) * lstride + (size_t)g * gstride;
        for (int e = threadIdx.x; e < (LY / 2 + 1) * GROUP; e += NT) {
            const int c = e % GROUP, ky = e / GROUP;
            const int x = x0 + c;
            if (x >= a.w) continue;
            const float2* ln = lines + (c >> 1) * PL;
            const float2 Z = ln[fft::pad16(ky)];
            const float2 Zm = ln[fft::pad16((LY - ky) & (LY - 1))];
            float2 v;
            if ((c & 1) == 0)
                v = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
            else
                v = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
            dst[e] = v;
        }
        if (threadIdx.x == 0) {
            a.flags[(size_t)(l - 1) * a.ngroups + g] = 1;
            a.any[l - 1] = 1;
            atomicOr(&a.fbits[(size_t)(l - 1) * a.fwpl + (g >> 5)], 1u << (g & 31));
        }
        __syncthreads();
    }
}

// ---- row pass ------------------------------------------------------------------

// Z = FFT(a + i b) of two real rows a, b; on entry z = Z(k), zc = conj Z(-k).
// On exit z = A(k) = (Z(k) + conj Z(-k))/2 and zc = B(k) = (Z(k) - conj Z(-k))/(2i).
__device__ __forceinline__ void split_real_pair(float2& z, float2& zc) {
    const float dx = z.x - zc.x, dy = z.y - zc.y;
    z = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    zc = make_float2(0.5f * dy, -0.5f * dx);
}

template <bool EXACT>
__device__ __forceinline__ float2 dm_transfer(const DmArgs& a, float tx, double ty, float usx, int sy,
                                              double K, float Kf, double sx2d) {
    const float u = fmaf(a.uy, (float)(sy * sy), usx);
    if (u >= 1.f) return make_float2(0.f, 0.f);
    float ph;
    if (!EXACT) {
        const float pu = u * u * fmaf(u, fmaf(u, fmaf(u, 7.f / 256.f, 5.f / 128.f), 1.f / 16.f), 0.125f);
        ph = tx + (float)ty + Kf * pu;
    } else {
        const double ud = a.lam2_sx2 * sx2d + a.lam2_sy2 * (double)sy * sy;
        if (ud >= 1.0) return make_float2(0.f, 0.f);
        double t = K * ud / (1.0 + sqrt(1.0 - ud));
        t -= rint(t);
        ph = (float)t;
    }
    return make_float2(cos_turns(ph), sin_turns(ph));
}

// Optional centred spectrum of one accumulated row, the encode crop, the
// inverse row transform and the store of the scene columns into T.
template <int LX>
__device__ __forceinline__ void finish_row(const DmArgs& a, const float2 (&acc)[16], int ky, float2* line,
                                           const float2* __restrict__ tw, int t) {
    constexpr int TPL = LX / 16;
    const int yc = (ky + a.Ny / 2) & (a.Ny - 1);
    if (a.spectrum) {
        // U_ref(y, x) = (-1)^(kx + ky) A(ky, kx), y, x centred indices
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int kx = t + m * TPL;
            const int xc = (kx + LX / 2) & (LX - 1);
            const float sg = ((kx + ky) & 1) ? -1.f : 1.f;
            a.spectrum[(size_t)yc * LX + xc] = make_float2(sg * acc[m].x, sg * acc[m].y);
        }
    }
    const bool kept = !a.encode || (yc >= a.cy1 && yc <= a.cy2);
    if (!kept) return;  // uniform over the workgroup
#pragma unroll
    for (int m = 0; m < 16; ++m) line[fft::lane_pad<LX>(t, m)] = acc[m];
    __syncthreads();
    fft::line_tw<LX, +1>(line, tw, t);
    float2* dst = a.T + (size_t)ky * a.w;
    for (int x = t; x < a.w; x += TPL) dst[x] = line[fft::pad16(x + a.ox)];
    __syncthreads();
}

template <int LX, bool EXACT>
__global__ __launch_bounds__(LX / 16) void dm_rows(DmArgs a, const float2* __restrict__ twx) {
    constexpr int TPL = LX / 16;
    constexpr int PL = fft::padded_len(LX);
    __shared__ float2 line[PL];
    const int t = threadIdx.x;
    const int kya = xcd_remap(blockIdx.x, gridDim.x);  // 0 .. Ny/2 - 1
    const bool pair0 = kya == 0;              // real rows 0 and Ny/2 packed as re/im
    const int kyb = pair0 ? a.Ny / 2 : a.Ny - kya;
    const int sya = sgnidx(kya, a.Ny) + 1;
    const int syb = sgnidx(kyb, a.Ny) + 1;
    const size_t nyq = (size_t)(a.Ny / 2) * GROUP;  // row Ny/2 relative to row 0
    const size_t gstride = (size_t)(a.Ny / 2 + 1) * GROUP;
    fft::Twiddles<LX> Tw;
    Tw.load(twx, t);
    float2 acca[16], accb[16];
    float usx[16];
    double sx2[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        acca[m] = accb[m] = make_float2(0.f, 0.f);
        const int sx = sgnidx(t + m * TPL, LX);
        sx2[m] = (double)sx * sx;
        usx[m] = a.ux * (float)(sx * sx);
    }
    const size_t lstride = gstride * a.ngroups;
    for (int l = a.lo; l <= a.hi; ++l) {
        if (!a.any[l - 1]) continue;
        const float2* src = a.I + (size_t)(l - 1) * lstride + (size_t)kya * GROUP;
        const unsigned char* fl = a.flags + (size_t)(l - 1) * a.ngroups;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int x = t + m * TPL - a.ox;
            float2 v = make_float2(0.f, 0.f);
            if (x >= 0 && x < a.w && fl[x / GROUP]) {
                const size_t o = (size_t)(x / GROUP) * gstride + x % GROUP;
                v = src[o];
                if (pair0) v.y = src[nyq + o].x;
            }
            line[fft::lane_pad<LX>(t, m)] = v;
        }
        __syncthreads();
        fft::line<LX, -1>(line, Tw, t);
        const double cx = a.lc[3 * (l - 1)], cy = a.lc[3 * (l - 1) + 1], K = a.lc[3 * (l - 1) + 2];
        const float Kf = (float)K;
        double tya = cy * (double)sya * sya;
        tya -= rint(tya);
        double tyb = cy * (double)syb * syb;
        tyb -= rint(tyb);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int kx = t + m * TPL;
            double txd = cx * sx2[m];
            txd -= rint(txd);
            const float tx = (float)txd;
            float2 gv = line[fft::lane_pad<LX>(t, m)];
            const float2 gm = line[fft::pad16((LX - kx) & (LX - 1))];
            float2 gmc = make_float2(gm.x, -gm.y);
            if (pair0) split_real_pair(gv, gmc);
            const float2 ha = dm_transfer<EXACT>(a, tx, tya, usx[m], sya, K, Kf, sx2[m]);
            acca[m] = fft::cadd(acca[m], fft::cmul(ha, gv));
            const float2 hb = dm_transfer<EXACT>(a, tx, tyb, usx[m], syb, K, Kf, sx2[m]);
            accb[m] = fft::cadd(accb[m], fft::cmul(hb, gmc));
        }
        __syncthreads();
    }
    // rho (the same for every layer), optional centred spectrum, encode crop,
    // inverse row transform
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        acca[m] = fft::cmul(acca[m], a.rho);
        accb[m] = fft::cmul(accb[m], a.rho);
    }
    finish_row<LX>(a, acca, kya, line, twx, t);
    finish_row<LX>(a, accb, kyb, line, twx, t);
}

// Row pass for equally spaced layers (no field lens).  The transfer of
// layer l is H_l = exp(i 2 pi z_l Phi(kx, ky)) with one Phi for all layers,
// so over a chunk of 16 consecutive layers starting at c,
//   sum_j H_{c+j} G_{c+j} = H_c sum_j R^j G_{c+j},   R = exp(i 2 pi dz Phi),
// evaluated by Horner from the top layer down (T = T R + G_l): one complex
// multiply-add per element and layer instead of the phase reduction, series
// and sin/cos, which are evaluated once per chunk for the anchor H_c (the
// recurrence's rounding stays within ~16 ulp).  The row transforms are
// radix-8 with LX/8 threads (8 elements per thread) so that R and T of both
// rows fit in registers next to the transform; the chunk sums S accumulate
// in LDS, each element touched only by its owner thread.  The twiddle table
// and a bitmask of the present layers are copied to LDS at the start, so
// that inside the layer loop only the prefetched rows are global loads.
template <int LX>
__host__ __device__ constexpr size_t rows_h_lds() {
    return (fft::padded_len(LX) + 2 * (size_t)LX + fft::table_len8(LX)) * sizeof(float2) +
           (size_t)(MAX_LAYERS / 32) * sizeof(unsigned);
}

// finish_row for the radix-8 layout (8 elements per thread)
template <int LX>
__device__ __forceinline__ void finish_row8(const DmArgs& a, const float2 (&acc)[8], int ky, float2* line,
                                            const float2* __restrict__ tw8, int t) {
    constexpr int TPL = LX / 8;
    const int yc = (ky + a.Ny / 2) & (a.Ny - 1);
    if (a.spectrum) {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int kx = t + m * TPL;
            const int xc = (kx + LX / 2) & (LX - 1);
            const float sg = ((kx + ky) & 1) ? -1.f : 1.f;
            a.spectrum[(size_t)yc * LX + xc] = make_float2(sg * acc[m].x, sg * acc[m].y);
        }
    }
    const bool kept = !a.encode || (yc >= a.cy1 && yc <= a.cy2);
    if (!kept) return;  // uniform over the workgroup
#pragma unroll
    for (int m = 0; m < 8; ++m) line[fft::pad16(t + m * TPL)] = acc[m];
    __syncthreads();
    fft::line_r8<LX, +1>(line, tw8, t);
    float2* dst = a.T + (size_t)ky * a.w;
    for (int x = t; x < a.w; x += TPL) dst[x] = line[fft::pad16(x + a.ox)];
    __syncthreads();
}

// PAIR0: workgroup 0, the real rows 0 and Ny/2 packed as re/im.  FBITS:
// the presence bits of every (layer, group) fit in LDS (a.fbits_lds).
// Copies n elements of a table to LDS with NT threads, eight loads per
// thread requested before any is stored (a plain loop waits for each load
// before the next is issued).
template <int NT, typename E>
__device__ __forceinline__ void lds_copy(E* dst, const E* __restrict__ src, int n, int t) {
    for (int i0 = 0; i0 < n; i0 += 8 * NT) {
        E v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = i0 + t + j * NT;
            if (i < n) v[j] = src[i];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = i0 + t + j * NT;
            if (i < n) dst[i] = v[j];
        }
    }
}

template <int LX, bool PAIR0, bool FBITS>
__device__ __forceinline__ void rows_h_body(const DmArgs& a, const float2* __restrict__ tw8, float2* sm, int kya) {
    constexpr int TPL = LX / 8;
    constexpr int PL = fft::padded_len(LX);
    constexpr int CH = 16;
    constexpr int NTW = fft::table_len8(LX);
    float2* line = sm;
    float2* Sa = sm + PL;
    float2* Sb = Sa + LX;
    float2* tws = Sb + LX;
    unsigned* amask = reinterpret_cast<unsigned*>(tws + NTW);
    // presence bits of every (layer, column group), when they fit in LDS:
    // then only present groups' rows of I are loaded
    unsigned* fb = amask + MAX_LAYERS / 32;
    const int t = threadIdx.x;
    lds_copy<TPL>(tws, tw8, NTW, t);
    for (int i = t; i < MAX_LAYERS / 32; i += TPL) amask[i] = 0u;
    if constexpr (FBITS) lds_copy<TPL>(fb, a.fbits, a.L * a.fwpl, t);
    __syncthreads();
    for (int i = t; i < a.L; i += TPL)
        if (a.any[i]) atomicOr(&amask[i >> 5], 1u << (i & 31));
    __syncthreads();
    auto present = [&](int l) -> bool { return (amask[(l - 1) >> 5] >> ((l - 1) & 31)) & 1u; };
    const int kyb = PAIR0 ? a.Ny / 2 : a.Ny - kya;
    const int sya = sgnidx(kya, a.Ny) + 1;
    const int syb = sgnidx(kyb, a.Ny) + 1;
    const size_t nyq = (size_t)(a.Ny / 2) * GROUP;
    const size_t gstride = (size_t)(a.Ny / 2 + 1) * GROUP;
    // Thread t owns row a at kx = t + m TPL and row b at kb = -kx mod LX:
    // row b's transform at kb is conj Z(kx) of the thread's own element of
    // row a's transform Z, and the transfer depends on kx only through kx^2,
    // which is the same at kb, so one set of per-element factors serves both.
    float2 Ra[8], Rb[8];
    {
        double tya = a.dcy * (double)sya * sya;
        tya -= rint(tya);
        double tyb = a.dcy * (double)syb * syb;
        tyb -= rint(tyb);
        const float dKf = (float)a.dK;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int kx = t + m * TPL;
            const int sx = sgnidx(kx, LX);
            const double sx2 = (double)sx * sx;
            double txd = a.dcx * sx2;
            txd -= rint(txd);
            const float usx = a.ux * (float)(sx * sx);
            Ra[m] = dm_transfer<false>(a, (float)txd, tya, usx, sya, a.dK, dKf, sx2);
            Rb[m] = dm_transfer<false>(a, (float)txd, tyb, usx, syb, a.dK, dKf, sx2);
            Sa[kx] = Sb[(LX - kx) & (LX - 1)] = make_float2(0.f, 0.f);
        }
    }
    const size_t lstride = gstride * a.ngroups;
    // Layers are taken chunk by chunk, top layer first within a chunk.  The
    // row of the next present layer in that order is loaded into registers
    // (values and presence flags) before the current layer's transform, so
    // its latency hides behind the transform.
    auto next_layer = [&](int l) -> int {
        for (;;) {
            const int c = a.lo + (l - a.lo) / CH * CH;
            l = l > c ? l - 1 : (c + CH <= a.hi ? min(a.hi, c + 2 * CH - 1) : 0);
            if (l == 0 || present(l)) return l;
        }
    };
    float2 nv[8];
    unsigned nf;  // presence of the prefetched elements, bit m for element m
    auto fetch = [&](int l) {
        const float2* src = a.I + (size_t)(l - 1) * lstride + (size_t)kya * GROUP;
        if constexpr (FBITS) {
            // flags from LDS first; then range-checked buffer loads whose
            // offset is out of range for absent groups (no memory access, a
            // zero result, no branch): the rows of absent groups were never
            // written and are not read, and the loaded values need no mask
            nf = 0xFFu;
            const unsigned* fl = fb + (size_t)(l - 1) * a.fwpl;
            const auto rs = oph_buf_rsrc(a.I + (size_t)(l - 1) * lstride, (unsigned)(lstride * sizeof(float2)));
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int x = t + m * TPL - a.ox;
                const int gx = x >> 2;  // x / GROUP for x >= 0
                const bool on = (x >= 0 && x < a.w) && ((fl[gx >> 5] >> (gx & 31)) & 1u);
                const unsigned o = (unsigned)(((size_t)gx * gstride + (x & 3) + (size_t)kya * GROUP) * sizeof(float2));
                if constexpr (PAIR0) {
                    const float2 lo = oph_buf_load_f2(rs, on ? o : OPH_BUF_OOB);
                    const float2 hi = oph_buf_load_f2(rs, on ? o + (unsigned)(nyq * sizeof(float2)) : OPH_BUF_OOB);
                    nv[m] = make_float2(lo.x, hi.x);
                } else {
                    nv[m] = oph_buf_load_f2(rs, on ? o : OPH_BUF_OOB);
                }
            }
            return;
        }
        const unsigned char* fl = a.flags + (size_t)(l - 1) * a.ngroups;
        nf = 0u;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int x = t + m * TPL - a.ox;
            const bool in = x >= 0 && x < a.w;
            const size_t o = (size_t)(x / GROUP) * gstride + x % GROUP;
            nf |= (in && fl[x / GROUP]) ? 1u << m : 0u;
            if constexpr (PAIR0)
                nv[m] = in ? make_float2(src[o].x, src[nyq + o].x) : make_float2(0.f, 0.f);
            else
                nv[m] = in ? src[o] : make_float2(0.f, 0.f);
        }
    };
    int pl = min(a.hi, a.lo + CH - 1);
    if (!present(pl)) pl = next_layer(pl);
    if (pl) fetch(pl);
    for (int cs = a.lo; cs <= a.hi; cs += CH) {
        const int ce = min(a.hi, cs + CH - 1);
        bool anyc = false;
        for (int l = cs; l <= ce; ++l) anyc |= present(l);
        if (!anyc) continue;
        float2 Ta[8], Tb[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) Ta[m] = Tb[m] = make_float2(0.f, 0.f);
        for (int l = ce; l >= cs; --l) {
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                Ta[m] = fft::cmul(Ta[m], Ra[m]);
                Tb[m] = fft::cmul(Tb[m], Rb[m]);
            }
            if (!present(l)) continue;
            // l is the prefetched layer
            float2 v[8];
#pragma unroll
            for (int m = 0; m < 8; ++m) v[m] = ((nf >> m) & 1u) ? nv[m] : make_float2(0.f, 0.f);
            pl = next_layer(l);
            if (pl) fetch(pl);
            __syncthreads();  // the previous transform's readers are done with the line
            fft::line_r8_regs<LX, -1>(v, line, tws, t);
            if constexpr (PAIR0) {
                // A(kx) = (Z(kx) + conj Z(-kx))/2 for row 0 and, row Ny/2
                // being real, B(-kx) = conj B(kx): exchange through the line
                __syncthreads();
#pragma unroll
                for (int m = 0; m < 8; ++m) line[fft::pad16(t + m * TPL)] = v[m];
                __syncthreads();
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int kx = t + m * TPL;
                    const float2 gm = line[fft::pad16((LX - kx) & (LX - 1))];
                    float2 z = v[m], zc = make_float2(gm.x, -gm.y);
                    split_real_pair(z, zc);
                    Ta[m] = fft::cadd(Ta[m], z);
                    Tb[m] = fft::cadd(Tb[m], make_float2(zc.x, -zc.y));
                }
            } else {
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    Ta[m] = fft::cadd(Ta[m], v[m]);
                    Tb[m] = fft::cadd(Tb[m], make_float2(v[m].x, -v[m].y));
                }
            }
        }
        // anchor: S += H_cs T
        const double cx = a.lc[3 * (cs - 1)], cy = a.lc[3 * (cs - 1) + 1], K = a.lc[3 * (cs - 1) + 2];
        const float Kf = (float)K;
        double tya = cy * (double)sya * sya;
        tya -= rint(tya);
        double tyb = cy * (double)syb * syb;
        tyb -= rint(tyb);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int kx = t + m * TPL;
            const int sx = sgnidx(kx, LX);
            const double sx2 = (double)sx * sx;
            double txd = cx * sx2;
            txd -= rint(txd);
            const float usx = a.ux * (float)(sx * sx);
            const float2 ha = dm_transfer<false>(a, (float)txd, tya, usx, sya, K, Kf, sx2);
            Sa[kx] = fft::cadd(Sa[kx], fft::cmul(ha, Ta[m]));
            const float2 hb = dm_transfer<false>(a, (float)txd, tyb, usx, syb, K, Kf, sx2);
            const int kb = (LX - kx) & (LX - 1);
            Sb[kb] = fft::cadd(Sb[kb], fft::cmul(hb, Tb[m]));
        }
    }
    __syncthreads();  // Sb entries were written by the threads owning -kx
    float2 acca[8], accb[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int kx = t + m * TPL;
        acca[m] = fft::cmul(Sa[kx], a.rho);
        accb[m] = fft::cmul(Sb[kx], a.rho);
    }
    finish_row8<LX>(a, acca, kya, line, tws, t);
    finish_row8<LX>(a, accb, kyb, line, tws, t);
}

// (two waves per SIMD: 256 registers, the budget of two 2048-point
// workgroups per CU or one 4096-point workgroup)
template <int LX, bool FBITS>
__global__ __launch_bounds__(LX / 8, 2) void dm_rows_h(DmArgs a, const float2* __restrict__ tw8) {
    extern __shared__ float2 sm[];
    // XCD-aware: neighbouring rows, which share the cache lines of I, run on
    // the same XCD at the same time
    const int kya = xcd_remap(blockIdx.x, gridDim.x);  // 0 .. Ny/2 - 1
    if (kya == 0)
        rows_h_body<LX, true, FBITS>(a, tw8, sm, 0);
    else
        rows_h_body<LX, false, FBITS>(a, tw8, sm, kya);
}

// ---- final column pass ---------------------------------------------------------

template <int LY, int B>
__global__ __launch_bounds__(B * (LY / 16)) void dm_final(DmArgs a, const float2* __restrict__ twy) {
    constexpr int TPL = LY / 16;
    constexpr int PL = fft::padded_len(LY);
    extern __shared__ float2 lds[];
    const int nthr = B * TPL;
    const int g = xcd_remap(blockIdx.x, gridDim.x);
    const int c0 = g * B;
    const int b = threadIdx.x / TPL, t = threadIdx.x % TPL;
    fft::Twiddles<LY> Tw;
    Tw.load(twy, t);
    // 16 elements per thread (LY B / nthr), all requested before any is
    // stored (range-checked buffer loads: dropped rows and columns beyond
    // the scene read zeros; T is at most 8192 x 8192 x 8 bytes)
    {
        const auto rt = oph_buf_rsrc(a.T, (unsigned)((size_t)a.Ny * a.w * sizeof(float2)));
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int e = threadIdx.x + m * nthr;
            const int ky = e / B, x = c0 + e % B;
            const int yc = (ky + a.Ny / 2) & (a.Ny - 1);
            const bool kept = !a.encode || (yc >= a.cy1 && yc <= a.cy2);
            v[m] = oph_buf_load_f2(rt, x < a.w && kept ? ((unsigned)ky * (unsigned)a.w + (unsigned)x) * 8u : OPH_BUF_OOB);
        }
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int e = threadIdx.x + m * nthr;
            lds[(e % B) * PL + fft::pad16(e / B)] = v[m];
        }
    }
    __syncthreads();
    fft::line<LY, +1>(lds + b * PL, Tw, t);
    const float sc = (float)a.scale;
    for (int e = threadIdx.x; e < a.h * B; e += nthr) {
        const int bb = e % B, r = e / B;
        const int x = c0 + bb;
        if (x >= a.w) continue;
        const float2 v0 = lds[bb * PL + fft::pad16(a.oy + r)];
        const float2 v = make_float2(v0.x * sc, v0.y * sc);
        const size_t o = (size_t)r * a.w + x;
        if (a.encode) {
            // exp(i 2 pi yy/(4 py)), yy = Sy/2 - py r - py, Sy = h py (CPU.cpp:376-390)
            double tt = a.h / 8.0 - (r + 1) / 4.0;
            tt -= rint(tt);
            const float ph = (float)tt;
            const float re = v.x * cos_turns(ph) - v.y * sin_turns(ph);
            if (a.fringe) a.fringe[o] = re;
            if (a.fringe32) a.fringe32[o] = re;
        } else {
            a.field[o] = v;
        }
    }
}

// ---- general lengths -------------------------------------------------------------
// The same three passes for grids whose sides are 2^a 3^b 5^c but not powers
// of two (the shipped 11520 x 2160 configuration, HD scenes transformed
// unpadded): mixed-radix line transforms (fft_mixed.h) by NT =
// fftm::threads(N) threads per line, thread t owning elements t + m NT
// (m < fftm::ept(N)), index arithmetic mod N (N even: the fftshifts are
// still the (-1)^k factors above).  No chunked recurrence: every layer's
// transfer is evaluated (dm_transfer).

#define OPH_DM_MIXED_LENGTHS(X) X(540) X(720) X(1080) X(1920) X(2160) X(3840) X(11520)

bool dm_mixed_length(int n) {
    switch (n) {
#define OPH_DMM_CASE(v) case v:
        OPH_DM_MIXED_LENGTHS(OPH_DMM_CASE)
#undef OPH_DMM_CASE
        return true;
        default: return false;
    }
}

template <int LY>
__global__ __launch_bounds__(2 * fftm::threads(LY)) void dm_cols_m(DmArgs a, int layers_per_wg,
                                                                 const float2* __restrict__ twy) {
    constexpr int NT = fftm::threads(LY);
    constexpr int EPT = fftm::ept(LY);
    constexpr int PL = fftm::padded_len(LY);
    extern __shared__ unsigned char smem[];
    float2* lines = reinterpret_cast<float2*>(smem);                 // 2 x PL
    unsigned* pix = reinterpret_cast<unsigned*>(lines + 2 * PL);     // h x GROUP: idx << 8 | img
    unsigned* mask = pix + (size_t)a.h * GROUP;                      // layer bitmask
    const int lin = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int g = lin % gridDim.x;
    const int x0 = g * GROUP;
    const int l_begin = a.lo + lin / gridDim.x * layers_per_wg;
    const int l_end = min(a.hi, l_begin + layers_per_wg - 1);
    const int nwords = (a.L + 32) / 32;
    const int p = threadIdx.x / NT, t = threadIdx.x % NT;
    float2* line = lines + p * PL;
    for (int i = threadIdx.x; i < nwords; i += 2 * NT) mask[i] = 0u;
    __syncthreads();
    for (int e = threadIdx.x; e < a.h * GROUP; e += 2 * NT) {
        const int c = e % GROUP, r = e / GROUP;
        const int x = x0 + c;
        unsigned v = 0u;
        if (x < a.w) {
            const size_t o = (size_t)r * a.w + x;
            const unsigned im = a.img[o];
            const unsigned id = a.idx[o];
            if (im > 0u && id >= (unsigned)l_begin && id <= (unsigned)l_end) {
                v = (id << 8) | im;
                atomicOr(&mask[id >> 5], 1u << (id & 31));
            }
        }
        pix[e] = v;
    }
    __syncthreads();
    const size_t gstride = (size_t)(a.Ny / 2 + 1) * GROUP;
    const size_t lstride = gstride * a.ngroups;
    for (int l = l_begin; l <= l_end; ++l) {
        if (!((mask[l >> 5] >> (l & 31)) & 1u)) continue;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int i = t + m * NT;
            if (EPT * NT == LY || i < LY) {
                const int r = i - a.oy;
                float2 v = make_float2(0.f, 0.f);
                if (r >= 0 && r < a.h) {
                    const unsigned wa = pix[r * GROUP + 2 * p], wb = pix[r * GROUP + 2 * p + 1];
                    if ((int)(wa >> 8) == l) v.x = (float)(wa & 255u) * (1.f / 255.f);
                    if ((int)(wb >> 8) == l) v.y = (float)(wb & 255u) * (1.f / 255.f);
                }
                line[fft::pad16(i)] = v;
            }
        }
        __syncthreads();
        fftm::line<LY, NT, -1>(line, twy, t);
        float2* dst = a.I + (size_t)(l - 1) * lstride + (size_t)g * gstride;
        for (int e = threadIdx.x; e < (LY / 2 + 1) * GROUP; e += 2 * NT) {
            const int c = e % GROUP, ky = e / GROUP;
            const int x = x0 + c;
            if (x >= a.w) continue;
            const float2* ln = lines + (c >> 1) * PL;
            const float2 Z = ln[fft::pad16(ky)];
            const float2 Zm = ln[fft::pad16(ky == 0 ? 0 : LY - ky)];
            float2 v;
            if ((c & 1) == 0)
                v = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
            else
                v = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
            dst[e] = v;
        }
        if (threadIdx.x == 0) {
            a.flags[(size_t)(l - 1) * a.ngroups + g] = 1;
            a.any[l - 1] = 1;
            atomicOr(&a.fbits[(size_t)(l - 1) * a.fwpl + (g >> 5)], 1u << (g & 31));
        }
        __syncthreads();
    }
}

template <int LX>
__device__ __forceinline__ void finish_row_m(const DmArgs& a, const float2 (&acc)[fftm::ept(LX)], int ky,
                                             float2* line, const float2* __restrict__ tw, int t) {
    constexpr int NT = fftm::threads(LX), EPT = fftm::ept(LX);
    const int yc = (ky + a.Ny / 2) % a.Ny;
    if (a.spectrum) {
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int kx = t + m * NT;
            if (EPT * NT == LX || kx < LX) {
                const int xc = (kx + LX / 2) % LX;
                const float sg = ((kx + ky) & 1) ? -1.f : 1.f;
                a.spectrum[(size_t)yc * LX + xc] = make_float2(sg * acc[m].x, sg * acc[m].y);
            }
        }
    }
    const bool kept = !a.encode || (yc >= a.cy1 && yc <= a.cy2);
    if (!kept) return;  // uniform over the workgroup
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int kx = t + m * NT;
        if (EPT * NT == LX || kx < LX) line[fft::pad16(kx)] = acc[m];
    }
    __syncthreads();
    fftm::line<LX, NT, +1>(line, tw, t);
    float2* dst = a.T + (size_t)ky * a.w;
    for (int x = t; x < a.w; x += NT) dst[x] = line[fft::pad16(x + a.ox)];
    __syncthreads();
}

template <int LX, bool EXACT>
__global__ __launch_bounds__(fftm::threads(LX)) void dm_rows_m(DmArgs a, const float2* __restrict__ twx) {
    constexpr int NT = fftm::threads(LX), EPT = fftm::ept(LX);
    extern __shared__ float2 line[];  // fftm::padded_len(LX)
    const int t = threadIdx.x;
    const int kya = xcd_remap(blockIdx.x, gridDim.x);  // 0 .. Ny/2 - 1
    const bool pair0 = kya == 0;              // real rows 0 and Ny/2 packed as re/im
    const int kyb = pair0 ? a.Ny / 2 : a.Ny - kya;
    const int sya = sgnidx(kya, a.Ny) + 1;
    const int syb = sgnidx(kyb, a.Ny) + 1;
    const size_t nyq = (size_t)(a.Ny / 2) * GROUP;
    const size_t gstride = (size_t)(a.Ny / 2 + 1) * GROUP;
    // (the per-element x factors are recomputed per layer: the two rows'
    // accumulators already take most of the registers of a 768-thread line)
    float2 acca[EPT], accb[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) acca[m] = accb[m] = make_float2(0.f, 0.f);
    const size_t lstride = gstride * a.ngroups;
    // a layer's slice of I addressed by 32-bit buffer offsets when it is
    // below 2 GiB (else plain loads with a branch per element)
    const bool bufs = lstride * sizeof(float2) < ((size_t)1 << 31);
    for (int l = a.lo; l <= a.hi; ++l) {
        if (!a.any[l - 1]) continue;
        // the thread index through a zero the compiler cannot see through:
        // the transforms' LDS addresses are then formed per layer instead of
        // being hoisted out of the loop into ~100 registers
        int zero;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
        const int tl = t + zero;
        if (bufs) {
            // range-checked buffer loads, eight elements at a time: the
            // presence flags of their column groups, then the values, whose
            // offset is out of range (a zero, no memory access) for absent
            // groups and columns outside the image -- each batch is
            // requested before any of it is used, with no branch per element
            const auto rf = oph_buf_rsrc(a.flags + (size_t)(l - 1) * a.ngroups, (unsigned)a.ngroups);
            const auto ri = oph_buf_rsrc(a.I + (size_t)(l - 1) * lstride, (unsigned)(lstride * sizeof(float2)));
            constexpr int CHB = 8;
#pragma unroll
            for (int m0 = 0; m0 < EPT; m0 += CHB) {
                unsigned char f[CHB];
                unsigned o[CHB];
#pragma unroll
                for (int j = 0; j < CHB; ++j) {
                    const int kx = tl + (m0 + j) * NT;
                    const int x = kx - a.ox;
                    const bool in = m0 + j < EPT && (EPT * NT == LX || kx < LX) && x >= 0 && x < a.w;
                    f[j] = __builtin_amdgcn_raw_buffer_load_b8(rf, in ? (unsigned)(x / GROUP) : OPH_BUF_OOB, 0, 0);
                    o[j] = in ? (unsigned)(((size_t)(x / GROUP) * gstride + x % GROUP + (size_t)kya * GROUP) *
                                           sizeof(float2))
                              : OPH_BUF_OOB;
                }
                float2 v[CHB];
                float vy[CHB];
#pragma unroll
                for (int j = 0; j < CHB; ++j) {
                    const unsigned oj = f[j] ? o[j] : OPH_BUF_OOB;
                    v[j] = oph_buf_load_f2(ri, oj);
                    // the packed Nyquist row (workgroup 0 only; elsewhere an
                    // out-of-range offset: no memory access)
                    const unsigned oy = pair0 && oj != OPH_BUF_OOB ? oj + (unsigned)(nyq * sizeof(float2)) : OPH_BUF_OOB;
                    vy[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ri, oy, 0, 0));
                }
#pragma unroll
                for (int j = 0; j < CHB; ++j) {
                    const int kx = tl + (m0 + j) * NT;
                    if (m0 + j < EPT && (EPT * NT == LX || kx < LX))
                        line[fft::pad16(kx)] = pair0 ? make_float2(v[j].x, vy[j]) : v[j];
                }
            }
        } else {
            const float2* src = a.I + (size_t)(l - 1) * lstride + (size_t)kya * GROUP;
            const unsigned char* fl = a.flags + (size_t)(l - 1) * a.ngroups;
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const int kx = tl + m * NT;
                if (EPT * NT == LX || kx < LX) {
                    const int x = kx - a.ox;
                    float2 v = make_float2(0.f, 0.f);
                    if (x >= 0 && x < a.w && fl[x / GROUP]) {
                        const size_t o = (size_t)(x / GROUP) * gstride + x % GROUP;
                        v = src[o];
                        if (pair0) v.y = src[nyq + o].x;
                    }
                    line[fft::pad16(kx)] = v;
                }
            }
        }
        __syncthreads();
        fftm::line<LX, NT, -1>(line, twx, tl);
        const double cx = a.lc[3 * (l - 1)], cy = a.lc[3 * (l - 1) + 1], K = a.lc[3 * (l - 1) + 2];
        const float Kf = (float)K;
        double tya = cy * (double)sya * sya;
        tya -= rint(tya);
        double tyb = cy * (double)syb * syb;
        tyb -= rint(tyb);
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int kx = tl + m * NT;
            if (EPT * NT == LX || kx < LX) {
                const int sx = sgnidx(kx, LX);
                const double sx2 = (double)sx * sx;
                const float usx = a.ux * (float)(sx * sx);
                double txd = cx * sx2;
                txd -= rint(txd);
                const float tx = (float)txd;
                float2 gv = line[fft::pad16(kx)];
                const float2 gm = line[fft::pad16(kx == 0 ? 0 : LX - kx)];
                float2 gmc = make_float2(gm.x, -gm.y);
                if (pair0) split_real_pair(gv, gmc);
                const float2 ha = dm_transfer<EXACT>(a, tx, tya, usx, sya, K, Kf, sx2);
                acca[m] = fft::cadd(acca[m], fft::cmul(ha, gv));
                const float2 hb = dm_transfer<EXACT>(a, tx, tyb, usx, syb, K, Kf, sx2);
                accb[m] = fft::cadd(accb[m], fft::cmul(hb, gmc));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        acca[m] = fft::cmul(acca[m], a.rho);
        accb[m] = fft::cmul(accb[m], a.rho);
    }
    finish_row_m<LX>(a, acca, kya, line, twx, t);
    finish_row_m<LX>(a, accb, kyb, line, twx, t);
}

template <int LY, int B>
__global__ __launch_bounds__(B * fftm::threads(LY)) void dm_final_m(DmArgs a, const float2* __restrict__ twy) {
    constexpr int NT = fftm::threads(LY);
    constexpr int PL = fftm::padded_len(LY);
    extern __shared__ float2 lds[];
    const int nthr = B * NT;
    const int g = xcd_remap(blockIdx.x, gridDim.x);
    const int c0 = g * B;
    const int b = threadIdx.x / NT, t = threadIdx.x % NT;
    for (int e = threadIdx.x; e < LY * B; e += nthr) {
        const int bb = e % B, ky = e / B;
        const int x = c0 + bb;
        const int yc = (ky + a.Ny / 2) % a.Ny;
        const bool kept = !a.encode || (yc >= a.cy1 && yc <= a.cy2);
        lds[bb * PL + fft::pad16(ky)] = (x < a.w && kept) ? a.T[(size_t)ky * a.w + x] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fftm::line<LY, NT, +1>(lds + b * PL, twy, t);
    const float sc = (float)a.scale;
    for (int e = threadIdx.x; e < a.h * B; e += nthr) {
        const int bb = e % B, r = e / B;
        const int x = c0 + bb;
        if (x >= a.w) continue;
        const float2 v0 = lds[bb * PL + fft::pad16(a.oy + r)];
        const float2 v = make_float2(v0.x * sc, v0.y * sc);
        const size_t o = (size_t)r * a.w + x;
        if (a.encode) {
            double tt = a.h / 8.0 - (r + 1) / 4.0;
            tt -= rint(tt);
            const float ph = (float)tt;
            const float re = v.x * cos_turns(ph) - v.y * sin_turns(ph);
            if (a.fringe) a.fringe[o] = re;
            if (a.fringe32) a.fringe32[o] = re;
        } else {
            a.field[o] = v;
        }
    }
}

// ---- launchers -------------------------------------------------------------------

template <int LY>
int launch_dm_cols(oph_ctx* ctx, const DmArgs& a, int chunks, int lpw) {
    const float2* tw = oph_twiddles(ctx, LY);
    if (!tw) return OPH_ENOMEM;
    const size_t shm = 2 * fft::padded_len(LY) * sizeof(float2) + (size_t)a.h * GROUP * 4 +
                       (size_t)((a.L + 32) / 32) * 4;
    if (shm > 160 * 1024) return oph_fail(ctx, OPH_ENOTSUP, "scene height %d too large for the column pass", a.h);
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&dm_cols<LY>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        hipGetLastError();
        attr = true;
    }
    OPH_LAUNCH(ctx, "dm_cols", dm_cols<LY>, dim3(a.ngroups, chunks), dim3(2 * (LY / 16)), shm, a, lpw, tw);
    return OPH_OK;
}

template <int LX>
int launch_dm_rows(oph_ctx* ctx, const DmArgs& a) {
    const float2* tw = oph_twiddles(ctx, LX);
    if (!tw) return OPH_ENOMEM;
    if constexpr (LX >= 64 && LX <= 4096) {
        if (a.horner) {
            const float2* tw8 = oph_twiddles(ctx, LX, 8);
            if (!tw8) return OPH_ENOMEM;
            static bool attr = false;
            if (!attr) {
                for (const void* k : {reinterpret_cast<const void*>(&dm_rows_h<LX, true>),
                                      reinterpret_cast<const void*>(&dm_rows_h<LX, false>)})
                    hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
                hipGetLastError();
                attr = true;
            }
            DmArgs b = a;
            const size_t fbytes = (size_t)a.L * a.fwpl * sizeof(unsigned);
            b.fbits_lds = rows_h_lds<LX>() + fbytes <= 160 * 1024 ? 1 : 0;
            const size_t shm = rows_h_lds<LX>() + (b.fbits_lds ? fbytes : 0);
            if (b.fbits_lds)
                OPH_LAUNCH(ctx, "dm_rows", (dm_rows_h<LX, true>), dim3(a.Ny / 2), dim3(LX / 8), shm, b, tw8);
            else
                OPH_LAUNCH(ctx, "dm_rows", (dm_rows_h<LX, false>), dim3(a.Ny / 2), dim3(LX / 8), shm, b, tw8);
            return OPH_OK;
        }
    }
    if (a.exact)
        OPH_LAUNCH(ctx, "dm_rows", (dm_rows<LX, true>), dim3(a.Ny / 2), dim3(LX / 16), 0, a, tw);
    else
        OPH_LAUNCH(ctx, "dm_rows", (dm_rows<LX, false>), dim3(a.Ny / 2), dim3(LX / 16), 0, a, tw);
    return OPH_OK;
}

template <int LY>
int launch_dm_final(oph_ctx* ctx, const DmArgs& a) {
    const float2* tw = oph_twiddles(ctx, LY);
    if (!tw) return OPH_ENOMEM;
    constexpr int B = fftk::col_lines(LY) >= 4 ? 4 : fftk::col_lines(LY);
    const size_t shm = (size_t)B * fft::padded_len(LY) * sizeof(float2);
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&dm_final<LY, B>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        attr = true;
    }
    OPH_LAUNCH(ctx, "dm_final", (dm_final<LY, B>), dim3((a.w + B - 1) / B), dim3(B * (LY / 16)), shm, a, tw);
    return OPH_OK;
}

template <int LY>
int launch_dm_cols_m(oph_ctx* ctx, const DmArgs& a, int chunks, int lpw) {
    const float2* tw = oph_twiddles_mixed(ctx, LY);
    if (!tw) return OPH_ENOMEM;
    if constexpr (2 * (size_t)fftm::padded_len(LY) * sizeof(float2) > 150 * 1024) {
        (void)chunks;
        (void)lpw;
        return oph_fail(ctx, OPH_ENOTSUP, "grid height %d too large for the column pass", LY);
    } else {
        const size_t shm = 2 * (size_t)fftm::padded_len(LY) * sizeof(float2) + (size_t)a.h * GROUP * 4 +
                           (size_t)((a.L + 32) / 32) * 4;
        if (shm > 160 * 1024)
            return oph_fail(ctx, OPH_ENOTSUP, "scene height %d too large for the column pass", a.h);
        static bool attr = false;
        if (!attr) {
            hipFuncSetAttribute(reinterpret_cast<const void*>(&dm_cols_m<LY>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            hipGetLastError();
            attr = true;
        }
        OPH_LAUNCH(ctx, "dm_cols", dm_cols_m<LY>, dim3(a.ngroups, chunks), dim3(2 * fftm::threads(LY)), shm, a,
                   lpw, tw);
        return OPH_OK;
    }
}

template <int LX>
int launch_dm_rows_m(oph_ctx* ctx, const DmArgs& a) {
    const float2* tw = oph_twiddles_mixed(ctx, LX);
    if (!tw) return OPH_ENOMEM;
    const size_t shm = (size_t)fftm::padded_len(LX) * sizeof(float2);
    static bool attr = false;
    if (!attr) {
        for (const void* k : {reinterpret_cast<const void*>(&dm_rows_m<LX, true>),
                              reinterpret_cast<const void*>(&dm_rows_m<LX, false>)})
            hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        attr = true;
    }
    if (a.exact)
        OPH_LAUNCH(ctx, "dm_rows", (dm_rows_m<LX, true>), dim3(a.Ny / 2), dim3(fftm::threads(LX)), shm, a, tw);
    else
        OPH_LAUNCH(ctx, "dm_rows", (dm_rows_m<LX, false>), dim3(a.Ny / 2), dim3(fftm::threads(LX)), shm, a, tw);
    return OPH_OK;
}

// columns per dm_final_m workgroup: up to 4, at most 1024 threads and ~140 KB
__host__ __device__ constexpr int final_m_cols(int LY) {
    int b = 4;
    while (b > 1 && (b * fftm::threads(LY) > 1024 || (size_t)b * fftm::padded_len(LY) * 8 > 140 * 1024)) b /= 2;
    return b;
}

template <int LY>
int launch_dm_final_m(oph_ctx* ctx, const DmArgs& a) {
    const float2* tw = oph_twiddles_mixed(ctx, LY);
    if (!tw) return OPH_ENOMEM;
    constexpr int B = final_m_cols(LY);
    const size_t shm = (size_t)B * fftm::padded_len(LY) * sizeof(float2);
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&dm_final_m<LY, B>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        attr = true;
    }
    OPH_LAUNCH(ctx, "dm_final", (dm_final_m<LY, B>), dim3((a.w + B - 1) / B), dim3(B * fftm::threads(LY)), shm,
               a, tw);
    return OPH_OK;
}

// Depth index of each u8 depth value d, as the reference forms it per pixel:
// mode 0 (FLAG_CHANGE_DEPTH_QUANTIZATION = 0): Q - d (0, no layer, below 0);
// mode 1, change_depth_quan (CPU.cpp:74-103): dmap = (1 - d/255)(far - near)
// + near, and every level l adds (l + 1) where dmap falls in
// [lev_l - dstep/2, lev_l + dstep/2) ([.., ..] for the last), capped at
// 65535.  Each product and sum is its own statement (no fused multiply-add),
// as the reference's double arithmetic.
void depth_index_table(int mode, int Q, const std::vector<double>& lev, double dstep, double near_m,
                       double far_m, double* table) {
    const int L = (int)lev.size();
    for (int d = 0; d < 256; ++d) {
        if (mode == 0) {
            const int v = Q - d;
            table[d] = (double)(unsigned short)(v < 0 ? 0 : v);
            continue;
        }
        const double dsrc = (double)d / 255.0;
        const double span = far_m - near_m;
        const double scaled = (1 - dsrc) * span;
        const double dmap = scaled + near_m;
        long long acc = 0;
        for (int l = 0; l < L; ++l) {
            const double half = dstep / 2.0;
            const double d1 = lev[l] - half, d2 = lev[l] + half;
            const bool t = (l < L - 1) ? (dmap >= d1 && dmap < d2) : (dmap >= d1 && dmap <= d2);
            if (t) acc += l + 1;
        }
        table[d] = (double)(acc > 65535 ? 65535 : acc);
    }
}

}  // namespace

extern "C" int oph_depthmap(oph_ctx* ctx, const oph_dm_spec* s, int ncolor, const double* lambdas,
                            const unsigned char* intensity, const unsigned char* depth, void* out, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!s || !lambdas || !intensity || !depth || !out || ncolor < 1)
        return oph_fail(ctx, OPH_EINVAL, "null argument");
    const int w = s->w, h = s->h, Nx = s->Nx, Ny = s->Ny, L = s->num_depth;
    if (w <= 0 || h <= 0 || w > Nx || h > Ny) return oph_fail(ctx, OPH_EINVAL, "scene %dx%d must fit the grid %dx%d", w, h, Nx, Ny);
    // both sides powers of two, or both among the general lengths
    const bool mixed = !(fftk::supported(Nx) && fftk::supported(Ny));
    if (mixed && !(dm_mixed_length(Nx) && dm_mixed_length(Ny)))
        return oph_fail(ctx, OPH_ENOTSUP,
                        "grid %dx%d: sides must both be powers of two in 16..8192 or both among 540, 720, 1080, "
                        "1920, 2160, 3840, 11520", Nx, Ny);
    if (L < 1 || L > MAX_LAYERS) return oph_fail(ctx, OPH_EINVAL, "num_depth must be in 1..%d", MAX_LAYERS);
    if (!(s->pitch_x > 0) || !(s->pitch_y > 0) || !(s->far_m >= s->near_m))
        return oph_fail(ctx, OPH_EINVAL, "bad pitch or depth range");
    if (s->quant_mode != 0 && s->quant_mode != 1) return oph_fail(ctx, OPH_EINVAL, "quant_mode must be 0 or 1");
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    for (int c = 0; c < ncolor; ++c)
        if (!(lambdas[c] > 0)) return oph_fail(ctx, OPH_EINVAL, "wavelength %d must be positive", c);
    const int fmt = flags & OPH_DM_OUT_MASK;
    const int lo = s->layer_lo >= 1 ? s->layer_lo : 1;
    const int hi = (s->layer_hi >= 1 && s->layer_hi <= L) ? s->layer_hi : L;

    // depth levels (GetDepthValues, HologramGenerator.cpp:393-401: repeated
    // addition; continued to exactly L entries) and the viewing-window
    // transform z_l = -f d/(d - f) (:432), or z_l = d without a field lens
    std::vector<double> lev(L);
    if (L > 1) {
        const double dstep = (s->far_m - s->near_m) / (L - 1);
        double v = s->near_m;
        for (int i = 0; i < L; ++i) { lev[i] = v; v += dstep; }
    } else {
        lev[0] = s->far_m - s->near_m;
    }
    const double dstep = L > 1 ? (s->far_m - s->near_m) / (L - 1) : (s->far_m + s->near_m) / 2;
    const double Sx = s->pitch_x * Nx, Sy = s->pitch_y * Ny;
    std::vector<double>& hc = ctx->host_scratch;
    // per colour and layer 3 constants, the L levels, the 256-entry depth-index table
    hc.assign((size_t)3 * L * ncolor + L + 256, 0.0);
    double zmax = 0.0;
    for (int c = 0; c < ncolor; ++c) {
        const double lam = lambdas[c];
        for (int l = 0; l < L; ++l) {
            const double d = lev[l];
            const double z = s->field_lens > 0.0 ? -s->field_lens * d / (d - s->field_lens) : d;
            double* o = &hc[(size_t)3 * (c * L + l)];
            o[0] = z * lam / (2.0 * Sx * Sx);
            o[1] = z * lam / (2.0 * Sy * Sy);
            o[2] = z / lam;
            zmax = fmax(zmax, fabs(z / lam));
        }
    }
    for (int l = 0; l < L; ++l) hc[(size_t)3 * L * ncolor + l] = lev[l];
    depth_index_table(s->quant_mode, s->default_q, lev, dstep, s->near_m, s->far_m, &hc[(size_t)3 * L * ncolor + L]);

    const size_t npx = (size_t)w * h;
    const int ngroups = (w + GROUP - 1) / GROUP;
    const size_t lstride = (size_t)(Ny / 2 + 1) * GROUP * ngroups;
    // device buffers
    unsigned char* dimg = const_cast<unsigned char*>(intensity);
    unsigned char* ddep = const_cast<unsigned char*>(depth);
    if (!(flags & OPH_DEVICE_PTR)) {
        dimg = static_cast<unsigned char*>(oph_scratch(ctx, SLOT_IN, npx * (ncolor + 1)));
        if (!dimg) return OPH_ENOMEM;
        ddep = dimg + npx * ncolor;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(dimg, intensity, npx * ncolor, hipMemcpyHostToDevice, ctx->stream));
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(ddep, depth, npx, hipMemcpyHostToDevice, ctx->stream));
    }
    const int fwpl = (ngroups + 31) / 32;
    const size_t aux_bytes = npx * 2 + (size_t)L * ngroups + L + 4 + (size_t)L * fwpl * 4 + 64 +
                             hc.size() * sizeof(double);
    unsigned char* aux = static_cast<unsigned char*>(oph_scratch(ctx, SLOT_AUX0, aux_bytes + 256));
    float2* I = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK2, (size_t)L * lstride * sizeof(float2)));
    float2* T = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK1, (size_t)Ny * w * sizeof(float2)));
    if (!aux || !I || !T) return OPH_ENOMEM;
    double* dlc = reinterpret_cast<double*>(aux);
    unsigned short* didx = reinterpret_cast<unsigned short*>(aux + hc.size() * sizeof(double));
    unsigned char* dflags = reinterpret_cast<unsigned char*>(didx + npx);
    unsigned char* dany = dflags + (size_t)L * ngroups;
    // the flag bits follow, at the next 4-byte boundary (word atomics)
    unsigned* dfbits = reinterpret_cast<unsigned*>((reinterpret_cast<uintptr_t>(dany + L) + 3) & ~(uintptr_t)3);
    const size_t nfl_bytes = reinterpret_cast<unsigned char*>(dfbits + (size_t)L * fwpl) - dflags;
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(dlc, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));

    // outputs
    const size_t osz = fmt == OPH_DM_OUT_FIELD ? sizeof(float2) : fmt == OPH_DM_OUT_FRINGE32 ? 4 : 8;
    void* dout = out;
    if (!(flags & OPH_DEVICE_PTR)) {
        dout = oph_scratch(ctx, SLOT_OUT, npx * ncolor * osz);
        if (!dout) return OPH_ENOMEM;
    }

    OPH_LAUNCH(ctx, "dm_index", dm_index, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, ddep, (int)npx,
               dlc + (size_t)3 * L * ncolor + L, didx);

    // truncated-series accuracy at the grid corner (see dm_transfer): the
    // first dropped term is 21/1024 u^6 z/lambda turns
    double lmax = 0;
    for (int c = 0; c < ncolor; ++c) lmax = fmax(lmax, lambdas[c]);
    const double fxm = 0.5 / s->pitch_x, fym = (Ny / 2 + 1.0) / Sy;
    const double umax = lmax * lmax * (fxm * fxm + fym * fym);
    const bool exact = umax >= 0.5 || zmax * pow(umax, 6) * (21.0 / 1024.0) > 1e-7;

    // equally spaced z (no field lens): the chunked recurrence of dm_rows_h
    // applies when every step equals the first to far below a phase ulp
    std::vector<double> zs(L);
    for (int l = 0; l < L; ++l)
        zs[l] = s->field_lens > 0.0 ? -s->field_lens * lev[l] / (lev[l] - s->field_lens) : lev[l];
    const double dz = L > 1 ? zs[1] - zs[0] : 0.0;
    double dev = 0.0;
    for (int l = 1; l < L; ++l) dev = fmax(dev, fabs((zs[l] - zs[l - 1]) - dz));
    const double phimax = lmax * (fxm * fxm + fym * fym) / 2.0 + 0.25 * umax * umax / lmax;  // turns per metre
    const bool horner = !mixed && !exact && L > 1 && hi > lo && dev * phimax < 1e-9;

    for (int c = 0; c < ncolor; ++c) {
        const double lam = lambdas[c];
        DmArgs a{};
        a.horner = horner ? 1 : 0;
        a.dcx = dz * lam / (2.0 * Sx * Sx);
        a.dcy = dz * lam / (2.0 * Sy * Sy);
        a.dK = dz / lam;
        a.w = w; a.h = h; a.Nx = Nx; a.Ny = Ny;
        a.ox = (Nx - w) / 2;
        a.oy = (Ny - h) / 2;
        a.L = L; a.lo = lo; a.hi = hi;
        a.ngroups = ngroups;
        a.img = dimg + npx * c;
        a.idx = didx;
        a.I = I;
        a.flags = dflags;
        a.any = dany;
        a.fbits = dfbits;
        a.fwpl = fwpl;
        a.lc = dlc + (size_t)3 * L * c;
        a.ux = (float)(lam * lam / (Sx * Sx));
        a.uy = (float)(lam * lam / (Sy * Sy));
        a.lam2_sx2 = lam * lam / (Sx * Sx);
        a.lam2_sy2 = lam * lam / (Sy * Sy);
        a.inv_lambda = 1.0 / lam;
        a.exact = exact ? 1 : 0;
        a.rho = make_float2((float)s->rho_re, (float)s->rho_im);
        a.encode = fmt != OPH_DM_OUT_FIELD;
        const int cropy = Ny / 2;
        a.cy1 = cropy - cropy / 2 - 1;
        a.cy2 = a.cy1 + cropy - 1;
        a.T = T;
        a.spectrum = nullptr;
        a.scale = 1.0 / ((double)Nx * Ny);
        char* ob = static_cast<char*>(dout) + npx * c * osz;
        a.fringe = fmt == OPH_DM_OUT_FRINGE ? reinterpret_cast<double*>(ob) : nullptr;
        a.fringe32 = fmt == OPH_DM_OUT_FRINGE32 ? reinterpret_cast<float*>(ob) : nullptr;
        a.field = fmt == OPH_DM_OUT_FIELD ? reinterpret_cast<float2*>(ob) : nullptr;

        const int nfl = (int)nfl_bytes;   // flags, any and the flag bits
        OPH_LAUNCH(ctx, "dm_zero", dm_zero, dim3((nfl + 255) / 256), dim3(256), 0, dflags, nfl);
        const int nl = hi - lo + 1;
        int chunks = (int)((4LL * ctx->cu_count + ngroups - 1) / ngroups);
        if (chunks > nl) chunks = nl;
        if (chunks < 1) chunks = 1;
        const int lpw = (nl + chunks - 1) / chunks;
        chunks = (nl + lpw - 1) / lpw;
        if (mixed) {
            switch (Ny) {
#define OPH_DMC(n) case n: rc = launch_dm_cols_m<n>(ctx, a, chunks, lpw); break;
                OPH_DM_MIXED_LENGTHS(OPH_DMC)
#undef OPH_DMC
            }
            if (rc) return rc;
            switch (Nx) {
#define OPH_DMR(n) case n: rc = launch_dm_rows_m<n>(ctx, a); break;
                OPH_DM_MIXED_LENGTHS(OPH_DMR)
#undef OPH_DMR
            }
            if (rc) return rc;
            switch (Ny) {
#define OPH_DMF(n) case n: rc = launch_dm_final_m<n>(ctx, a); break;
                OPH_DM_MIXED_LENGTHS(OPH_DMF)
#undef OPH_DMF
            }
            if (rc) return rc;
            continue;
        }
        switch (Ny) {
#define OPH_DMC(n) case n: rc = launch_dm_cols<n>(ctx, a, chunks, lpw); break;
            OPH_FFT_LENGTHS(OPH_DMC)
#undef OPH_DMC
        }
        if (rc) return rc;
        switch (Nx) {
#define OPH_DMR(n) case n: rc = launch_dm_rows<n>(ctx, a); break;
            OPH_FFT_LENGTHS(OPH_DMR)
#undef OPH_DMR
        }
        if (rc) return rc;
        switch (Ny) {
#define OPH_DMF(n) case n: rc = launch_dm_final<n>(ctx, a); break;
            OPH_FFT_LENGTHS(OPH_DMF)
#undef OPH_DMF
        }
        if (rc) return rc;
    }
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(out, dout, npx * ncolor * osz, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}
