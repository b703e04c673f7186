// fft.hip -- 2-D complex64 FFT for gfx950 (replaces FFTW / cuFFT Z2Z of the
// reference's fftwShift / cudaFFT, DM/Hologram/src/HologramGenerator_CPU.cpp:
// 285-322, DM/Hologram/src/HologramKernel.cu:189-225).
//
// Row pass: one LDS line per L/16 threads, several lines per workgroup for
// short rows; rows are read and written with unit stride.  For L = 4096 a
// persistent kernel loads the next row into registers while it transforms
// the current one (first and last stages on registers).
// Column pass: a workgroup stages B neighbouring columns (B*8-byte runs per
// row) in LDS and transforms them together; workgroup ids are remapped so
// neighbouring column groups run on the same XCD and share its L2.
// Twiddle tables are computed once per length in FP64 and cached in the
// context.  The plan is implicit (length -> kernel instance); nothing is
// created or destroyed per call.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "fft_engine.h"
#include "fft_mixed.h"
#include "fft_kernels.h"

const float2* oph_twiddles(oph_ctx* ctx, int L, int radix) {
    int lg = 0;
    while ((1 << lg) < L) ++lg;
    if ((1 << lg) != L || lg >= 24 || (radix != 16 && radix != 8)) {
        oph_fail(ctx, OPH_EINVAL, "FFT length %d / radix %d unsupported", L, radix);
        return nullptr;
    }
    void** cache = radix == 16 ? ctx->tw : ctx->tw8;
    if (cache[lg]) return static_cast<const float2*>(cache[lg]);
    // per-stage tables in the order the line transforms walk the stages
    std::vector<float2> h;
    int Ns = 1, rem = L;
    while (rem > 1) {
        const int R = (rem % radix == 0) ? radix : rem;
        if (Ns > 1)
            for (int r = 1; r < R; ++r)
                for (int k = 0; k < Ns; ++k) {
                    const double a = -2.0 * M_PI * (double)k * r / ((double)Ns * R);
                    h.push_back(make_float2((float)cos(a), (float)sin(a)));
                }
        Ns *= R;
        rem /= R;
    }
    if (h.empty()) h.push_back(make_float2(1.f, 0.f));
    void* d = nullptr;
    if (hipMalloc(&d, sizeof(float2) * h.size()) != hipSuccess) {
        hipGetLastError();
        oph_fail(ctx, OPH_ENOMEM, "twiddle table allocation failed");
        return nullptr;
    }
    if (hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipGetLastError();
        hipFree(d);
        oph_fail(ctx, OPH_EDEVICE, "twiddle table upload failed");
        return nullptr;
    }
    cache[lg] = d;
    return static_cast<const float2*>(d);
}

const float2* oph_twiddles_mixed(oph_ctx* ctx, int N) {
    const fftm::Plan P = fftm::plan(N);
    if (N < 2 || P.n <= 0) {
        oph_fail(ctx, OPH_EINVAL, "FFT length %d is not 2^a 3^b 5^c", N);
        return nullptr;
    }
    int slot = -1;
    for (int i = 0; i < 16; ++i) {
        if (ctx->twm_len[i] == N) return static_cast<const float2*>(ctx->twm[i]);
        if (slot < 0 && !ctx->twm[i]) slot = i;
    }
    if (slot < 0) {
        oph_fail(ctx, OPH_ENOTSUP, "too many mixed-radix FFT lengths in one context");
        return nullptr;
    }
    // per-stage tables in the order the plan walks the stages
    std::vector<float2> h;
    int Ns = 1;
    for (int s = 0; s < P.n; ++s) {
        const int R = P.r[s];
        if (Ns > 1)
            for (int r = 1; r < R; ++r)
                for (int k = 0; k < Ns; ++k) {
                    const double a = -2.0 * M_PI * (double)k * r / ((double)Ns * R);
                    h.push_back(make_float2((float)cos(a), (float)sin(a)));
                }
        Ns *= R;
    }
    if (h.empty()) h.push_back(make_float2(1.f, 0.f));
    void* d = nullptr;
    if (hipMalloc(&d, sizeof(float2) * h.size()) != hipSuccess) {
        hipGetLastError();
        oph_fail(ctx, OPH_ENOMEM, "twiddle table allocation failed");
        return nullptr;
    }
    if (hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipGetLastError();
        hipFree(d);
        oph_fail(ctx, OPH_EDEVICE, "twiddle table upload failed");
        return nullptr;
    }
    ctx->twm[slot] = d;
    ctx->twm_len[slot] = N;
    return static_cast<const float2*>(d);
}

int fftk::col_width(int L, int bmax) {
    static int env = -1;
    if (env < 0) {
        const char* e = getenv("OPH_FFT_COLUMNS");
        env = e ? atoi(e) : 0;
    }
    int b = env > 0 ? env : (L >= 4096 ? 2 : (L >= 2048 ? 4 : 8));
    if (b > bmax) b = bmax;
    if (b < 1) b = 1;
    return b;
}

namespace {

// ---- plain row / column kernels --------------------------------------------

template <int L, int SIGN>
__global__ __launch_bounds__(512) void fft_rows_kernel(float2* data, int rows, int stride, float scale,
                                                       const float2* __restrict__ tw) {
    constexpr int TPL = L / 16;
    constexpr int LPW = fftk::row_lines(L);
    constexpr int PL = fft::padded_len(L);
    __shared__ float2 lds[LPW * PL];
    const int sub = threadIdx.x / TPL, t = threadIdx.x % TPL;
    const int row = blockIdx.x * LPW + sub;
    float2* buf = lds + sub * PL;
    const bool ok = row < rows;
    float2* src = data + (size_t)(ok ? row : 0) * stride;
    fft::Twiddles<L> T;
    T.load(tw, t);
#pragma unroll
    for (int m = 0; m < 16; ++m) buf[fft::lane_pad<L>(t, m)] = ok ? src[t + m * TPL] : make_float2(0.f, 0.f);
    __syncthreads();
    fft::line<L, SIGN>(buf, T, t);
    if (ok) {
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const float2 v = buf[fft::lane_pad<L>(t, m)];
            src[t + m * TPL] = make_float2(v.x * scale, v.y * scale);
        }
    }
}

// Persistent row transform for L = 16^3: thread t holds elements t + m L/16
// in the first stage's inputs and in the last stage's outputs, so a row goes
// from global memory straight into the first butterfly and from the last
// butterfly straight back.  The next row is loaded (into the registers the
// current row was read from) while the current one is transformed; the
// middle stage's twiddles sit in LDS and the last stage's 15 per thread are
// rebuilt from two powers, so no global load but the row's is issued in the
// loop, and rows are range-checked buffer accesses without branches.  With
// the first row peeled off the loop, the memory wait before each row's
// first butterfly covers that row's loads only (the wait counter is in
// order).  In place: each row is loaded and stored by one workgroup.
template <int L, int SIGN>
__global__ __launch_bounds__(L / 16) void fft_rows_f(float2* data, int rows, int stride, float scale,
                                                     const float2* __restrict__ tw) {
    static_assert(fft::regs_fusable<L>() && L / 16 >= 64, "L = 16^a, a >= 3");
    constexpr int TPL = L / 16;
    constexpr int PL = fft::padded_len(L);
    constexpr int TWO = fft::last_stage_tw_offset<L>();
    __shared__ float2 buf[PL];
    __shared__ float2 twm[TWO];
    const int t = threadIdx.x;
    for (int i = t; i < TWO; i += TPL) twm[i] = tw[i];
    const float2 w1 = tw[TWO + t], w4 = tw[TWO + 3 * TPL + t];
    // the whole matrix, below 2^31 bytes (checked by the launcher)
    const __amdgpu_buffer_rsrc_t R = oph_buf_rsrc(data, (unsigned)((size_t)rows * stride * sizeof(float2)));
    float2 in[16];
    auto fetch = [&](int row) {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            in[m] = oph_buf_load_f2(
                R, row < rows ? (unsigned)(((size_t)row * stride + t + m * TPL) * sizeof(float2)) : OPH_BUF_OOB);
    };
    auto row = [&](int r) {
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = in[m];
        __syncthreads();  // the previous row's last-stage reads are done
        fft::stage0_from_regs<L, SIGN>(v, buf, t);
        fetch(r + gridDim.x);
        __syncthreads();
        fft::middle_stages_tw<L, SIGN>(buf, twm, t);
        float2 wl[16];
        fft::twiddle_powers(w1, w4, wl);
        fft::last_stage_to_regs_w<L, SIGN>(buf, wl, t, v);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            oph_buf_store_f2(make_float2(v[m].x * scale, v[m].y * scale), R,
                             (unsigned)(((size_t)r * stride + t + m * TPL) * sizeof(float2)));
    };
    int r = blockIdx.x;
    fetch(r);
    if (r < rows) {
        row(r);
        for (r += gridDim.x; r < rows; r += gridDim.x) row(r);
    }
}

// wide: every column of the workgroup exists, the stride is even and the
// L x stride array is below 2 GiB; then each lane moves two neighbouring
// columns of a row (16 bytes) with range-checked buffer accesses, and its 8
// loads are all issued before the first is used (a load loop with a bounds
// branch completes each load before issuing the next).
template <int L, int SIGN, int B>
__global__ __launch_bounds__(1024) void fft_cols_kernel(float2* data, int ncols, int stride, float scale,
                                                        const float2* __restrict__ tw, int wide) {
    constexpr int TPL = L / 16;
    constexpr int PL = fft::padded_len(L);
    extern __shared__ float2 lds[];
    const int nthr = TPL * B;
    const int g = xcd_remap(blockIdx.x, gridDim.x);
    const int c0 = g * B;
    constexpr int PER = (L * B / 2) / (TPL * B);  // 16-byte pieces per lane
    const bool w2 = B % 2 == 0 && wide && c0 + B <= ncols;
    const auto rs = oph_buf_rsrc(data, (unsigned)((size_t)L * stride * sizeof(float2)));
    if constexpr (B % 2 == 0 && L <= 4096) {
        if (w2) {
            float4 ld[PER];
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int idx = threadIdx.x + j * nthr;
                const int pr = idx % (B / 2), i = idx / (B / 2);
                ld[j] = oph_buf_load_f4(rs, (unsigned)(((size_t)i * stride + c0 + 2 * pr) * sizeof(float2)));
            }
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int idx = threadIdx.x + j * nthr;
                const int pr = idx % (B / 2), i = idx / (B / 2);
                lds[(2 * pr) * PL + fft::pad16(i)] = make_float2(ld[j].x, ld[j].y);
                lds[(2 * pr + 1) * PL + fft::pad16(i)] = make_float2(ld[j].z, ld[j].w);
            }
        }
    }
    if (!w2 || !(B % 2 == 0 && L <= 4096)) {
        for (int idx = threadIdx.x; idx < L * B; idx += nthr) {
            const int b = idx % B, i = idx / B;
            const int c = c0 + b;
            lds[b * PL + fft::pad16(i)] = c < ncols ? data[(size_t)i * stride + c] : make_float2(0.f, 0.f);
        }
    }
    // (twiddles after the column data: their loads do not hold up the data's)
    fft::Twiddles<L> T;
    T.load(tw, threadIdx.x % TPL);
    __syncthreads();
    const int b = threadIdx.x / TPL, t = threadIdx.x % TPL;
    fft::line<L, SIGN>(lds + b * PL, T, t);
    if (w2) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int idx = threadIdx.x + j * nthr;
            const int pr = idx % (B / 2), i = idx / (B / 2);
            const float2 u = lds[(2 * pr) * PL + fft::pad16(i)], v = lds[(2 * pr + 1) * PL + fft::pad16(i)];
            oph_buf_store_f4(make_float4(u.x * scale, u.y * scale, v.x * scale, v.y * scale), rs,
                             (unsigned)(((size_t)i * stride + c0 + 2 * pr) * sizeof(float2)));
        }
    } else {
        for (int idx = threadIdx.x; idx < L * B; idx += nthr) {
            const int bb = idx % B, i = idx / B;
            const int c = c0 + bb;
            if (c < ncols) {
                const float2 v = lds[bb * PL + fft::pad16(i)];
                data[(size_t)i * stride + c] = make_float2(v.x * scale, v.y * scale);
            }
        }
    }
}

// 4096-point columns, persistent: workgroups of four columns (1024 threads,
// 140 KB of LDS: one per CU) walk column groups.  Thread T transforms column
// T mod 4 as butterfly t = T/4 of the row transform's register-fused form:
// its stage-0 inputs and last-stage outputs are rows t + 256 m (m < 16), so
// the columns go from global memory straight into the first butterfly and
// from the last butterfly straight back -- four LDS passes per column
// instead of eight -- and a wave's 8-byte accesses cover 16 rows x 32
// bytes (the four columns).  The next group is loaded into the registers
// the current one was read from while it is transformed.  Groups are split
// by XCD (the hardware places workgroup b on XCD b mod 8): XCD x owns groups
// [x G/8, (x + 1) G/8), its workgroups take consecutive groups at each step.
// The middle stage's twiddles sit in LDS and the last stage's 15 per thread
// are rebuilt from two powers (also in LDS), so no global load but the
// columns' is issued in the loop; with the first group peeled off the loop,
// the memory wait before a group's first butterfly covers that group's loads
// only (the wait counter is in order).  ncols is a multiple of 32, the grid a
// multiple of 8.
constexpr int COLP_L = 4096, COLP_B = 4, COLP_THREADS = COLP_B * (COLP_L / 16);
// LDS entries ahead of the lines: the middle stage's 240 twiddles, then W^t
// and W^(4t) (W = exp(-2 pi i/4096)) for t < 256
constexpr int COLP_TW = 256 + 2 * 256;

template <int SIGN>
__global__ __launch_bounds__(COLP_THREADS) void fft_cols_p(float2* data, int ncols, int stride, float scale,
                                                          const float2* __restrict__ tw) {
    constexpr int L = COLP_L, B = COLP_B, TPL = L / 16, PL = fft::padded_len(L);
    constexpr int TWO = fft::last_stage_tw_offset<L>();
    static_assert(TWO <= COLP_TW, "middle-stage table");
    extern __shared__ float2 lds[];
    float2* twm = lds;
    const int b = threadIdx.x % B, t = threadIdx.x / B;
    float2* buf = lds + COLP_TW + b * PL;
    float2* wp = lds + 256;   // [t]: W^t, [TPL + t]: W^(4t)
    for (int i = threadIdx.x; i < TWO; i += COLP_THREADS) twm[i] = tw[i];
    for (int i = threadIdx.x; i < 2 * TPL; i += COLP_THREADS) wp[i] = tw[TWO + (i < TPL ? i : 2 * TPL + i)];
    const int per = (ncols / B) / 8, steps = gridDim.x / 8;
    const int xcd = blockIdx.x % 8;
    const int gend = (xcd + 1) * per;
    const auto R = oph_buf_rsrc(data, (unsigned)((size_t)L * stride * sizeof(float2)));
    // element (t + m TPL, B grp + b): a per-lane offset plus m TPL rows as
    // the uniform (scalar) offset
    const unsigned rowstep = (unsigned)((size_t)TPL * stride * sizeof(float2));
    auto lane_off = [&](int grp) { return (unsigned)(((size_t)t * stride + B * grp + b) * sizeof(float2)); };
    float2 in[16];
    auto fetch = [&](int grp) {
        const unsigned o = lane_off(grp);
#pragma unroll
        for (int m = 0; m < 16; ++m) in[m] = oph_buf_load_f2(R, o, m * rowstep);
    };
    auto group = [&](int grp) {
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = in[m];
        __syncthreads();  // the previous group's last-stage reads are done
        fft::stage0_from_regs<L, SIGN>(v, buf, t);
        if (grp + steps < gend) fetch(grp + steps);
        __syncthreads();
        fft::middle_stages_tw<L, SIGN>(buf, twm, t);
        float2 wl[16];
        fft::twiddle_powers(wp[t], wp[TPL + t], wl);
        fft::last_stage_to_regs_w<L, SIGN>(buf, wl, t, v);
        const unsigned o = lane_off(grp);
#pragma unroll
        for (int m = 0; m < 16; ++m) oph_buf_store_f2(make_float2(v[m].x * scale, v[m].y * scale), R, o, m * rowstep);
    };
    int g = xcd * per + blockIdx.x / 8;
    if (g < gend) {
        fetch(g);
        group(g);
        for (g += steps; g < gend; g += steps) group(g);
    }
}

template <int L, int SIGN>
int launch_rows(oph_ctx* ctx, float2* data, int rows, int stride, float scale) {
    const float2* tw = oph_twiddles(ctx, L);
    if (!tw) return OPH_ENOMEM;
    if constexpr (fft::regs_fusable<L>() && L / 16 >= 64) {
        if ((size_t)rows * stride * sizeof(float2) < (1u << 31)) {
            // persistent: the resident workgroups, at most 3 per CU (4096 x
            // 4096: 42.9 us at 3, 43.4 at the occupancy maximum of 4, 44.0 at 2)
            static int per_cu = 0;
            if (!per_cu) {
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &per_cu, reinterpret_cast<const void*>(&fft_rows_f<L, SIGN>), L / 16, 0) != hipSuccess ||
                    per_cu < 1) {
                    hipGetLastError();
                    per_cu = 1;
                }
                per_cu = std::min(per_cu, 3);
            }
            const int grid = std::min(rows, per_cu * std::max(ctx->cu_count, 1));
            OPH_LAUNCH(ctx, "fft_rows", (fft_rows_f<L, SIGN>), dim3(grid), dim3(L / 16), 0, data, rows, stride,
                       scale, tw);
            return OPH_OK;
        }
    }
    constexpr int LPW = fftk::row_lines(L);
    const int grid = (rows + LPW - 1) / LPW;
    OPH_LAUNCH(ctx, "fft_rows", (fft_rows_kernel<L, SIGN>), dim3(grid), dim3(LPW * (L / 16)), 0, data, rows,
               stride, scale, tw);
    return OPH_OK;
}

template <int L, int SIGN, int B>
int launch_cols_b(oph_ctx* ctx, float2* data, int ncols, int stride, float scale, const float2* tw) {
    const size_t shm = (size_t)B * fft::padded_len(L) * sizeof(float2);
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_cols_kernel<L, SIGN, B>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        attr = true;
    }
    const int grid = (ncols + B - 1) / B;
    const int wide = stride % 2 == 0 && (size_t)L * stride * sizeof(float2) < (1u << 31);
    OPH_LAUNCH(ctx, "fft_cols", (fft_cols_kernel<L, SIGN, B>), dim3(grid), dim3(B * (L / 16)), shm, data,
               ncols, stride, scale, tw, wide);
    return OPH_OK;
}

template <int SIGN>
int launch_cols_p(oph_ctx* ctx, float2* data, int ncols, int stride, float scale, const float2* tw) {
    const size_t shm = ((size_t)COLP_B * fft::padded_len(COLP_L) + COLP_TW) * sizeof(float2);
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_cols_p<SIGN>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        attr = true;
    }
    // one workgroup per CU (LDS), a multiple of 8, at most one per group
    const int grid = std::max(8, std::min(ncols / COLP_B, std::max(ctx->cu_count, 1)) / 8 * 8);
    OPH_LAUNCH(ctx, "fft_cols", (fft_cols_p<SIGN>), dim3(grid), dim3(COLP_THREADS), shm, data, ncols, stride,
               scale, tw);
    return OPH_OK;
}

template <int L, int SIGN>
int launch_cols(oph_ctx* ctx, float2* data, int ncols, int stride, float scale) {
    const float2* tw = oph_twiddles(ctx, L);
    if (!tw) return OPH_ENOMEM;
    constexpr int BMAX = fftk::col_lines(L);
    static int b = 0;
    if (!b) {
        const char* e = getenv("OPH_FFT_COLUMNS");
        b = e ? atoi(e) : BMAX;
        if (b > BMAX || b < 1 || (b & (b - 1))) b = BMAX;
    }
    if constexpr (L == COLP_L) {
        // the persistent form: four columns, 32-byte pieces, buffer offsets in 32 bits
        if (b == COLP_B && ncols % (8 * COLP_B) == 0 && stride % 2 == 0 &&
            (size_t)L * stride * sizeof(float2) < (1u << 31))
            return launch_cols_p<SIGN>(ctx, data, ncols, stride, scale, tw);
    }
    switch (b) {
        case 1: return launch_cols_b<L, SIGN, 1>(ctx, data, ncols, stride, scale, tw);
        case 2:
            if constexpr (BMAX >= 2) return launch_cols_b<L, SIGN, 2>(ctx, data, ncols, stride, scale, tw);
            break;
        case 4:
            if constexpr (BMAX >= 4) return launch_cols_b<L, SIGN, 4>(ctx, data, ncols, stride, scale, tw);
            break;
        case 8:
            if constexpr (BMAX >= 8) return launch_cols_b<L, SIGN, 8>(ctx, data, ncols, stride, scale, tw);
            break;
    }
    return launch_cols_b<L, SIGN, BMAX>(ctx, data, ncols, stride, scale, tw);
}

template <int SIGN>
int rows_dispatch(oph_ctx* ctx, int L, float2* data, int rows, int stride, float scale) {
    switch (L) {
#define OPH_ROW_CASE(n) \
    case n: return launch_rows<n, SIGN>(ctx, data, rows, stride, scale);
        OPH_FFT_LENGTHS(OPH_ROW_CASE)
#undef OPH_ROW_CASE
    }
    return oph_fail(ctx, OPH_ENOTSUP, "FFT length %d not supported (16..8192)", L);
}

template <int SIGN>
int cols_dispatch(oph_ctx* ctx, int L, float2* data, int ncols, int stride, float scale) {
    switch (L) {
#define OPH_COL_CASE(n) \
    case n: return launch_cols<n, SIGN>(ctx, data, ncols, stride, scale);
        OPH_FFT_LENGTHS(OPH_COL_CASE)
#undef OPH_COL_CASE
    }
    return oph_fail(ctx, OPH_ENOTSUP, "FFT length %d not supported (16..8192)", L);
}

}  // namespace

int oph_fft_rows(oph_ctx* ctx, int L, int sign, float2* data, int rows, int stride, float scale) {
    return sign < 0 ? rows_dispatch<-1>(ctx, L, data, rows, stride, scale)
                    : rows_dispatch<+1>(ctx, L, data, rows, stride, scale);
}

int oph_fft_cols(oph_ctx* ctx, int L, int sign, float2* data, int ncols, int stride, float scale) {
    return sign < 0 ? cols_dispatch<-1>(ctx, L, data, ncols, stride, scale)
                    : cols_dispatch<+1>(ctx, L, data, ncols, stride, scale);
}

extern "C" int oph_fft2d(oph_ctx* ctx, oph_c64* data, int nx, int ny, int direction, int normalize,
                         int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!data) return oph_fail(ctx, OPH_EINVAL, "null data");
    if (direction != -1 && direction != 1) return oph_fail(ctx, OPH_EINVAL, "direction must be -1 or +1");
    if (!fftk::supported(nx) || !fftk::supported(ny))
        return oph_fail(ctx, OPH_ENOTSUP, "FFT size %dx%d: each side must be a power of two in 16..8192", nx, ny);
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    const size_t bytes = (size_t)nx * ny * sizeof(float2);
    float2* d = reinterpret_cast<float2*>(data);
    if (!(flags & OPH_DEVICE_PTR)) {
        d = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK0, bytes));
        if (!d) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(d, data, bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    const float scale = normalize ? (float)(1.0 / ((double)nx * ny)) : 1.f;
    if ((rc = oph_fft_rows(ctx, nx, direction, d, ny, nx, 1.f))) return rc;
    if ((rc = oph_fft_cols(ctx, ny, direction, d, nx, nx, scale))) return rc;
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(data, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}
