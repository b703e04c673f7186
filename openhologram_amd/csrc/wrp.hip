// wrp.hip -- wavefront-recording-plane accumulation on gfx950, and the full
// WRP hologram pipeline (WRP, then angular-spectrum propagation).
//
// Replaces OPHWRP::SingleWRP (W/openwrp/WRP.cpp:75-142, with CWO::AddPixel,
// CWO/cwo.cpp:4298-4309).  Per point n, with dz = z_wrp - z_n:
//   centre  tx = (int)(x_n / p) + Nx/2, ty = (int)(y_n / p) + Ny/2  (:114-115)
//   window  wx, wy in [-w, w)                                      (:119-120)
//   r = sign(dz) sqrt((wx p)^2 + (wy p)^2 + dz^2)                   (:122-128)
//   wrp[ty+wy][tx+wx] += a_n exp(i k r) / (r + 0.05)                (:131-136)
// Half-width w: mode 0 the grating-angle radius (int)(|dz| tan(asin(lambda/
// (2p)))/p) (CWO/cwoPLS.cpp:613), mode 1 the reference's
// (int)|lambda dz/(2p^2) + 0.5| * 2 - 1 (WRP.cpp:111).  Mode 2 is CWO's own
// statement cwoPLS::WRP (CWO/cwoPLS.cpp:522-591): Fresnel phase
// r = dz + rho^2/(2 dz), amplitude a (no 1/(r + 0.05)), half-width
// (int)(|dz| lambda/(2p)/p) (:557; the reference takes |dz| from its call
// argument, which is dz when the points sit at z = 0).  Mode 3 is the
// MATLAB statement (OpenHolo_WRP/WRP_matlab_ver 1.0/Hologram_WRP.m:19-36):
// full width N = round(|lambda dz/p^2/2| + 0.5) 2 - 1, window
// [-(N-1)/2, (N-1)/2], centre round(x_n/p) + (Nx-1)/2 - 1, amplitude
// a_n/r, and the point's phase (the script's random phase) added.
//
// The reference scatters with a racing AddPixel; here the plane is gathered
// per 64x64 tile, with no global atomics:
//   1. wrp_bin_count   per chunk of points: the point records (centre,
//                      window, FP64-reduced phase constants) and per-tile
//                      counts in LDS;
//   2. wrp_bin_colscan per tile: prefix of the counts over chunks;
//   3. wrp_bin_fill    per chunk: tile offsets (scan of the tile totals, in
//                      LDS), then point indices into the tile lists;
//   4. wrp_gather      one workgroup per tile: each wave takes points of the
//                      list; a window value depends on (|wx|, |wy|) only, so
//                      each is evaluated once and added to its mirrored
//                      pixels; contributions are summed in LDS as 32-bit
//                      fixed point scaled per list segment so the sum
//                      cannot overflow (integer adds are associative, so the
//                      result does not depend on the list order:
//                      deterministic without sorting).
// Per pixel, with u = (wx^2 + wy^2) p^2 / dz^2 (<= 2 tan^2 ~ 2.2e-3 for the
// grating-angle window):
//   (r - |dz|)/|dz| = P(u) = u/2 - u^2/8 + u^3/16 - 5u^4/128,
//   phase (turns) = frac(sign |dz|/lambda + phi) + sign |dz|/lambda P(u),
//   1/(r + 0.05)  = 1/D (1 - sign |dz| P/D),  D = sign |dz| + 0.05,
// so no square root or division per pixel; points whose window makes the
// truncated series worse than 1e-7 turns use an exact FP64 evaluation.
#include <cmath>
#include <mutex>
#include <type_traits>

#include "oph_internal.h"
#include "pipelines.h"

namespace {

constexpr int TILE = 64;
// threads per gather workgroup: 8 waves per tile (4 tiles fit per CU by LDS,
// so 8 waves per SIMD hide the LDS-atomic and transcendental latencies)
constexpr int GATHER_THREADS = 512;

// evaluation path of a point (uniform over its window)
constexpr int PATH_SERIES = 0, PATH_EXACT = 1, PATH_FRESNEL = 2, PATH_FRESNEL_EXACT = 3;

struct WrpPoint {
    int tx, ty, w, exact;  // exact: one of PATH_*
    float f0;     // frac(sign |dz| / lambda + phase/2pi), turns
    float ks;     // sign |dz| / lambda
    float p2dz2;  // p^2 / dz^2
    float ampd;   // a / D
    float cd;     // -sign |dz| / D
    float pad0, pad1, pad2;
    double dz;    // for the exact path
    double ph;    // initial phase (turns), mode 3; 0 otherwise
};

struct WrpArgs {
    const double* x;
    const double* y;
    const double* z;
    const double* amp;
    const double* phase;     // radians, mode 3 only (may be null)
    int64_t n;
    int nx, ny, ntx, nty, ntiles;
    double pitch, inv_lambda, z_wrp, grating, lambda;
    int mode;
    WrpPoint* pts;
    unsigned* chunk_counts;  // [ntiles][nchunks]: counts, then exclusive prefix over chunks
    int nchunks;
    int chunk;               // points per binning workgroup, a multiple of BIN_THREADS
    unsigned* totals;        // [ntiles]
    unsigned* offsets;       // [ntiles + 1]
    float* chunk_max;        // [nchunks] max |a/D|
    float* gmax;             // [1]
    unsigned* list;
    unsigned capacity;
    unsigned* overflow;      // this call's flag (cleared by wrp_bin_colscan)
    unsigned* sticky;        // the context's flag for asynchronous calls, or null
};

constexpr int BIN_THREADS = 256;
constexpr int MAX_CHUNKS = 1024;     // chunks per call (bounds the [tile][chunk] count matrix)
constexpr int MAX_TILES = 16384;     // LDS counters of one binning workgroup

// points per binning workgroup for n points: one per thread while that keeps
// the chunk count at most MAX_CHUNKS
__host__ __device__ constexpr int64_t wrp_chunk(int64_t np) {
    return ((np + MAX_CHUNKS - 1) / MAX_CHUNKS + BIN_THREADS - 1) / BIN_THREADS * BIN_THREADS;
}
// words of binning bookkeeping: chunk counts, totals, offsets (ntiles + 1),
// overflow, chunk maxima, global max
__host__ __device__ constexpr size_t wrp_book_words(int64_t nchunks, int64_t ntiles) {
    return (size_t)nchunks * ntiles + 2 * (size_t)ntiles + 2 + nchunks + 1;
}
// first tile-list capacity (entries) for n points
__host__ __device__ constexpr size_t wrp_list_entries(int64_t np) { return (size_t)24 * np + 65536; }

__device__ __forceinline__ void tile_range(const WrpPoint& q, const WrpArgs& a, int& x0, int& x1, int& y0,
                                           int& y1) {
    // window [t - w, t + w), or [t - w + 1, t + w) for mode 3 (odd width)
    const int lo_add = a.mode == 3 ? 1 : 0;
    int lo = q.tx - q.w + lo_add, hi = q.tx + q.w - 1;
    if (lo < 0) lo = 0;
    if (hi > a.nx - 1) hi = a.nx - 1;
    x0 = lo / TILE;
    x1 = hi / TILE;
    if (lo > hi) x1 = x0 - 1;
    lo = q.ty - q.w + lo_add;
    hi = q.ty + q.w - 1;
    if (lo < 0) lo = 0;
    if (hi > a.ny - 1) hi = a.ny - 1;
    y0 = lo / TILE;
    y1 = hi / TILE;
    if (lo > hi) y1 = y0 - 1;
}

__device__ WrpPoint make_point(const WrpArgs& a, int64_t i) {
    // every mode's quantities are formed and the mode's selected at the end
    // (no branch on the mode: one straight-line path for all four)
    const double dz = a.z_wrp - a.z[i];
    const bool fres = a.mode == 2, mat = a.mode == 3;
    WrpPoint q;
    // centres and half-widths are clamped so that tx +- w stays inside int:
    // |tx|, |ty| <= 2^29 and w <= 2^30 (a window that large covers the whole
    // plane of any allowed size from any clamped centre that can reach it).
    // Mode 3 rounds the centre (the script's round(x/p) + (res-1)/2, 1-based)
    const double xs = a.x[i] / a.pitch, ys = a.y[i] / a.pitch;
    const double cx = fmin(fmax(mat ? round(xs) : trunc(xs), -536870912.0), 536870912.0);
    const double cy = fmin(fmax(mat ? round(ys) : trunc(ys), -536870912.0), 536870912.0);
    q.tx = (int)cx + (mat ? (a.nx - 1) / 2 - 1 : (a.nx >> 1));
    q.ty = (int)cy + (mat ? (a.ny - 1) / 2 - 1 : (a.ny >> 1));
    const double w0 = trunc(fabs(dz) * a.grating / a.pitch);
    const double w1 = trunc(fabs(a.lambda * dz / a.pitch / a.pitch / 2 + 0.5)) * 2 - 1;
    const double w2 = trunc(fabs(dz) * (a.lambda / (2.0 * a.pitch)) / a.pitch);
    // mode 3: full width N = round(|lambda dz/p^2/2| + 0.5) 2 - 1, stored as
    // w = (N - 1)/2 + 1 over [t - w + 1, t + w); nothing for dz = 0 (r = 0)
    const double n3 = round(fabs(a.lambda * dz / (a.pitch * a.pitch) / 2) + 0.5) * 2 - 1;
    const double w3 = dz == 0.0 ? 0.0 : (n3 - 1) / 2 + 1;
    const double wd = a.mode == 0 ? w0 : a.mode == 1 ? w1 : a.mode == 2 ? w2 : w3;
    q.w = (int)fmin(fmax(wd, 0.0), 1073741824.0);
    const double sgn = dz > 0.0 ? 1.0 : -1.0;
    const double adz = fabs(dz);
    const double am = a.amp ? a.amp[i] : 1.0;
    const double ph = (mat && a.phase) ? a.phase[i] * (1.0 / (2.0 * M_PI)) : 0.0;
    double t = sgn * adz * a.inv_lambda + ph;
    t -= rint(t);
    q.f0 = (float)t;
    q.ks = (float)(sgn * adz * a.inv_lambda);
    const double p2 = adz > 0.0 ? a.pitch * a.pitch / (dz * dz) : 0.0;
    q.p2dz2 = (float)p2;
    // spherical amplitude a/(r + off) = a/D (1 + cd P), D = sign|dz| + off:
    // off = 0.05 (WRP.cpp) or 0 (mode 3, a/r); Fresnel (mode 2): a
    const double D = sgn * adz + (mat ? 0.0 : 0.05);
    q.ampd = fres ? (float)am : (float)(am / D);
    q.cd = fres ? 0.f : (float)(-sgn * adz / D);
    q.pad0 = q.pad1 = q.pad2 = 0.f;
    q.dz = dz;
    q.ph = ph;
    const double umax = 2.0 * (double)q.w * q.w * p2;
    // spherical: series error at the window corner, (|dz|/lambda) 7/256 u^5
    // turns, and the amplitude's dropped second-order term (cd P)^2
    const double u5 = umax * umax * umax * umax * umax;
    const bool sph_exact =
        umax > 0.05 || adz * a.inv_lambda * (7.0 / 256.0) * u5 > 1e-7 || fabs(-sgn * adz / D) * 0.5 * umax > 1e-3;
    // Fresnel: phase (dz + rho^2/(2 dz))/lambda = f0 + ks u/2, u = rho^2/dz^2,
    // in FP32 while the quadratic part stays within 16 turns at the corner
    const bool fr_exact = adz * a.inv_lambda * 0.5 * umax > 16.0 || adz == 0.0;
    q.exact = fres ? (fr_exact ? PATH_FRESNEL_EXACT : PATH_FRESNEL) : (sph_exact ? PATH_EXACT : PATH_SERIES);
    return q;
}

// pass 1: per chunk of points, the point records and per-tile counts
__global__ __launch_bounds__(BIN_THREADS) void wrp_bin_count(WrpArgs a) {
    __shared__ unsigned cnt[MAX_TILES];
    __shared__ float wmax[BIN_THREADS / 64];
    for (int t = threadIdx.x; t < a.ntiles; t += BIN_THREADS) cnt[t] = 0u;
    __syncthreads();
    // chunks numbered XCD by XCD, so that neighbouring columns of the count
    // matrix are written through the same L2
    const int chunk = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t base = (int64_t)chunk * a.chunk;
    float mx = 0.f;
    for (int k = threadIdx.x; k < a.chunk; k += BIN_THREADS) {
        const int64_t i = base + k;
        if (i >= a.n) break;
        const WrpPoint q = make_point(a, i);
        a.pts[i] = q;
        if (q.w <= 0) continue;
        mx = fmaxf(mx, fabsf(q.ampd));
        int x0, x1, y0, y1;
        tile_range(q, a, x0, x1, y0, y1);
        for (int ty = y0; ty <= y1; ++ty)
            for (int tx = x0; tx <= x1; ++tx) atomicAdd(&cnt[ty * a.ntx + tx], 1u);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    for (int t = threadIdx.x; t < a.ntiles; t += BIN_THREADS)
        a.chunk_counts[(size_t)t * a.nchunks + chunk] = cnt[t];
    if (threadIdx.x == 0) {
        float m = 0.f;
        for (int w = 0; w < BIN_THREADS / 64; ++w) m = fmaxf(m, wmax[w]);
        a.chunk_max[chunk] = m;
    }
}

// pass 2: per tile (one wave each), exclusive prefix of the chunk counts and
// the tile total; the first wave also reduces the chunk maxima
__global__ __launch_bounds__(256) void wrp_bin_colscan(WrpArgs a) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave == 0) {
        float m = 0.f;
        for (int c = lane; c < a.nchunks; c += 64) m = fmaxf(m, a.chunk_max[c]);
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (lane == 0) {
            *a.gmax = m;
            *a.overflow = 0u;
        }
    }
    if (wave >= a.ntiles) return;
    unsigned* p = a.chunk_counts + (size_t)wave * a.nchunks;
    // all of the tile's counts are loaded first (one memory round trip),
    // then scanned 64 at a time
    constexpr int PER_LANE = MAX_CHUNKS / 64;
    unsigned v[PER_LANE];
#pragma unroll
    for (int j = 0; j < PER_LANE; ++j) {
        const int c = j * 64 + lane;
        v[j] = c < a.nchunks ? p[c] : 0u;
    }
    unsigned run = 0;
#pragma unroll
    for (int j = 0; j < PER_LANE; ++j) {
        if (j * 64 >= a.nchunks) break;
        unsigned incl = v[j];
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned o = __shfl_up(incl, off);
            if (lane >= off) incl += o;
        }
        const int c = j * 64 + lane;
        if (c < a.nchunks) p[c] = run + incl - v[j];
        run += __shfl(incl, 63);
    }
    if (lane == 0) a.totals[wave] = run;
}

// pass 3: per chunk, point indices into the tile lists (slot = tile offset +
// chunk prefix + rank within the chunk from an LDS counter).  Every
// workgroup forms the tile offsets (exclusive prefix of the tile totals) in
// LDS itself and adds its own chunk's prefixes, so the point loop reads only
// LDS; chunk 0 also stores the offsets for the gather.  Dynamic LDS: 2
// ntiles words.
__global__ __launch_bounds__(BIN_THREADS) void wrp_bin_fill(WrpArgs a) {
    extern __shared__ unsigned fill_lds[];
    __shared__ unsigned part[BIN_THREADS];
    unsigned* cur = fill_lds;
    unsigned* base = fill_lds + a.ntiles;  // tile offset + this chunk's prefix
    const int n = a.ntiles, t = threadIdx.x;
    const int per = (n + BIN_THREADS - 1) / BIN_THREADS;
    const int b = min(n, t * per), e = min(n, b + per);
    const int chunk = xcd_remap(blockIdx.x, gridDim.x);
    // The thread's tile totals and this chunk's prefixes in those tiles, read
    // once; up to FAST tiles per thread the loads are all issued before any
    // is used (range-checked buffer loads, out-of-range entries read 0)
    constexpr int FAST = 4;
    unsigned tv[FAST], cv[FAST];
    unsigned sum = 0;
    if (per <= FAST) {
        const auto rt = oph_buf_rsrc(a.totals, (unsigned)(n * sizeof(unsigned)));
        const auto rc = oph_buf_rsrc(a.chunk_counts, (unsigned)((size_t)n * a.nchunks * sizeof(unsigned)));
#pragma unroll
        for (int j = 0; j < FAST; ++j) {
            const int i = b + j;
            const bool in = j < per && i < e;
            tv[j] = __builtin_amdgcn_raw_buffer_load_b32(rt, in ? (unsigned)(i * sizeof(unsigned)) : OPH_BUF_OOB, 0, 0);
            cv[j] = __builtin_amdgcn_raw_buffer_load_b32(
                rc, in ? (unsigned)(((size_t)i * a.nchunks + chunk) * sizeof(unsigned)) : OPH_BUF_OOB, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < FAST; ++j) sum += tv[j];
    } else {
        for (int i = b; i < e; ++i) sum += a.totals[i];
    }
    for (int i = b; i < e; ++i) cur[i] = 0u;
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < BIN_THREADS; off <<= 1) {
        const unsigned v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned run = part[t] - sum;  // exclusive prefix of this thread's range
    if (per <= FAST) {
#pragma unroll
        for (int j = 0; j < FAST; ++j) {
            const int i = b + j;
            if (j < per && i < e) {
                base[i] = run + cv[j];
                if (chunk == 0) a.offsets[i] = run;
                run += tv[j];
            }
        }
    } else {
        for (int i = b; i < e; ++i) {
            base[i] = run + a.chunk_counts[(size_t)i * a.nchunks + chunk];
            if (chunk == 0) a.offsets[i] = run;
            run += a.totals[i];
        }
    }
    if (chunk == 0 && t == BIN_THREADS - 1) a.offsets[n] = part[BIN_THREADS - 1];
    __syncthreads();
    const int64_t pbase = (int64_t)chunk * a.chunk;
    // the chunk's records; a record's first 16 bytes (tx, ty, w, exact) in
    // one load ahead of the test on w
    const auto rp = oph_buf_rsrc(a.pts + pbase, (unsigned)((size_t)a.chunk * sizeof(WrpPoint)));
    for (int k = threadIdx.x; k < a.chunk; k += BIN_THREADS) {
        const int64_t i = pbase + k;
        if (i >= a.n) break;
        const float4 hf = oph_buf_load_f4(rp, (unsigned)(k * sizeof(WrpPoint)));
        WrpPoint q;
        q.tx = __float_as_int(hf.x);
        q.ty = __float_as_int(hf.y);
        q.w = __float_as_int(hf.z);
        if (q.w <= 0) continue;
        int x0, x1, y0, y1;
        tile_range(q, a, x0, x1, y0, y1);
        for (int ty = y0; ty <= y1; ++ty)
            for (int tx = x0; tx <= x1; ++tx) {
                const int tile = ty * a.ntx + tx;
                const unsigned slot = base[tile] + atomicAdd(&cur[tile], 1u);
                if (slot < a.capacity) {
                    a.list[slot] = (unsigned)i;
                } else {
                    *a.overflow = 1u;
                    if (a.sticky) *a.sticky = 1u;
                }
            }
    }
}

struct GatherArgs {
    const WrpPoint* pts;
    const unsigned* offsets;
    const unsigned* list;
    const unsigned* chunk_pre;  // [ntiles][nchunks] exclusive prefix over chunks
    const unsigned* overflow;
    const float* gmax;
    int nchunks;
    unsigned capacity;
    int nx, ny, ntx;
    double pitch, inv_lambda;
    float r_off;     // spherical amplitude a/(r + r_off): 0.05, or 0 for mode 3
    int lo_add;      // 1 for mode 3 (window [t - w + 1, t + w)), else 0
    float2* wrp;
    int accumulate;
};

// |w| range [lo, hi] covered by window offsets w in [w0, w1)
__device__ __forceinline__ void abs_range(int w0, int w1, int& lo, int& hi) {
    if (w0 >= 0) {
        lo = w0;
        hi = w1 - 1;
    } else if (w1 <= 1) {
        lo = 1 - w1;
        hi = -w0;
    } else {
        lo = 0;
        hi = max(-w0, w1 - 1);
    }
}

// Per (point, tile) setup of the gather, formed once by the thread that
// stages the point (one lane per point) and read by the wave that walks the
// point's window (an LDS broadcast), so the wave-uniform work of a window --
// its clipping to the tile, the walk's steps, the fixed-point scale and the
// evaluation coefficients -- is not repeated on 64 lanes.  Series paths
// (spherical and Fresnel) evaluate, with v = wx^2 + wy^2,
//   phase = c0 + v (c1 + v (c2 + v (c3 + v c4)))        (turns)
//   amplitude = a0 + v (a1 + v a2)
// where for the spherical modes the coefficients fold ks, p^2/dz^2, the
// series of P(u) and the amplitude slope cd in (the amplitude keeps P's first
// two terms; the next is below 1e-10 relative for the windows this path
// takes), and for the Fresnel mode c1 = ks p^2/(2 dz^2), the rest 0.  The
// exact paths use a0 (amplitude x segment scale) and the point record.
struct GSetup {
    int ax0, ay0, na, nay;        // |wx| in [ax0, ax0 + na), |wy| in [ay0, ay0 + nay); na = 0: no overlap
    int sx, sy;                   // advancing 64 walk positions = sy rows plus sx columns
    float inv_na;
    int cbyte;                    // byte offset of the window centre in the accumulator
    int wx0, wy0;                 // this tile's part of the window offsets: w in [w0, w0 + dw)
    unsigned dwx, dwy, nwx, nwy;  // mirror tests: |w| - w0 < dw (positive side), |w| - 1 < nw (negative)
    float c0, c1, c2, c3, c4;
    float a0, a1, a2;
    unsigned ip;                  // point index << 2 | evaluation path
    unsigned pad;
};
constexpr int REC_BATCH = 64;

// The window value depends on (|wx|, |wy|) only: each (ax, ay) is evaluated
// once and added to the up-to-four mirrored pixels (+-ax, +-ay) that lie in
// this tile's part of the window.  Sums are 32-bit fixed point with a scale
// chosen so that (points summed) x (largest contribution) fits in 2^30:
// exact integer adds, independent of the list order.  Both parts of a pixel
// share one 64-bit word, 2^32 im + re, added with one 64-bit LDS atomic: the
// word is exact modulo 2^64, and with |sum re|, |sum im| < 2^31 re is its
// sign-extended low half and im = (word - re) >> 32.  A tile's list is
// summed in segments of at most SEGMENT points (one segment for all but
// dense tiles), each with its own scale, so the fixed-point unit stays below
// 2^-19 of the largest contribution; segment sums are added to the tile's
// output in segment order.  Segments end on binning-chunk boundaries (the
// largest boundary within SEGMENT entries, or the next one when a single
// chunk holds more): the order of entries inside a chunk's run comes from an
// LDS atomic in wrp_bin_fill, but which points form each segment then does
// not depend on it, so the result is deterministic.
constexpr unsigned SEGMENT = 2048;

// End (relative to the tile's list start) of the segment that starts at the
// chunk boundary r0: pre[c] is the exclusive prefix of chunk c's entries in
// this tile (c < nchunks), and total the tile's entry count.
__device__ unsigned segment_end(const unsigned* pre, int nchunks, unsigned total, unsigned r0) {
    const unsigned target = r0 + SEGMENT;
    if (target >= total) return total;
    // largest c with pre[c] <= target (pre[0] = 0)
    int lo = 0, hi = nchunks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= target)
            lo = mid;
        else
            hi = mid - 1;
    }
    if (pre[lo] > r0) return pre[lo];
    // the chunk starting at r0 alone exceeds SEGMENT: the next boundary above r0
    lo = 0;
    hi = nchunks;  // boundary nchunks is total > r0
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const unsigned v = mid < nchunks ? pre[mid] : total;
        if (v > r0)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo < nchunks ? pre[lo] : total;
}

// FRESNEL: the call's window mode is 2 (Fresnel phase, unit amplitude
// factor); the spherical modes 0 and 1 use the other instantiation, so each
// carries only its own two evaluation paths.
template <bool FRESNEL_MODE>
__global__ __launch_bounds__(GATHER_THREADS) void wrp_gather(GatherArgs g) {
    __shared__ unsigned long long acc[TILE * TILE];
    __shared__ GSetup recs[REC_BATCH];
    __shared__ unsigned seg_hi;
    // an overflowed binning leaves the output untouched (the call is redone
    // at the exact size, or reported)
    if (*g.overflow) return;
    const int tile = blockIdx.x;
    const int tyi = tile / g.ntx, txi = tile % g.ntx;
    const int X0 = txi * TILE, Y0 = tyi * TILE;
    unsigned b = g.offsets[tile], e = g.offsets[tile + 1];
    if (e > g.capacity) e = g.capacity;
    if (b > e) b = e;
    const float mx = *g.gmax;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int NW = GATHER_THREADS / 64;
    const unsigned* pre = g.chunk_pre + (size_t)tile * g.nchunks;
    for (unsigned sb = b;;) {
        __syncthreads();  // the previous segment's output reads are done
        if (threadIdx.x == 0) seg_hi = e - b <= SEGMENT ? e : b + segment_end(pre, g.nchunks, e - b, sb - b);
        __syncthreads();
        const unsigned se = seg_hi;
        const float bound = mx * (float)max(se - sb, 1u);
        const int S = bound > 0.f ? max(-60, min(60, 30 - (int)ceilf(log2f(bound)))) : 30;
        const float scale = ldexpf(1.f, S);
        for (int i = threadIdx.x; i < TILE * TILE; i += GATHER_THREADS) acc[i] = 0ull;
        for (unsigned bs = sb; bs < se; bs += REC_BATCH) {
            const int cnt_b = (int)min((unsigned)REC_BATCH, se - bs);
            __syncthreads();
            if ((int)threadIdx.x < cnt_b) {
                const unsigned idx = g.list[bs + threadIdx.x];
                const WrpPoint q = g.pts[idx];
                GSetup r;
                const int tx = q.tx, ty = q.ty, w = q.w;
                // window offsets of this tile's part: wx in [wx0, wx1), wy in [wy0, wy1)
                const int wx0 = max(tx - w + g.lo_add, max(X0, 0)) - tx, wx1 = min(tx + w, min(X0 + TILE, g.nx)) - tx;
                const int wy0 = max(ty - w + g.lo_add, max(Y0, 0)) - ty, wy1 = min(ty + w, min(Y0 + TILE, g.ny)) - ty;
                int ax0 = 0, ax1 = -1, ay0 = 0, ay1 = -1;
                if (wx1 > wx0 && wy1 > wy0) {
                    abs_range(wx0, wx1, ax0, ax1);
                    abs_range(wy0, wy1, ay0, ay1);
                }
                const int na = ax1 - ax0 + 1;
                r.ax0 = ax0;
                r.ay0 = ay0;
                r.na = na;
                r.nay = ay1 - ay0 + 1;
                // lane positions qq = lane + 64 j walk the na x nay grid of
                // (|wx|, |wy|) row by row; advancing 64 positions is sy rows
                // plus sx columns, so no division or multiply per position
                const float inv_na = na > 0 ? 1.f / (float)na : 0.f;
                r.inv_na = inv_na;
                r.sy = (int)(64.5f * inv_na);
                r.sx = 64 - r.sy * na;
                r.cbyte = ((ty - Y0) * TILE + (tx - X0)) * 8;
                r.wx0 = wx0;
                r.wy0 = wy0;
                r.dwx = (unsigned)(wx1 - wx0);
                r.dwy = (unsigned)(wy1 - wy0);
                r.nwx = (unsigned)max(-wx0, 0);
                r.nwy = (unsigned)max(-wy0, 0);
                // the point's amplitude with the segment's fixed-point scale folded in
                const float amps = q.ampd * scale;
                const float P1 = q.p2dz2, P2 = P1 * P1, P3 = P2 * P1, P4 = P3 * P1;
                r.c0 = q.f0;
                r.a0 = amps;
                if constexpr (FRESNEL_MODE) {
                    r.c1 = q.ks * 0.5f * P1;
                    r.c2 = r.c3 = r.c4 = 0.f;
                    r.a1 = r.a2 = 0.f;
                } else {
                    r.c1 = q.ks * 0.5f * P1;
                    r.c2 = q.ks * -0.125f * P2;
                    r.c3 = q.ks * (1.f / 16.f) * P3;
                    r.c4 = q.ks * (-5.f / 128.f) * P4;
                    r.a1 = amps * q.cd * 0.5f * P1;
                    r.a2 = amps * q.cd * -0.125f * P2;
                }
                r.ip = idx << 2 | (unsigned)q.exact;
                r.pad = 0u;
                recs[threadIdx.x] = r;
            }
            __syncthreads();
            for (int kk = wave; kk < cnt_b; kk += NW) {
                const GSetup q = recs[kk];
                if (q.na <= 0) continue;
                const int na = q.na, nay = q.nay, ax0 = q.ax0, ay0 = q.ay0, sx = q.sx, sy = q.sy;
                int iy = (int)(((float)lane + 0.5f) * q.inv_na);
                int ix = lane - iy * na;
                const int cbyte = q.cbyte;
                const unsigned qidx = q.ip >> 2;
                const int qpath = (int)(q.ip & 3u);
                const float amps = q.a0;
                const int wx0 = q.wx0, wy0 = q.wy0;
                const unsigned dwx = q.dwx, dwy = q.dwy, nwx = q.nwx, nwy = q.nwy;
                // the setup's LDS reads complete here, before the walk: left to
                // the first use inside the loop, the wait would also cover the
                // previous position's LDS atomics on every iteration
                __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
                // one loop per evaluation path (the path is uniform over the point)
                auto walk = [&](auto exact_tag, auto fresnel_tag) {
                    constexpr bool EXACT = decltype(exact_tag)::value;
                    constexpr bool FRESNEL = decltype(fresnel_tag)::value;
                    while (iy < nay) {
                        const int ax = ax0 + ix, ay = ay0 + iy;
                        float re, im;
                        if constexpr (!EXACT) {
                            // (24-bit multiplies: full rate, where a 32-bit one is quarter rate)
                            const float v = (float)(__umul24(ax, ax) + __umul24(ay, ay));
                            const float ph = fmaf(v, fmaf(v, fmaf(v, fmaf(v, q.c4, q.c3), q.c2), q.c1), q.c0);
                            const float am = fmaf(v, fmaf(v, q.a2, q.a1), q.a0);
                            re = am * cos_turns(ph);
                            im = am * sin_turns(ph);
                        } else if constexpr (FRESNEL) {
                            const double dz = g.pts[qidx].dz;
                            const double dx = ax * g.pitch, dy = ay * g.pitch;
                            double t = (dz + (dx * dx + dy * dy) / (2.0 * dz)) * g.inv_lambda;
                            t -= rint(t);
                            re = amps * cos_turns((float)t);
                            im = amps * sin_turns((float)t);
                        } else {
                            const double dz = g.pts[qidx].dz;
                            const double dx = ax * g.pitch, dy = ay * g.pitch;
                            const double sgn = dz > 0.0 ? 1.0 : -1.0;
                            const double r = sgn * sqrt(dx * dx + dy * dy + dz * dz);
                            double t = r * g.inv_lambda + g.pts[qidx].ph;
                            t -= rint(t);
                            // amplitude a/(r + off) = (a/D) D/(r + off), D = sign|dz| + off
                            const double off = (double)g.r_off;
                            const double D = sgn * fabs(dz) + off;
                            const float am = (float)((double)amps * D / (r + off));
                            re = am * cos_turns((float)t);
                            im = am * sin_turns((float)t);
                        }
                        // fixed point, truncated toward zero (any fixed rounding
                        // keeps the sum exact; the unit is 2^-30 of the bound)
                        const int ire = (int)re;
                        const int iim = (int)im;
                        // 2^32 im + re modulo 2^64
                        const unsigned long long pk =
                            (unsigned long long)(long long)ire + ((unsigned long long)(long long)iim << 32);
                        const bool xp = (unsigned)(ax - wx0) < dwx, xm = (unsigned)(ax - 1) < nwx;
                        const bool yp = (unsigned)(ay - wy0) < dwy, ym = (unsigned)(ay - 1) < nwy;
                        // byte offsets of the mirrored pixels (8 bytes per pixel)
                        const int rp = cbyte + (ay << 9), rm = cbyte - (ay << 9), cx = ax << 3;
                        char* ba = reinterpret_cast<char*>(acc);
                        if (xp && yp) atomicAdd(reinterpret_cast<unsigned long long*>(ba + rp + cx), pk);
                        if (xm && yp) atomicAdd(reinterpret_cast<unsigned long long*>(ba + rp - cx), pk);
                        if (xp && ym) atomicAdd(reinterpret_cast<unsigned long long*>(ba + rm + cx), pk);
                        if (xm && ym) atomicAdd(reinterpret_cast<unsigned long long*>(ba + rm - cx), pk);
                        ix += sx;
                        iy += sy;
                        if (ix >= na) {
                            ix -= na;
                            ++iy;
                        }
                    }
                };
                if constexpr (FRESNEL_MODE) {
                    if (qpath == PATH_FRESNEL)
                        walk(std::false_type{}, std::true_type{});
                    else
                        walk(std::true_type{}, std::true_type{});
                } else {
                    if (qpath == PATH_SERIES)
                        walk(std::false_type{}, std::false_type{});
                    else
                        walk(std::true_type{}, std::false_type{});
                }
            }
        }
        __syncthreads();
        // the segment's sums: written (first segment, unless accumulating)
        // or added to the tile's output
        const bool add = sb != b || g.accumulate;
        const float inv = ldexpf(1.f, -S);
        for (int p = threadIdx.x; p < TILE * TILE; p += GATHER_THREADS) {
            const int row = Y0 + p / TILE, col = X0 + p % TILE;
            if (row >= g.ny || col >= g.nx) continue;
            const unsigned long long wv = acc[p];
            const int sre = (int)(unsigned)wv;
            const int sim = (int)((long long)(wv - (unsigned long long)(long long)sre) >> 32);
            const float re = (float)sre * inv;
            const float im = (float)sim * inv;
            float2* o = g.wrp + (size_t)row * g.nx + col;
            if (add) {
                const float2 v = *o;
                *o = make_float2(re + v.x, im + v.y);
            } else {
                *o = make_float2(re, im);
            }
        }
        if (se >= e) break;
        sb = se;
    }
}

}  // namespace

namespace {

// Where one WRP call keeps its binning state.  From the context's grow-only
// slots (oph_wrp_device) or from the static array of the allocation-free
// entry point (fixed: a list overflow is then an error, not a retry).
struct WrpScratch {
    bool fixed;          // true: pts/book/list given, capacity list_cap
    WrpPoint* pts;
    unsigned* book;
    unsigned* list;
    size_t list_cap;     // entries
};

int wrp_core(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, double z_wrp, int mode,
             const double* x, const double* y, const double* z, const double* amp, const double* phase,
             int64_t n, float2* wrp, int accumulate, int check, const WrpScratch& ws) {
    WrpArgs a{};
    a.x = x;
    a.y = y;
    a.z = z;
    a.amp = amp;
    a.phase = phase;
    a.n = n;
    a.nx = nx;
    a.ny = ny;
    a.ntx = (nx + TILE - 1) / TILE;
    a.nty = (ny + TILE - 1) / TILE;
    a.ntiles = a.ntx * a.nty;
    if (a.ntiles > MAX_TILES)
        return oph_fail(ctx, OPH_ENOTSUP, "WRP plane of %dx%d exceeds %d tiles of 64x64", nx, ny, MAX_TILES);
    a.pitch = pitch;
    a.lambda = lambda;
    a.inv_lambda = 1.0 / lambda;
    a.z_wrp = z_wrp;
    a.grating = tan(asin(lambda / (2.0 * pitch)));
    a.mode = mode;
    const int ntiles = a.ntiles;
    if (n > (int64_t)1 << 30)
        return oph_fail(ctx, OPH_EINVAL, "%lld points: a WRP call takes at most 2^30", (long long)n);
    const int64_t np = n > 0 ? n : 1;
    // one point per thread while that keeps the chunk count bounded
    const int64_t chunk = wrp_chunk(np);
    a.chunk = (int)chunk;
    const int nchunks = (int)((np + chunk - 1) / chunk);
    a.nchunks = nchunks;
    // bookkeeping: chunk counts, totals, offsets (ntiles + 1), overflow, chunk
    // maxima, global max
    const size_t book = wrp_book_words(nchunks, ntiles);
    unsigned* bk;
    if (ws.fixed) {
        a.pts = ws.pts;
        bk = ws.book;
    } else {
        a.pts = static_cast<WrpPoint*>(oph_scratch(ctx, SLOT_AUX0, sizeof(WrpPoint) * np));
        bk = static_cast<unsigned*>(oph_scratch(ctx, SLOT_AUX1, sizeof(unsigned) * book));
    }
    if (!a.pts || !bk) return OPH_ENOMEM;
    a.chunk_counts = bk;
    a.totals = bk + (size_t)nchunks * ntiles;
    a.offsets = a.totals + ntiles;
    a.overflow = a.offsets + ntiles + 1;
    a.chunk_max = reinterpret_cast<float*>(a.overflow + 1);
    a.gmax = a.chunk_max + nchunks;
    a.sticky = check ? nullptr : ctx->wrp_flag;
    size_t cap = ws.fixed ? ws.list_cap : ctx->slot_size[SLOT_TABLE2] / sizeof(unsigned);
    const size_t want = wrp_list_entries(np);
    if (!ws.fixed && cap < want) cap = want;
    for (int attempt = 0; attempt < 2; ++attempt) {
        a.list = ws.fixed ? ws.list : static_cast<unsigned*>(oph_scratch(ctx, SLOT_TABLE2, cap * sizeof(unsigned)));
        if (!a.list) return OPH_ENOMEM;
        a.capacity = (unsigned)(cap < 0xFFFFFFFFull ? cap : 0xFFFFFFFFull);
        OPH_LAUNCH(ctx, "wrp_bin_count", wrp_bin_count, dim3(nchunks), dim3(BIN_THREADS), 0, a);
        OPH_LAUNCH(ctx, "wrp_bin_colscan", wrp_bin_colscan, dim3((ntiles + 3) / 4), dim3(256), 0, a);
        static bool fill_attr = false;
        if (!fill_attr) {
            hipFuncSetAttribute(reinterpret_cast<const void*>(&wrp_bin_fill),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)(2 * sizeof(unsigned) * MAX_TILES));
            hipGetLastError();
            fill_attr = true;
        }
        OPH_LAUNCH(ctx, "wrp_bin_fill", wrp_bin_fill, dim3(nchunks), dim3(BIN_THREADS),
                   2 * sizeof(unsigned) * (size_t)ntiles, a);
        GatherArgs g{};
        g.pts = a.pts;
        g.offsets = a.offsets;
        g.list = a.list;
        g.chunk_pre = a.chunk_counts;
        g.overflow = a.overflow;
        g.nchunks = nchunks;
        g.gmax = a.gmax;
        g.capacity = a.capacity;
        g.nx = nx;
        g.ny = ny;
        g.ntx = a.ntx;
        g.pitch = pitch;
        g.inv_lambda = a.inv_lambda;
        g.r_off = mode == 3 ? 0.f : 0.05f;
        g.lo_add = mode == 3 ? 1 : 0;
        g.wrp = wrp;
        g.accumulate = accumulate;
        if (mode == 2)
            OPH_LAUNCH(ctx, "wrp_gather", wrp_gather<true>, dim3(ntiles), dim3(GATHER_THREADS), 0, g);
        else
            OPH_LAUNCH(ctx, "wrp_gather", wrp_gather<false>, dim3(ntiles), dim3(GATHER_THREADS), 0, g);
        // the list size is only known on the device; a capacity overflow is
        // detected after the fact (the gather then writes nothing, so an
        // accumulating call is still intact) and the pass is redone with the
        // exact size.  Asynchronous calls skip the read-back (the capacity is
        // 24 entries per point, far above what grating-angle windows need):
        // their overflow sets the context's flag, which oph_sync() reports.
        if (!check) break;
        unsigned hdr[2] = {0, 0};
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(hdr, a.offsets + ntiles, 2 * sizeof(unsigned), hipMemcpyDeviceToHost,
                                          ctx->stream));
        OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (!hdr[1]) break;
        if (ws.fixed)
            return oph_fail(ctx, OPH_ENOMEM, "WRP tile lists need %u entries, above the static capacity %zu; "
                                             "the plane was left unwritten", hdr[0], cap);
        if (attempt == 1) return oph_fail(ctx, OPH_ENOMEM, "WRP tile lists overflowed at the exact size");
        cap = (size_t)hdr[0] + 65536;
    }
    return OPH_OK;
}

}  // namespace

int oph_wrp_device(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, double z_wrp, int mode,
                   const double* x, const double* y, const double* z, const double* amp, const double* phase,
                   int64_t n, float2* wrp, int accumulate, int check) {
    return wrp_core(ctx, nx, ny, pitch, lambda, z_wrp, mode, x, y, z, amp, phase, n, wrp, accumulate, check,
                    WrpScratch{false, nullptr, nullptr, nullptr, 0});
}

extern "C" int oph_wrp(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, double z_wrp, int mode,
                       const oph_points* pts, oph_c64* wrp, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!pts || !wrp) return oph_fail(ctx, OPH_EINVAL, "null argument");
    if (nx <= 0 || ny <= 0 || nx > 65536 || ny > 65536) return oph_fail(ctx, OPH_EINVAL, "bad WRP size");
    if (!(pitch > 0) || !(lambda > 0) || lambda >= 2 * pitch)
        return oph_fail(ctx, OPH_EINVAL, "pitch and wavelength must be positive with lambda < 2 pitch");
    if (mode < 0 || mode > 3) return oph_fail(ctx, OPH_EINVAL, "mode must be 0, 1, 2 or 3");
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    if (pts->n < 0 || (pts->n > 0 && (!pts->x || !pts->y || !pts->z)))
        return oph_fail(ctx, OPH_EINVAL, "bad point arrays");
    const int64_t n = pts->n;
    const size_t pb = (size_t)(n > 0 ? n : 1) * sizeof(double);
    const int sf = n > 0 ? flags : (flags | OPH_DEVICE_PTR);
    // the phase is read by mode 3 only
    const double* hph = mode == 3 ? pts->phase : nullptr;
    if (!(flags & OPH_DEVICE_PTR) && !oph_scratch(ctx, SLOT_PTS, 5 * pb)) return OPH_ENOMEM;
    const double* dx = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->x, pb, sf, 0));
    const double* dy = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->y, pb, sf, pb));
    const double* dz = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->z, pb, sf, 2 * pb));
    const double* da = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->amp, pb, sf, 3 * pb));
    const double* dp = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, hph, pb, sf, 4 * pb));
    if (n > 0 && (!dx || !dy || !dz || (pts->amp && !da) || (hph && !dp))) return OPH_ENOMEM;
    const size_t ob = (size_t)nx * ny * sizeof(float2);
    float2* out = reinterpret_cast<float2*>(wrp);
    if (!(flags & OPH_DEVICE_PTR)) {
        out = static_cast<float2*>(oph_scratch(ctx, SLOT_OUT, ob));
        if (!out) return OPH_ENOMEM;
        if (flags & OPH_ACCUMULATE)
            OPH_HIP_CHECK(ctx, hipMemcpyAsync(out, wrp, ob, hipMemcpyHostToDevice, ctx->stream));
    }
    if ((rc = oph_wrp_device(ctx, nx, ny, pitch, lambda, z_wrp, mode, dx, dy, dz, da, dp, n, out,
                             (flags & OPH_ACCUMULATE) ? 1 : 0, !(flags & OPH_ASYNC))))
        return rc;
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(wrp, out, ob, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}

extern "C" int oph_wrp_hologram(oph_ctx* ctx, int n_side, double pitch, double lambda, double z_wrp,
                                double distance, int pad, int band_limit, int mode, const oph_points* pts,
                                oph_c64* out, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (pad != 1 && pad != 2) return oph_fail(ctx, OPH_EINVAL, "pad must be 1 or 2");
    if (mode < 0 || mode > 3) return oph_fail(ctx, OPH_EINVAL, "mode must be 0, 1, 2 or 3");
    const int np = n_side * pad;
    if (n_side <= 0 || (np & (np - 1)) || np < 16 || np > 8192)
        return oph_fail(ctx, OPH_ENOTSUP, "padded grid %d must be a power of two in 16..8192", np);
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    // WRP into the output buffer (device or staged), then ASM in place
    if (!(flags & OPH_DEVICE_PTR)) {
        // host buffers: run the device core on staged copies
        const int64_t n = pts->n;
        const size_t pb = (size_t)(n > 0 ? n : 1) * sizeof(double);
        const int sf = n > 0 ? flags : (flags | OPH_DEVICE_PTR);
        const double* hph = mode == 3 ? pts->phase : nullptr;
        if (!oph_scratch(ctx, SLOT_PTS, 5 * pb)) return OPH_ENOMEM;
        const double* dx = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->x, pb, sf, 0));
        const double* dy = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->y, pb, sf, pb));
        const double* dz = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->z, pb, sf, 2 * pb));
        const double* da = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->amp, pb, sf, 3 * pb));
        const double* dp = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, hph, pb, sf, 4 * pb));
        if (n > 0 && (!dx || !dy || !dz || (pts->amp && !da) || (hph && !dp))) return OPH_ENOMEM;
        const size_t ob = (size_t)n_side * n_side * sizeof(float2);
        float2* d = static_cast<float2*>(oph_scratch(ctx, SLOT_OUT, ob));
        float2* work = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK1, (size_t)np * np * sizeof(float2)));
        if (!d || !work) return OPH_ENOMEM;
        if ((rc = oph_wrp_device(ctx, n_side, n_side, pitch, lambda, z_wrp, mode, dx, dy, dz, da, dp, n, d, 0, 1)))
            return rc;
        if ((rc = oph_asm_device(ctx, d, n_side, n_side, np, np, pitch, pitch, lambda, distance, band_limit,
                                 work)))
            return rc;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(out, d, ob, hipMemcpyDeviceToHost, ctx->stream));
        return oph_finish(ctx, flags);
    }
    float2* d = reinterpret_cast<float2*>(out);
    float2* work = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK1, (size_t)np * np * sizeof(float2)));
    if (!work) return OPH_ENOMEM;
    if ((rc = oph_wrp_device(ctx, n_side, n_side, pitch, lambda, z_wrp, mode, pts->x, pts->y, pts->z, pts->amp,
                             mode == 3 ? pts->phase : nullptr, pts->n, d, 0, !(flags & OPH_ASYNC))))
        return rc;
    if ((rc = oph_asm_device(ctx, d, n_side, n_side, np, np, pitch, pitch, lambda, distance, band_limit, work)))
        return rc;
    return oph_finish(ctx, flags);
}

// ---------------------------------------------------------------------------
// Allocation-free WRP (oph_wrp_noalloc): the binning state lives in one
// file-scope device array, `scratch`, sized for at most NOALLOC_MAX_POINTS
// points on planes of at most NOALLOC_MAX_TILES 64x64 tiles (2048 x 2048),
// with a fixed tile-list capacity.  Each device holds its own copy for the
// life of the process.  In host code `scratch` names no device address: its
// address on a device comes from hipGetSymbolAddress, taken once per device
// (with that device current) when a context is created there, and only that
// address is passed to kernels.  Every call on a device shares its copy and
// overwrites everything it reads (records, counts, offsets, flags and lists
// are all written by the binning passes before use), so nothing needs
// clearing; calls that overlap on one device (two contexts' streams at once)
// race on it.
namespace {

constexpr int64_t NOALLOC_MAX_POINTS = 131072;
constexpr int NOALLOC_MAX_TILES = 1024;
constexpr size_t NOALLOC_LIST = wrp_list_entries(NOALLOC_MAX_POINTS);
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NOALLOC_PTS_BYTES = align256(sizeof(WrpPoint) * NOALLOC_MAX_POINTS);
constexpr size_t NOALLOC_BOOK_BYTES = align256(
    sizeof(unsigned) * wrp_book_words((NOALLOC_MAX_POINTS + wrp_chunk(NOALLOC_MAX_POINTS) - 1) /
                                          wrp_chunk(NOALLOC_MAX_POINTS),
                                      NOALLOC_MAX_TILES));
constexpr size_t NOALLOC_LIST_BYTES = align256(sizeof(unsigned) * NOALLOC_LIST);
constexpr size_t NOALLOC_BYTES = NOALLOC_PTS_BYTES + NOALLOC_BOOK_BYTES + NOALLOC_LIST_BYTES;
static_assert(NOALLOC_BYTES < (size_t)24 << 20, "the static WRP scratch stays below 24 MiB");

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex noalloc_mutex;
void* noalloc_ptr[MAX_DEVICES] = {};

}  // namespace

// Called by oph_create with `device` current: records the device address of
// `scratch` on that device (once per device).
int oph_wrp_noalloc_setup(oph_ctx* ctx, int device) {
    if (device < 0 || device >= MAX_DEVICES)
        return oph_fail(ctx, OPH_ENOTSUP, "device %d: the allocation-free WRP supports devices 0..%d", device,
                        MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (noalloc_ptr[device]) return OPH_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        hipGetLastError();
        return oph_fail(ctx, OPH_EDEVICE, "hipGetSymbolAddress of the static WRP scratch failed: %s",
                        hipGetErrorString(e));
    }
    noalloc_ptr[device] = p;
    return OPH_OK;
}

extern "C" int oph_wrp_noalloc(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, double z_wrp, int mode,
                               const oph_points* pts, oph_c64* wrp, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!pts || !wrp) return oph_fail(ctx, OPH_EINVAL, "null argument");
    if (!(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "oph_wrp_noalloc takes device pointers only (OPH_DEVICE_PTR)");
    if (nx <= 0 || ny <= 0) return oph_fail(ctx, OPH_EINVAL, "bad WRP size");
    const int64_t ntiles = (int64_t)((nx + TILE - 1) / TILE) * ((ny + TILE - 1) / TILE);
    if (ntiles > NOALLOC_MAX_TILES)
        return oph_fail(ctx, OPH_EINVAL, "WRP plane %dx%d has %lld tiles of 64x64; the allocation-free entry point "
                                         "takes at most %d (2048 x 2048)",
                        nx, ny, (long long)ntiles, NOALLOC_MAX_TILES);
    if (pts->n < 0 || pts->n > NOALLOC_MAX_POINTS)
        return oph_fail(ctx, OPH_EINVAL, "%lld points; the allocation-free entry point takes at most %lld",
                        (long long)pts->n, (long long)NOALLOC_MAX_POINTS);
    if (!(pitch > 0) || !(lambda > 0) || lambda >= 2 * pitch)
        return oph_fail(ctx, OPH_EINVAL, "pitch and wavelength must be positive with lambda < 2 pitch");
    if (mode < 0 || mode > 3) return oph_fail(ctx, OPH_EINVAL, "mode must be 0, 1, 2 or 3");
    if (pts->n > 0 && (!pts->x || !pts->y || !pts->z)) return oph_fail(ctx, OPH_EINVAL, "bad point arrays");
    void* base = ctx->device < MAX_DEVICES ? noalloc_ptr[ctx->device] : nullptr;
    if (!base) return oph_fail(ctx, OPH_EDEVICE, "the static WRP scratch has no address on device %d", ctx->device);
    char* b = static_cast<char*>(base);
    const WrpScratch ws{true, reinterpret_cast<WrpPoint*>(b), reinterpret_cast<unsigned*>(b + NOALLOC_PTS_BYTES),
                        reinterpret_cast<unsigned*>(b + NOALLOC_PTS_BYTES + NOALLOC_BOOK_BYTES), NOALLOC_LIST};
    if ((rc = wrp_core(ctx, nx, ny, pitch, lambda, z_wrp, mode, pts->x, pts->y, pts->z, pts->amp,
                       mode == 3 ? pts->phase : nullptr, pts->n,
                       reinterpret_cast<float2*>(wrp), (flags & OPH_ACCUMULATE) ? 1 : 0, !(flags & OPH_ASYNC), ws)))
        return rc;
    return oph_finish(ctx, flags);
}
