// fft_kernels.h -- launch geometry of the FFT kernels and the row/column
// entry points other translation units reuse.
#pragma once

#include "oph_internal.h"

#define OPH_FFT_LENGTHS(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

namespace fftk {

// lines per workgroup of the row kernel: 256 threads when the line is short
__host__ __device__ constexpr int row_lines(int L) { return L / 16 >= 256 ? 1 : 256 / (L / 16); }

// columns per workgroup of the column kernel: up to 16 neighbouring columns
// (128-byte runs), at most 1024 threads and about 140 KB of LDS
__host__ __device__ constexpr int col_lines(int L) {
    int b = 16;
    while (b > 1 && (b * (L / 16) > 1024 || (size_t)b * (L + L / 16) * 8 > 140 * 1024)) b /= 2;
    return b;
}

inline bool supported(int n) { return n >= 16 && n <= 8192 && (n & (n - 1)) == 0; }

// Columns per workgroup actually used for a length-L column pass: the
// default below, or OPH_FFT_COLUMNS (a power of two) when set, capped at bmax.
int col_width(int L, int bmax);

}  // namespace fftk

// In-place batched transforms on device data (row-major, `stride` elements
// between rows), launched on the context stream; scale multiplies the output.
int oph_fft_rows(oph_ctx* ctx, int L, int sign, float2* data, int rows, int stride, float scale);
int oph_fft_cols(oph_ctx* ctx, int L, int sign, float2* data, int ncols, int stride, float scale);
