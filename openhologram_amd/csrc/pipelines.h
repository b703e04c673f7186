// pipelines.h -- device-pointer cores shared between the C-ABI entry points.
#pragma once

#include "oph_internal.h"

// ASM propagation of an ny x nx complex64 field in place on an npx x npy FFT
// grid (zero padded, centred at (npx/2 - nx/2, npy/2 - ny/2)); `work` holds
// npx*npy complex64.  Launches on the context stream.
int oph_asm_device(oph_ctx* ctx, float2* field, int nx, int ny, int npx, int npy, double px, double py,
                   double lambda, double z, int band_limit, float2* work);

// WRP accumulation into a zeroed-or-overwritten nx x ny complex64 plane.
int oph_wrp_device(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, double z_wrp, int mode,
                   const double* x, const double* y, const double* z, const double* amp, const double* phase,
                   int64_t n, float2* wrp, int accumulate, int check);

// Records the device address of the allocation-free WRP's static scratch on
// `device` (current); called once per context creation.
int oph_wrp_noalloc_setup(oph_ctx* ctx, int device);
