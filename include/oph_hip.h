/*
 * oph_hip.h -- C ABI of the MI355X (gfx950) hologram-generation library
 * liboph_hip.so.
 *
 * Plain C types only: pointers, sizes, POD structs.  Every entry point returns
 * OPH_OK (0) or a negative OPH_E* code and never exits the process; the text
 * of the last error of a context is returned by oph_last_error().
 *
 * Memory: array arguments are host pointers unless the call's flags contain
 * OPH_DEVICE_PTR, in which case every array argument of that call is a device
 * pointer (from oph_device_alloc or any HIP allocation on the context's
 * device).  The context owns its device scratch and one HIP stream; calls are
 * synchronous unless OPH_ASYNC is given (device-pointer calls only), after
 * which oph_sync() waits.  A context is used by one host thread at a time;
 * several contexts may run concurrently.  Results are deterministic for a
 * fixed input (no floating-point atomics on any path).
 *
 * Each generator cites the reference interface it replaces (paths relative to
 * the reference tree: K/ = OpenHolo_KETI/VS/Openholo_Generation_PointCloud_
 * Library_beta_v1.0, DM/ = OpenHolo_DepthMap, W/ = OpenHolo_WRP/openwrp_
 * console_ver 1.0, CWO/ = W/cwo/src).
 */
#ifndef OPH_HIP_H
#define OPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define OPH_OK        0
#define OPH_EINVAL   -1  /* bad argument, size or spec */
#define OPH_ENOMEM   -2  /* device or host allocation failed */
#define OPH_EDEVICE  -3  /* HIP runtime error */
#define OPH_EIO      -4  /* file missing or malformed */
#define OPH_ENUMERIC -5  /* non-finite input or output */
#define OPH_ENOTSUP  -6  /* combination not supported */

/* ---- flags ---- */
#define OPH_DEVICE_PTR  0x0001 /* all array arguments are device pointers */
#define OPH_ACCUMULATE  0x0002 /* add into the output instead of overwriting */
#define OPH_ASYNC       0x0004 /* do not wait for the stream (device pointers) */
/* output element format of the point-cloud generators (one of) */
#define OPH_OUT_C128    0x0000 /* complex128, {re, im} doubles */
#define OPH_OUT_C64     0x0010 /* complex64, {re, im} floats */
#define OPH_OUT_FRINGE  0x0020 /* float Re(U): the KETI real fringe */
#define OPH_OUT_MASK    0x0030

typedef struct oph_ctx oph_ctx;
typedef struct { double re, im; } oph_c128;
typedef struct { float re, im; } oph_c64;

/* ---- context ---- */
int oph_create(int device, oph_ctx** out);
void oph_destroy(oph_ctx* ctx);
const char* oph_last_error(const oph_ctx* ctx);
int oph_sync(oph_ctx* ctx);
/* Version string of the library ("oph_hip <major>.<minor> gfx950"). */
const char* oph_version(void);

/* ---- device memory and timing helpers (for callers that keep data
 * resident in HBM) ---- */
int oph_device_alloc(oph_ctx* ctx, size_t bytes, void** dptr);
int oph_device_free(oph_ctx* ctx, void* dptr);
int oph_copy_to_device(oph_ctx* ctx, void* dst, const void* src, size_t bytes);
int oph_copy_to_host(oph_ctx* ctx, void* dst, const void* src, size_t bytes);
int oph_device_zero(oph_ctx* ctx, void* dptr, size_t bytes);
/* Device-to-device copy on the context stream (synchronous). */
int oph_device_copy(oph_ctx* ctx, void* dst, const void* src, size_t bytes);
/* HIP events recorded on the context's stream around whatever the caller
 * launches between begin and end; end waits and returns milliseconds. */
int oph_timer_begin(oph_ctx* ctx);
int oph_timer_end(oph_ctx* ctx, float* ms);

/* Per-kernel timing: when enabled, every kernel the library launches on the
 * context's stream is bracketed by a HIP event pair (enable clears previous
 * records).  oph_profile_get sums the recorded durations of one kernel name
 * (e.g. "rs_accumulate", "fft_rows") in milliseconds. */
int oph_profile_enable(oph_ctx* ctx, int on);
int oph_profile_get(oph_ctx* ctx, const char* kernel, int* count, double* total_ms);

/* ---- hologram spec (replaces KETI OphSpec, K/ophgen_pointcloud/
 * ophgen_pointcloud.h:45-70, in double: a float wavelength or depth alone
 * moves k r by ~0.06-0.09 rad at 0.1 m) ---- */
typedef struct oph_spec {
    int nx, ny;            /* SLM pixels */
    double pitch_x, pitch_y;
    double lambda;         /* metres */
    double scale[3];       /* point-cloud scale factors x, y, z */
    double offset_z;       /* added to the scaled z */
    double tilt_deg[2];    /* reference-wave tilt, degrees */
} oph_spec;

/* Point cloud, structure of arrays, n points.  amp and phase may be NULL
 * (amplitude 1, phase 0).  phase is in radians. */
typedef struct oph_points {
    const double* x;
    const double* y;
    const double* z;
    const double* amp;
    const double* phase;
    int64_t n;
} oph_points;

/* Parses a 17-line KETI spec file (`key = value`, keys ignored, exactly 17
 * lines) -- replaces ophLoadSpecConfig, K/ophgen_pointcloud/
 * ophgen_pointcloud.cpp:112-158.  Wavelength, pitches, scales and offset are
 * parsed directly to double.  OPH_EIO if the file is missing or has != 17
 * lines. */
int oph_spec_from_keti_config(const char* path, oph_spec* spec);

/* The 17 value strings of the same file in file order (the reference keeps
 * field 8, the filter shape, as text), for callers that map them onto the
 * reference's float OphSpec themselves. */
int oph_keti_config_fields(const char* path, char values[17][64]);

/* Loads a KETI .dat point cloud (first line P, then `idx x y z phase amp`
 * per line, idx == line index) -- replaces ophLoadPointCloud,
 * K/ophgen_pointcloud/ophgen_pointcloud.cpp:13-109.  With xyzpa == NULL only
 * *n is set; otherwise up to cap points are written as 5 doubles each
 * (x, y, z, phase, amplitude: the file's column order).  OPH_EIO on a missing
 * file or an index mismatch. */
int oph_load_keti_points(const char* path, double* xyzpa, int64_t cap, int64_t* n);

/* ---- point cloud: exact Rayleigh-Sommerfeld field ----
 * Replaces ophGenCghPointCloud (K/ophgen_pointcloud/ophgen_pointcloud.cpp:
 * 195-365) and gpu::ophGenCghPointCloud_cuda (K/ophgen_pointcloud_gpu/
 * ophgen_pointcloud_gpu.cpp:391-503).  For every pixel (row, col):
 *   X = (col+1/2) px - nx px/2,   Y = ny py/2 - (row+1/2) py
 *   U = sum_n a_n exp(i (k r_n + phi_n)) exp(-i k (X sin tx + Y sin ty))
 *   r_n = sqrt((X - sx x_n)^2 + (Y - sy y_n)^2 + (sz z_n + offset_z)^2)
 * out is ny*nx elements, row-major, in the format given by OPH_OUT_*;
 * OPH_OUT_FRINGE writes Re(U) as float, the reference's fringe (which ignores
 * the point phase: pass phase = NULL for parity).  OPH_ACCUMULATE adds into
 * out as the reference's `dst +=`. */
int oph_pc_rs(oph_ctx* ctx, const oph_spec* spec, const oph_points* pts,
              void* out, int flags);

/* ---- 2-D FFT, complex64, in place ----
 * Replaces the FFTW / cuFFT Z2Z transform of the depth-map generator's
 * fftwShift / cudaFFT (DM/Hologram/src/HologramGenerator_CPU.cpp:285-322,
 * DM/Hologram/src/HologramKernel.cu:189-225) without the shifts.
 * data is ny rows of nx; nx, ny powers of two in 16..8192.
 * direction -1: X[k] = sum_n x[n] exp(-2 pi i k n / N) (forward),
 * direction +1: the unnormalised inverse; normalize != 0 divides by nx*ny. */
int oph_fft2d(oph_ctx* ctx, oph_c64* data, int nx, int ny, int direction,
              int normalize, int flags);

/* ---- angular-spectrum propagation, complex64, in place ----
 * Replaces CWO::Diffract with CWO_ANGULAR (CWO/cwo.cpp:2198-2278,
 * 2907-2936; transfer CWO/cwo_lib.h:66-101): the ny x nx field is zero-padded
 * to (pad*ny) x (pad*nx) at offset (pad*nx/2 - nx/2, pad*ny/2 - ny/2),
 * transformed, multiplied by H = exp(i 2 pi z sqrt(1/lambda^2 - fx^2 - fy^2))
 * (0 for evanescent frequencies; fx = kx/(pad nx px) in FFT order),
 * inverse-transformed with 1/(Npx Npy), and cropped back.  band_limit != 0
 * applies the Matsushima band limit of OpenHolo_LFCGH/
 * fn_FresnelPropagation_as.m:33-38.  pad is 1 or 2; pad*nx and pad*ny must
 * be powers of two in 16..8192. */
int oph_asm_propagate(oph_ctx* ctx, oph_c64* field, int nx, int ny, double px,
                      double py, double lambda, double z, int pad, int band_limit,
                      int flags);

/* ---- wavefront-recording plane ----
 * Replaces OPHWRP::SingleWRP (W/openwrp/WRP.cpp:75-142 with CWO::AddPixel,
 * CWO/cwo.cpp:4298-4309).  Coordinates are along the propagation axis; for
 * each point, dz = z_wrp - z_n, centre tx = (int)(x/p) + nx/2,
 * ty = (int)(y/p) + ny/2, and over the window wx, wy in [-w, w):
 *   wrp[ty+wy][tx+wx] += a exp(i k r) / (r + 0.05),
 *   r = sign(dz) sqrt((wx p)^2 + (wy p)^2 + dz^2)
 * (pixels outside the plane skipped).  mode 0: grating-angle half-width
 * w = (int)(|dz| tan(asin(lambda/(2p)))/p) (CWO/cwoPLS.cpp:613); mode 1:
 * the reference's w = (int)|lambda dz/(2 p^2) + 0.5| * 2 - 1 (WRP.cpp:111);
 * mode 2: CWO's cwoPLS::WRP (CWO/cwoPLS.cpp:522-591), Fresnel phase
 *   wrp[..] += a exp(i k (dz + ((wx p)^2 + (wy p)^2)/(2 dz))),
 * half-width w = (int)(|dz| lambda/(2p)/p) (:557; the reference takes |dz|
 * from its call argument, equal to dz for points at z = 0); mode 3: the
 * MATLAB statement (OpenHolo_WRP/WRP_matlab_ver 1.0/Hologram_WRP.m:19-36,
 * fn_wrp.m): full width N = round(|lambda dz/p^2/2| + 0.5) 2 - 1, square
 * window wx, wy in [-(N-1)/2, (N-1)/2], centre round(x/p) + (nx-1)/2 - 1
 * (likewise y; the script's 1-based round(x/p) + (res-1)/2), and
 *   wrp[..] += a exp(i phase) exp(i k r) / r
 * with pts->phase the script's per-point random phase (radians; NULL: 0);
 * points with dz = 0 add nothing.  The script's matrix holds x along rows:
 * it is the transpose of wrp.  Modes 0-2 ignore pts->phase (the reference
 * points carry none).  wrp is ny x nx complex64, row index = y.
 * OPH_ACCUMULATE adds into wrp. */
int oph_wrp(oph_ctx* ctx, int nx, int ny, double pitch, double lambda,
            double z_wrp, int mode, const oph_points* pts, oph_c64* wrp, int flags);

/* ---- allocation-free WRP ----
 * No counterpart in the reference: oph_wrp for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  Same arguments, kernels and result as oph_wrp
 * (bit-identical output), but every array is a device pointer (flags must
 * contain OPH_DEVICE_PTR; OPH_ACCUMULATE and OPH_ASYNC as in oph_wrp) and
 * the binning state lives in a static device array of the library (under
 * 24 MiB per device, held until the process ends) instead of the context's
 * scratch, so the call allocates nothing.  Limits: at most 131072 points,
 * planes of at most 1024 tiles of 64x64 (e.g. 2048 x 2048), and at most
 * 3,211,264 (point, tile) list entries (24 per point plus 65536); larger
 * calls return OPH_EINVAL, and a list overflow returns OPH_ENOMEM (reported
 * by oph_sync for OPH_ASYNC) with the plane left unwritten.  All calls on
 * one device share the array: calls that overlap on a device (from two
 * contexts at once) race. */
int oph_wrp_noalloc(oph_ctx* ctx, int nx, int ny, double pitch, double lambda,
                    double z_wrp, int mode, const oph_points* pts, oph_c64* wrp, int flags);

/* ---- WRP hologram: oph_wrp on an n_side^2 plane, then oph_asm_propagate
 * over `distance` with zero padding `pad` (replaces SingleWRP followed by
 * OPHWRP::Fresnelpropagation, W/openwrp/WRP.cpp:65-69, whose CWO Fresnel
 * convolution kernels are closed; the angular-spectrum transfer is used).
 * out: n_side^2 complex64 hologram. */
int oph_wrp_hologram(oph_ctx* ctx, int n_side, double pitch, double lambda,
                     double z_wrp, double distance, int pad, int band_limit,
                     int mode, const oph_points* pts, oph_c64* out, int flags);

/* ---- Huygens (Rayleigh-Sommerfeld) point sum on CWO's grid ----
 * cwoPLS::Huygens (CWO/cwoPLS.cpp:316-368) with the point amplitude and
 * initial phase:
 *   U[i][j] = sum_n a_n exp(i (k r + ph + phi_n)) / r,
 *   r = sqrt((X_j - x_n)^2 + (Y_i - y_n)^2 + (z_n + z)^2),
 *   X_j = (j - nx/2) pitch, Y_i = (i - ny/2) pitch (row index i is y).
 * The reference computes r in double and sums in float; here the phase is
 * reduced in FP64 and summed as oph_pc_rs does.  flags: OPH_OUT_C128 or
 * OPH_OUT_C64, OPH_DEVICE_PTR, OPH_ACCUMULATE, OPH_ASYNC. */
int oph_pc_huygens(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, double z,
                   double ph, const oph_points* pts, void* out, int flags);

/* ---- Fresnel point-cloud hologram ----
 * The direct Fresnel sum of cwoPLS::Fresnel (CWO/cwoPLS.cpp:370-422) with the
 * point amplitude, initial phase and plane carrier:
 *   U[i][j] = sum_n a_n exp(i (k z_n + phi_n)) exp(i k ((X_j - x_n)^2 + (Y_i - y_n)^2)/(2 z_n)),
 *   X_j = (j - nx/2) pitch, Y_i = (i - ny/2) pitch (row index i is y).
 * method OPH_FRESNEL_VALU sums it directly (FP64 phase reduction, FP32
 * accumulation); the MFMA methods evaluate the same field as the separable
 * complex contraction (pixels x points)(points x pixels) with chirp factors
 * rounded to BF16 or FP16 (optionally split hi + lo) and FP32 accumulation.
 * When the points lie on few depth planes (at most 4096 distinct z, >= 256
 * points per plane on average) the contraction runs plane by plane with the
 * plane's chirp exp(i k (X^2 + Y^2)/(2 z_p)) applied once per plane (same
 * field; the environment variable OPH_FRESNEL_PLANES=0 disables this form).
 * out: ny x nx complex64.  z_n > 0 (distance of the point's plane). */
#define OPH_FRESNEL_VALU             0
#define OPH_FRESNEL_MFMA_BF16        1
#define OPH_FRESNEL_MFMA_BF16_SPLIT  2
#define OPH_FRESNEL_MFMA_FP16        3
#define OPH_FRESNEL_MFMA_FP16_SPLIT  4
int oph_pc_fresnel(oph_ctx* ctx, int nx, int ny, double pitch, double lambda,
                   const oph_points* pts, int method, oph_c64* out, int flags);

/* ---- depth-map layer hologram ----
 * Replaces HologramGenerator::GenerateHologram's compute (DM/Hologram/src/
 * HologramGenerator.cpp:261-294 with the CPU path HologramGenerator_CPU.cpp:
 * 44-270 and the GPU path HologramGenerator_GPU.cpp:194-368): depth
 * quantisation (FLAG_CHANGE_DEPTH_QUANTIZATION 0 or 1), the viewing-window
 * transform z_l = -f d_l/(d_l - f) (or z_l = d_l without a field lens), per
 * layer l = layer_lo..layer_hi:
 *   u_l = img/255 [img > 0] [depth index == l] rho exp(i k z_l)
 *   U  += H_l . fftshift(FFT2(fftshift(u_l))),
 *   H_l = exp(i k (-z_l) sqrt(1 - (lambda fx)^2 - (lambda fy)^2)) on the
 *         reference's frequency grids (HologramGenerator_CPU.cpp:196-216),
 * then Encoding_Symmetrization(0, 1) (HologramGenerator.cpp:481-522,
 * HologramGenerator_CPU.cpp:231-270) to the real fringe.  The scene (w x h,
 * u8 intensity per colour and one u8 depth map) sits centred in an Nx x Ny
 * transform grid (powers of two >= w, h; Nx = w, Ny = h is the reference's
 * unpadded case).  Depth levels follow the reference's repeated addition
 * (HologramGenerator.cpp:393-401), continued to exactly num_depth entries.
 * Evanescent frequencies (absent at pitches above lambda/2) give 0 where the
 * reference would produce NaN. */
typedef struct oph_dm_spec {
    int w, h;              /* scene size (pixels) */
    int Nx, Ny;            /* transform grid */
    double pitch_x, pitch_y;
    double near_m, far_m;  /* depth range */
    int num_depth;         /* number of depth levels L (<= 4096) */
    int quant_mode;        /* 0: index = default_q - d;  1: requantise to L levels */
    int default_q;         /* DEFAULT_DEPTH_QUANTIZATION (256) */
    double field_lens;     /* focal length of the field lens; <= 0: none */
    double rho_re, rho_im; /* the random-phase factor (1, 0 when RANDOM_PHASE = 0) */
    int layer_lo, layer_hi;/* rendered layers, 1-based inclusive; 0 = all */
} oph_dm_spec;

/* output of oph_depthmap (one of; ORed into flags) */
#define OPH_DM_OUT_FRINGE   0x0000 /* double fringe (u255_fringe_), ncolor x h x w */
#define OPH_DM_OUT_FRINGE32 0x0100 /* float fringe */
#define OPH_DM_OUT_FIELD    0x0200 /* complex64 hologram-plane field before encoding */
#define OPH_DM_OUT_MASK     0x0300

/* intensity: ncolor x h x w u8 (contiguous), depth: h x w u8, lambdas: ncolor
 * wavelengths; out per OPH_DM_OUT_*, ncolor x h x w. */
int oph_depthmap(oph_ctx* ctx, const oph_dm_spec* spec, int ncolor, const double* lambdas,
                 const unsigned char* intensity, const unsigned char* depth, void* out,
                 int flags);

/* ---- learned correction (no counterpart in the reference) ----
 * A residual CNN that maps a coarse (few-layer) depth-map hologram field to
 * the full-layer one: h0 = conv(x) (2 -> 32 channels), h_b = h_{b-1} +
 * conv(relu(conv(h_{b-1}))) for b = 1..blocks, y = x + conv(h_blocks)
 * (32 -> 2); 3x3 convolutions with zero padding and bias.  Tensors are NHWC
 * float32 (B x H x W x 2: real and imaginary parts); the convolutions run on
 * MFMA with BF16-rounded operands and FP32 accumulation.  Parameters are one
 * float32 vector: head W[32][9][2], b[32]; per block W1[32][9][32], b1[32],
 * W2[32][9][32], b2[32]; tail W[2][9][32], b[2]. */
typedef struct oph_cnn oph_cnn;
int oph_cnn_create(oph_ctx* ctx, int blocks, int channels, uint64_t seed, oph_cnn** out);
void oph_cnn_destroy(oph_ctx* ctx, oph_cnn* net);
int64_t oph_cnn_num_params(const oph_cnn* net);
/* which = 0: parameters, 1: gradients of the last oph_cnn_gradients/train_step */
int oph_cnn_get_params(oph_ctx* ctx, oph_cnn* net, float* host, int which);
int oph_cnn_set_params(oph_ctx* ctx, oph_cnn* net, const float* host);
/* flags: OPH_DEVICE_PTR, OPH_ASYNC, and OPH_CNN_BF16_ACTS: keep the
 * activations between layers in BF16 (inference; the convolution operands
 * are BF16-rounded in both forms, this also rounds the stored intermediate
 * and residual tensors, halving their HBM traffic; each residual block is one
 * kernel, for images below 2^25 pixels, OPH_ENOTSUP above).  Training always
 * uses float32 activations. */
#define OPH_CNN_BF16_ACTS 0x1000
int oph_cnn_forward(oph_ctx* ctx, oph_cnn* net, const float* x, float* y, int batch, int h, int w,
                    int flags);
/* loss = mean over all elements of (y - target)^2; gradients only, or
 * gradients then one Adam step (beta 0.9 / 0.999, eps 1e-8, rate lr) */
int oph_cnn_gradients(oph_ctx* ctx, oph_cnn* net, const float* x, const float* target, int batch,
                      int h, int w, double* loss, int flags);
int oph_cnn_train_step(oph_ctx* ctx, oph_cnn* net, const float* x, const float* target, int batch,
                       int h, int w, float lr, double* loss, int flags);
/* binary file: "OPHCNN1\0", int32 blocks, int32 channels, int64 n, n float32 */
int oph_cnn_save(oph_ctx* ctx, oph_cnn* net, const char* path);
int oph_cnn_load(oph_ctx* ctx, oph_cnn* net, const char* path);

/* ---- encoding and image files ----
 * 8-bit normalisation of a float fringe (nx x ny, row-major): min and max
 * over the array, vertical flip, dst[(ny-1-y) nx + x] =
 * (unsigned char)(((v - min)/(max - min)) * 255 + 0.5) with the reference's
 * float/double operation order -- replaces ophNormalize,
 * K/oph_common/oph_common.cpp:92-116 (bit-exact).  A constant input (max ==
 * min, undefined in the reference) gives all zeros.  flags: OPH_DEVICE_PTR. */
int oph_normalize_u8(oph_ctx* ctx, const float* src, unsigned char* dst, int nx, int ny, int flags);

/* 8-bit scaling of a double array without flip: dst[i] = (unsigned char)
 * (unsigned)((v - min)/(max - min) * 255) -- the scaling of
 * HologramGenerator::Write_Result_image and Write_Simulation_image
 * (DM/Hologram/src/HologramGenerator.cpp:528-565,
 * HologramGenerator_CPU.cpp:619-658).  Constant input gives zeros.
 * flags: OPH_DEVICE_PTR. */
int oph_scale_u8(oph_ctx* ctx, const double* src, unsigned char* dst, int64_t n, int flags);

/* Encodings of a complex field (nx x ny complex64, row-major) to a float
 * array -- replaces the MATLAB toolbox's fn_encoding
 * (OpenHolo_Encoding/fn_encoding.m:1-98, with fn_myfft2.m / fn_myifft2.m):
 *   OPH_ENC_SSB         single sideband (:29-47): centred spectrum with the
 *                       1-based block floor(N/2)..N (left/top) or 1..floor(N/2)
 *                       (right/bottom) zeroed, inverse, (Re - min)/(max - min);
 *                       nx, ny powers of two in 16..8192
 *   OPH_ENC_NI          |U + max|U||^2 / its maximum (:49-54)
 *   OPH_ENC_P2A         angle(U) + 0.5, the reference's angle + pi/(2 pi) (:56-58)
 *   OPH_ENC_REAL        Re(U) (:20-21)
 *   OPH_ENC_BURCKHARDT  A1, A2, A3 interleaved, out is ny x 3nx (:60-83), with
 *                       the reference's branch terms as written
 *   OPH_ENC_TWO_PHASE   angle + pi +- acos(|U|/max|U|) interleaved, out is
 *                       ny x 2nx (:85-98)
 * passband (SSB only): OPH_PASS_LEFT, _RIGHT, _TOP, _BOTTOM.
 * flags: OPH_DEVICE_PTR, OPH_ASYNC (with device pointers). */
#define OPH_ENC_SSB        0
#define OPH_ENC_NI         1
#define OPH_ENC_P2A        2
#define OPH_ENC_REAL       3
#define OPH_ENC_BURCKHARDT 4
#define OPH_ENC_TWO_PHASE  5
#define OPH_PASS_LEFT      0
#define OPH_PASS_RIGHT     1
#define OPH_PASS_TOP       2
#define OPH_PASS_BOTTOM    3
int oph_encode_field(oph_ctx* ctx, const oph_c64* field, int nx, int ny, int method, int passband,
                     float* out, int flags);

/* 8-bit grayscale BMP with a 256-entry gray palette, byte-for-byte the file
 * ophSaveFileBmp writes (K/oph_common/oph_common.cpp:49-90): 1078-byte
 * header, rows stored as given with no row padding.  `path` is used as is
 * (the reference appends ".bmp"). */
int oph_save_bmp8(const unsigned char* pixels, int width, int height, const char* path);

/* 24-bit BMP of packed RGB888 pixels (3*width bytes per row, top row
 * first): the file Write_Result_image saves through QImage
 * (HologramGenerator.cpp:553-559): bottom-up rows, BGR order, rows padded
 * to 4 bytes. */
int oph_save_bmp24(const unsigned char* rgb, int width, int height, const char* path);

/* Reads an uncompressed 8-bit (palette), 24-bit or 32-bit BMP as 8-bit gray,
 * top row first: the Grayscale8 input of ReadImageDepth
 * (HologramGenerator.cpp:304-383; colour pixels use the gray formula
 * (11 r + 16 g + 5 b)/32).  With dst == NULL only the size is returned;
 * otherwise width*height bytes are written (OPH_EINVAL if cap is smaller). */
int oph_load_bmp_gray8(const char* path, int* width, int* height, unsigned char* dst, int64_t cap);

/* ---- point-cloud files of the WRP path ----
 * Whitespace-separated x y z triples (any line layout) -- OPHWRP::readxyz,
 * W/openwrp/WRP.cpp:7-55.  With xyz == NULL only *n is set; otherwise up to
 * cap points are written as 3 doubles each. */
int oph_load_xyz(const char* path, double* xyz, int64_t cap, int64_t* n);

/* CWO binary point files -- cwoPLS::Load, CWO/cwoPLS.cpp:48-95: int32 count,
 * then count (x, y, z) triples of int32 (".3d") or float32 (".3df"), chosen
 * by the file extension. */
int oph_load_cwo_points(const char* path, double* xyz, int64_t cap, int64_t* n);

/* ---- depth-map configuration file ----
 * The positional config_openholo.txt read by HologramGenerator::readConfig
 * (DM/Hologram/src/HologramGenerator.cpp:65-229): the same line skips and
 * fields; the pixel number and pitch fields are arithmetic expressions
 * (numbers, + - * /, parentheses, unary signs) evaluated in double;
 * RENDER_DEPTH is "a:b" or a list of level numbers.  OPH_EIO on a missing
 * file or an unparsable field. */
#define OPH_DM_MAX_RENDER 4096
typedef struct {
    char source_folder[256], image_prefix[128], depth_prefix[128];
    char result_folder[256], result_prefix[128], sim_prefix[128];
    int flag_static_image, start_of_frame_numbering, number_of_frame, number_of_digit;
    int transform_method, propagation_method, encoding_method;
    double field_lens, lambda;
    int pn[2];
    double pp[2];
    double near_depthmap, far_depthmap;   /* ordered: near <= far */
    int flag_change_depth_quantization, default_depth_quantization, number_of_depth_quantization;
    int num_of_depth;                     /* as the reference selects it */
    int n_render_depth;
    int render_depth[OPH_DM_MAX_RENDER];
    int random_phase;
    double test_pixel_number_scale, eye_length, eye_pupil_diameter, eye_center_xy[2], focus_distance;
    int sim_type, sim_step_num;
    double sim_from, sim_to;
} oph_dm_config;
int oph_dm_config_load(const char* path, oph_dm_config* cfg);

/* ---- numerical reconstruction (eye model) ----
 * Propagation of a real fringe (nx x ny, double) to the eye pupil through
 * the field lens and imaging onto the retina with the eye focused at
 * focus_distance -- HologramGenerator::ReconstructImage /
 * Test_Propagation_to_Eye_Pupil / Reconstruction
 * (DM/Hologram/src/HologramGenerator_CPU.cpp:415-614): shifted FFT, field-lens
 * chirp, eye-centre circular shift, eye-lens chirp with the anti-aliasing and
 * pupil masks, shifted FFT, magnitude.  nx, ny: powers of two (16..8192).
 * out: nx x ny doubles (the reference's sim_final_).  flags: OPH_DEVICE_PTR. */
typedef struct {
    double pitch_x, pitch_y, lambda, field_lens;
    double eye_length, eye_pupil_diameter, eye_center_x, eye_center_y, focus_distance;
} oph_eye_spec;
int oph_dm_reconstruct(oph_ctx* ctx, const oph_eye_spec* spec, const double* fringe, int nx, int ny,
                       double* out, int flags);

#ifdef __cplusplus
}
#endif
#endif /* OPH_HIP_H */
