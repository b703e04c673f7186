"""ctypes binding of liboph_hip.so (the C ABI declared in include/oph_hip.h).

The shared library is loaded from this package directory only; if it is
missing or does not load, every call raises -- there is no CPU fallback.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboph_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "oph_hip.h")

OPH_OK, OPH_EINVAL, OPH_ENOMEM, OPH_EDEVICE, OPH_EIO, OPH_ENUMERIC, OPH_ENOTSUP = 0, -1, -2, -3, -4, -5, -6
DEVICE_PTR, ACCUMULATE, ASYNC = 0x1, 0x2, 0x4
CNN_BF16_ACTS = 0x1000
OUT_C128, OUT_C64, OUT_FRINGE = 0x00, 0x10, 0x20

_ERRNAMES = {-1: "OPH_EINVAL", -2: "OPH_ENOMEM", -3: "OPH_EDEVICE", -4: "OPH_EIO",
             -5: "OPH_ENUMERIC", -6: "OPH_ENOTSUP"}


class OphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class Spec(C.Structure):
    """oph_spec"""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("pitch_x", C.c_double), ("pitch_y", C.c_double),
                ("lambda_", C.c_double), ("scale", C.c_double * 3), ("offset_z", C.c_double),
                ("tilt_deg", C.c_double * 2)]


class DmSpec(C.Structure):
    """oph_dm_spec"""
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("Nx", C.c_int), ("Ny", C.c_int),
                ("pitch_x", C.c_double), ("pitch_y", C.c_double), ("near_m", C.c_double),
                ("far_m", C.c_double), ("num_depth", C.c_int), ("quant_mode", C.c_int),
                ("default_q", C.c_int), ("field_lens", C.c_double), ("rho_re", C.c_double),
                ("rho_im", C.c_double), ("layer_lo", C.c_int), ("layer_hi", C.c_int)]


class DmConfig(C.Structure):
    """oph_dm_config: the positional depth-map configuration file."""
    _fields_ = [("source_folder", C.c_char * 256), ("image_prefix", C.c_char * 128),
                ("depth_prefix", C.c_char * 128), ("result_folder", C.c_char * 256),
                ("result_prefix", C.c_char * 128), ("sim_prefix", C.c_char * 128),
                ("flag_static_image", C.c_int), ("start_of_frame_numbering", C.c_int),
                ("number_of_frame", C.c_int), ("number_of_digit", C.c_int),
                ("transform_method", C.c_int), ("propagation_method", C.c_int), ("encoding_method", C.c_int),
                ("field_lens", C.c_double), ("lambda_", C.c_double), ("pn", C.c_int * 2), ("pp", C.c_double * 2),
                ("near_depthmap", C.c_double), ("far_depthmap", C.c_double),
                ("flag_change_depth_quantization", C.c_int), ("default_depth_quantization", C.c_int),
                ("number_of_depth_quantization", C.c_int), ("num_of_depth", C.c_int),
                ("n_render_depth", C.c_int), ("render_depth", C.c_int * 4096), ("random_phase", C.c_int),
                ("test_pixel_number_scale", C.c_double), ("eye_length", C.c_double),
                ("eye_pupil_diameter", C.c_double), ("eye_center_xy", C.c_double * 2),
                ("focus_distance", C.c_double), ("sim_type", C.c_int), ("sim_step_num", C.c_int),
                ("sim_from", C.c_double), ("sim_to", C.c_double)]


class EyeSpec(C.Structure):
    """oph_eye_spec: eye model of the reconstruction."""
    _fields_ = [(n, C.c_double) for n in ("pitch_x", "pitch_y", "lambda_", "field_lens", "eye_length",
                                          "eye_pupil_diameter", "eye_center_x", "eye_center_y",
                                          "focus_distance")]


FRESNEL_METHODS = {"valu": 0, "bf16": 1, "bf16_split": 2, "fp16": 3, "fp16_split": 4}
DM_OUT = {"fringe": (0x0000, np.float64), "fringe32": (0x0100, np.float32), "field": (0x0200, np.complex64)}


class Points(C.Structure):
    """oph_points"""
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("amp", C.c_void_p),
                ("phase", C.c_void_p), ("n", C.c_int64)]


_lib = None


def load():
    """Loads liboph_hip.so; raises if it is absent (build with
    `python -m openhologram_amd.build`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m openhologram_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P, I, D, I64, SZ = C.c_void_p, C.c_int, C.c_double, C.c_int64, C.c_size_t
    sig = {
        "oph_create": (I, [I, C.POINTER(P)]),
        "oph_destroy": (None, [P]),
        "oph_last_error": (C.c_char_p, [P]),
        "oph_sync": (I, [P]),
        "oph_version": (C.c_char_p, []),
        "oph_device_alloc": (I, [P, SZ, C.POINTER(P)]),
        "oph_device_free": (I, [P, P]),
        "oph_copy_to_device": (I, [P, P, P, SZ]),
        "oph_copy_to_host": (I, [P, P, P, SZ]),
        "oph_device_zero": (I, [P, P, SZ]),
        "oph_device_copy": (I, [P, P, P, SZ]),
        "oph_timer_begin": (I, [P]),
        "oph_timer_end": (I, [P, C.POINTER(C.c_float)]),
        "oph_profile_enable": (I, [P, I]),
        "oph_profile_get": (I, [P, C.c_char_p, C.POINTER(I), C.POINTER(D)]),
        "oph_spec_from_keti_config": (I, [C.c_char_p, C.POINTER(Spec)]),
        "oph_load_keti_points": (I, [C.c_char_p, P, I64, C.POINTER(I64)]),
        "oph_pc_rs": (I, [P, C.POINTER(Spec), C.POINTER(Points), P, I]),
        "oph_pc_huygens": (I, [P, I, I, D, D, D, D, C.POINTER(Points), P, I]),
        "oph_fft2d": (I, [P, P, I, I, I, I, I]),
        "oph_asm_propagate": (I, [P, P, I, I, D, D, D, D, I, I, I]),
        "oph_wrp": (I, [P, I, I, D, D, D, I, C.POINTER(Points), P, I]),
        "oph_wrp_noalloc": (I, [P, I, I, D, D, D, I, C.POINTER(Points), P, I]),
        "oph_wrp_hologram": (I, [P, I, D, D, D, D, I, I, I, C.POINTER(Points), P, I]),
        "oph_depthmap": (I, [P, C.POINTER(DmSpec), I, P, P, P, P, I]),
        "oph_pc_fresnel": (I, [P, I, I, D, D, C.POINTER(Points), I, P, I]),
        "oph_cnn_create": (I, [P, I, I, C.c_uint64, C.POINTER(P)]),
        "oph_cnn_destroy": (None, [P, P]),
        "oph_cnn_num_params": (I64, [P]),
        "oph_cnn_get_params": (I, [P, P, P, I]),
        "oph_cnn_set_params": (I, [P, P, P]),
        "oph_cnn_forward": (I, [P, P, P, P, I, I, I, I]),
        "oph_cnn_gradients": (I, [P, P, P, P, I, I, I, C.POINTER(D), I]),
        "oph_cnn_train_step": (I, [P, P, P, P, I, I, I, C.c_float, C.POINTER(D), I]),
        "oph_cnn_save": (I, [P, P, C.c_char_p]),
        "oph_cnn_load": (I, [P, P, C.c_char_p]),
        "oph_normalize_u8": (I, [P, P, P, I, I, I]),
        "oph_encode_field": (I, [P, P, I, I, I, I, P, I]),
        "oph_scale_u8": (I, [P, P, P, I64, I]),
        "oph_save_bmp8": (I, [P, I, I, C.c_char_p]),
        "oph_save_bmp24": (I, [P, I, I, C.c_char_p]),
        "oph_load_bmp_gray8": (I, [C.c_char_p, C.POINTER(I), C.POINTER(I), P, I64]),
        "oph_load_xyz": (I, [C.c_char_p, P, I64, C.POINTER(I64)]),
        "oph_load_cwo_points": (I, [C.c_char_p, P, I64, C.POINTER(I64)]),
        "oph_dm_config_load": (I, [C.c_char_p, C.POINTER(DmConfig)]),
        "oph_dm_reconstruct": (I, [P, C.POINTER(EyeSpec), P, I, I, P, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def header_functions(path=HEADER):
    """Names of the functions declared in a C header of include/."""
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(oph_\w+)\s*\(", src, flags=re.M)))


def ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, int):
        return a
    return a.ctypes.data_as(C.c_void_p).value


class Context:
    """One oph_ctx: a device, a HIP stream and its scratch."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.oph_create(device, C.byref(h))
        if rc != OPH_OK:
            raise OphError(rc, f"oph_create(device={device}) failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.oph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc):
        if rc != OPH_OK:
            raise OphError(rc, self.lib.oph_last_error(self.h).decode())
        return rc

    # --- device memory ---
    def empty(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        self.check(self.lib.oph_copy_to_device(self.h, d.ptr, ptr(a), a.nbytes))
        return d

    def copy(self, dst, src, nbytes=None):
        """Device-to-device copy between DeviceArrays."""
        self.check(self.lib.oph_device_copy(self.h, ptr(dst), ptr(src), nbytes or src.nbytes))

    def sync(self):
        self.check(self.lib.oph_sync(self.h))

    def timer_begin(self):
        self.check(self.lib.oph_timer_begin(self.h))

    def timer_end(self):
        ms = C.c_float()
        self.check(self.lib.oph_timer_end(self.h, C.byref(ms)))
        return ms.value

    def profile(self, on=True):
        self.check(self.lib.oph_profile_enable(self.h, int(on)))

    def profile_get(self, kernel):
        """(launch count, total ms) of one kernel since profile(True)."""
        n, t = C.c_int(), C.c_double()
        self.check(self.lib.oph_profile_get(self.h, kernel.encode(), C.byref(n), C.byref(t)))
        return n.value, t.value

    # --- encoding and reconstruction ---
    def normalize_u8(self, src, device=False, dst=None):
        """ophNormalize: float fringe (ny, nx) -> flipped u8 (ny, nx)."""
        ny, nx = src.shape
        if not device:
            src = np.ascontiguousarray(src, np.float32)
            dst = np.zeros((ny, nx), np.uint8)
        self.check(self.lib.oph_normalize_u8(self.h, ptr(src), ptr(dst), nx, ny, DEVICE_PTR if device else 0))
        return dst

    def encode_field(self, field, method, passband="left", device=False, out=None, asynchronous=False):
        """MATLAB-toolbox encoding of a complex field (oph_encode_field).
        method: ssb | ni | p2a | real | burckhardt | two_phase; the output is
        (ny, nx) float32, (ny, 3nx) for burckhardt, (ny, 2nx) for two_phase."""
        m = ENCODINGS[method]
        ny, nx = field.shape
        k = 3 if method == "burckhardt" else 2 if method == "two_phase" else 1
        if not device:
            field = np.ascontiguousarray(field, np.complex64)
        if out is None:
            out = self.empty((ny, k * nx), np.float32) if device else np.zeros((ny, k * nx), np.float32)
        flags = (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0)
        self.check(self.lib.oph_encode_field(self.h, ptr(field), nx, ny, m, PASSBANDS[passband], ptr(out),
                                             flags))
        return out

    def scale_u8(self, src, device=False, dst=None):
        """Write_Result_image scaling of a double array -> u8 (same shape)."""
        if not device:
            src = np.ascontiguousarray(src, np.float64)
            dst = np.zeros(src.shape, np.uint8)
        n = int(np.prod(src.shape))
        self.check(self.lib.oph_scale_u8(self.h, ptr(src), ptr(dst), n, DEVICE_PTR if device else 0))
        return dst

    def dm_reconstruct(self, fringe, pitch, lambda_, field_lens, eye_length, pupil, eye_center=(0.0, 0.0),
                       focus=0.0, device=False, out=None):
        """Eye-model reconstruction magnitude of a real fringe (ny, nx)."""
        px, py = pitch if hasattr(pitch, "__len__") else (pitch, pitch)
        ny, nx = fringe.shape
        spec = EyeSpec(px, py, lambda_, field_lens, eye_length, pupil, eye_center[0], eye_center[1], focus)
        if not device:
            fringe = np.ascontiguousarray(fringe, np.float64)
            out = np.zeros((ny, nx), np.float64)
        self.check(self.lib.oph_dm_reconstruct(self.h, C.byref(spec), ptr(fringe), nx, ny, ptr(out),
                                               DEVICE_PTR if device else 0))
        return out

    # --- generators ---
    def pc_rs(self, spec, x, y, z, amp=None, phase=None, out=None, fmt=OUT_C128, accumulate=False,
              device=False, asynchronous=False):
        """Exact RS field (oph_pc_rs).  Host numpy in / out unless device=True
        (then x..phase and out are DeviceArray)."""
        s = spec if isinstance(spec, Spec) else make_spec(**spec)
        pts, keep = _points(x, y, z, amp, phase, device)
        if out is None:
            dt = {OUT_C128: np.complex128, OUT_C64: np.complex64, OUT_FRINGE: np.float32}[fmt]
            out = self.empty((s.ny, s.nx), dt) if device else np.zeros((s.ny, s.nx), dt)
        flags = fmt | (ACCUMULATE if accumulate else 0) | (DEVICE_PTR if device else 0) | \
            (ASYNC if asynchronous else 0)
        self.check(self.lib.oph_pc_rs(self.h, C.byref(s), C.byref(pts), ptr(out), flags))
        del keep
        return out


    def huygens(self, nx, ny, pitch, lambda_, z, x, y, zz, amp=None, phase=None, ph=0.0, out=None,
                fmt=OUT_C128, accumulate=False, device=False, asynchronous=False):
        """Huygens point sum on CWO's grid (oph_pc_huygens), (ny, nx) complex."""
        pts, keep = _points(x, y, zz, amp, phase, device)
        dt = np.complex128 if fmt == OUT_C128 else np.complex64
        if out is None:
            out = self.empty((ny, nx), dt) if device else np.zeros((ny, nx), dt)
        flags = fmt | (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0) | \
            (ACCUMULATE if accumulate else 0)
        self.check(self.lib.oph_pc_huygens(self.h, nx, ny, pitch, lambda_, z, ph, C.byref(pts), ptr(out), flags))
        del keep
        return out

    def fft2d(self, data, direction=-1, normalize=False, device=False, asynchronous=False):
        """In-place complex64 2-D FFT (oph_fft2d).  Host arrays are copied."""
        if not device:
            data = np.ascontiguousarray(data, np.complex64).copy()
        ny, nx = data.shape
        flags = (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0)
        self.check(self.lib.oph_fft2d(self.h, ptr(data), nx, ny, direction, int(normalize), flags))
        return data

    def asm(self, field, pitch, lambda_, z, pad=2, band_limit=False, device=False, asynchronous=False):
        """Angular-spectrum propagation in place (oph_asm_propagate)."""
        px, py = pitch if hasattr(pitch, "__len__") else (pitch, pitch)
        if not device:
            field = np.ascontiguousarray(field, np.complex64).copy()
        ny, nx = field.shape
        flags = (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0)
        self.check(self.lib.oph_asm_propagate(self.h, ptr(field), nx, ny, px, py, lambda_, z, pad,
                                              int(band_limit), flags))
        return field

    def wrp(self, nx, ny, pitch, lambda_, z_wrp, x, y, z, amp=None, mode=0, out=None, accumulate=False,
            device=False, asynchronous=False, phase=None):
        """WRP plane (oph_wrp), (ny, nx) complex64.  phase (radians) is read
        by mode 3 (the MATLAB statement's per-point random phase) only."""
        pts, keep = _points(x, y, z, amp, phase, device)
        if out is None:
            out = self.empty((ny, nx), np.complex64) if device else np.zeros((ny, nx), np.complex64)
        flags = (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0) | \
            (ACCUMULATE if accumulate else 0)
        self.check(self.lib.oph_wrp(self.h, nx, ny, pitch, lambda_, z_wrp, mode, C.byref(pts), ptr(out),
                                    flags))
        del keep
        return out

    def wrp_hologram(self, n_side, pitch, lambda_, z_wrp, distance, x, y, z, amp=None, pad=2,
                     band_limit=False, mode=0, out=None, device=False, asynchronous=False, phase=None):
        """WRP + ASM hologram (oph_wrp_hologram), (n, n) complex64."""
        pts, keep = _points(x, y, z, amp, phase, device)
        if out is None:
            out = self.empty((n_side, n_side), np.complex64) if device else \
                np.zeros((n_side, n_side), np.complex64)
        flags = (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0)
        self.check(self.lib.oph_wrp_hologram(self.h, n_side, pitch, lambda_, z_wrp, distance, pad,
                                             int(band_limit), mode, C.byref(pts), ptr(out), flags))
        del keep
        return out


    def fresnel(self, nx, ny, pitch, lambda_, x, y, z, amp=None, phase=None, method="valu", out=None,
                device=False, asynchronous=False):
        """Fresnel point-cloud field (oph_pc_fresnel), (ny, nx) complex64.
        method: valu | bf16 | bf16_split | fp16 | fp16_split."""
        m = FRESNEL_METHODS[method]
        pts, keep = _points(x, y, z, amp, phase, device)
        if out is None:
            out = self.empty((ny, nx), np.complex64) if device else np.zeros((ny, nx), np.complex64)
        flags = (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0)
        self.check(self.lib.oph_pc_fresnel(self.h, nx, ny, pitch, lambda_, C.byref(pts), m, ptr(out), flags))
        del keep
        return out

    def depthmap(self, w, h, Nx, Ny, pitch, near, far, L, lambdas, intensity, depth, out="fringe",
                 quant_mode=0, Q=256, field_lens=0.0, rho=(1.0, 0.0), layers=(0, 0), device=False,
                 asynchronous=False, dst=None):
        """Depth-map hologram (oph_depthmap).  intensity: (ncolor, h, w) u8,
        depth: (h, w) u8; returns (ncolor, h, w) per `out`."""
        px, py = pitch if hasattr(pitch, "__len__") else (pitch, pitch)
        spec = DmSpec(w, h, Nx, Ny, px, py, near, far, L, quant_mode, Q, field_lens, rho[0], rho[1],
                      layers[0], layers[1])
        lam = np.ascontiguousarray(np.atleast_1d(lambdas), np.float64)
        flag, dt = DM_OUT[out]
        nc = len(lam)
        if not device:
            intensity = np.ascontiguousarray(intensity, np.uint8).reshape(nc, h, w)
            depth = np.ascontiguousarray(depth, np.uint8).reshape(h, w)
        if dst is None:
            dst = self.empty((nc, h, w), dt) if device else np.zeros((nc, h, w), dt)
        flags = flag | (DEVICE_PTR if device else 0) | (ASYNC if asynchronous else 0)
        self.check(self.lib.oph_depthmap(self.h, C.byref(spec), nc, ptr(lam), ptr(intensity), ptr(depth),
                                         ptr(dst), flags))
        return dst


class Cnn:
    """oph_cnn: the residual correction network (NHWC float32 tensors)."""

    def __init__(self, ctx, blocks=8, channels=32, seed=1):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(ctx.lib.oph_cnn_create(ctx.h, blocks, channels, seed, C.byref(h)))
        self.h = h
        self.blocks, self.channels = blocks, channels
        self.n = ctx.lib.oph_cnn_num_params(h)

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.oph_cnn_destroy(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, which=0):
        out = np.zeros(self.n, np.float32)
        self.ctx.check(self.ctx.lib.oph_cnn_get_params(self.ctx.h, self.h, ptr(out), which))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n
        self.ctx.check(self.ctx.lib.oph_cnn_set_params(self.ctx.h, self.h, ptr(p)))

    def forward(self, x, out=None, device=False, shape=None, bf16_acts=False):
        """y = net(x); x (B, H, W, 2) float32 (a complex64 (B, H, W) field
        is the same memory).  `shape` overrides x.shape for device arrays.
        bf16_acts: BF16 activations between layers (inference form)."""
        B, H, W, _ = shape or x.shape
        if not device:
            x = np.ascontiguousarray(x, np.float32)
        if out is None:
            out = self.ctx.empty((B, H, W, 2), np.float32) if device else np.zeros(x.shape, np.float32)
        flags = (DEVICE_PTR if device else 0) | (CNN_BF16_ACTS if bf16_acts else 0)
        self.ctx.check(self.ctx.lib.oph_cnn_forward(self.ctx.h, self.h, ptr(x), ptr(out), B, H, W, flags))
        return out

    def gradients(self, x, t, device=False):
        B, H, W, _ = x.shape
        if not device:
            x = np.ascontiguousarray(x, np.float32)
            t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        self.ctx.check(self.ctx.lib.oph_cnn_gradients(self.ctx.h, self.h, ptr(x), ptr(t), B, H, W,
                                                      C.byref(loss), DEVICE_PTR if device else 0))
        return loss.value, self.params(1)

    def train_step(self, x, t, lr, device=False):
        B, H, W, _ = x.shape
        if not device:
            x = np.ascontiguousarray(x, np.float32)
            t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        self.ctx.check(self.ctx.lib.oph_cnn_train_step(self.ctx.h, self.h, ptr(x), ptr(t), B, H, W, lr,
                                                       C.byref(loss), DEVICE_PTR if device else 0))
        return loss.value

    def save(self, path):
        self.ctx.check(self.ctx.lib.oph_cnn_save(self.ctx.h, self.h, path.encode()))

    def load(self, path):
        self.ctx.check(self.ctx.lib.oph_cnn_load(self.ctx.h, self.h, path.encode()))


class DeviceArray:
    """A device allocation owned by a Context (freed on deletion)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx.lib.oph_device_alloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    def numpy(self):
        a = np.empty(self.shape, self.dtype)
        self.ctx.check(self.ctx.lib.oph_copy_to_host(self.ctx.h, ptr(a), self.ptr, self.nbytes))
        return a

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        self.ctx.check(self.ctx.lib.oph_copy_to_device(self.ctx.h, self.ptr, ptr(a), a.nbytes))

    def zero(self):
        self.ctx.check(self.ctx.lib.oph_device_zero(self.ctx.h, self.ptr, self.nbytes))

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.ctx.lib.oph_device_free(self.ctx.h, self.ptr)
        except Exception:
            pass
        self.ptr = None


ENCODINGS = {"ssb": 0, "ni": 1, "p2a": 2, "real": 3, "burckhardt": 4, "two_phase": 5}
PASSBANDS = {"left": 0, "right": 1, "top": 2, "bottom": 3}


def make_spec(nx, ny, pitch, lambda_, scale=(1.0, 1.0, 1.0), offset_z=0.0, tilt_deg=(0.0, 0.0)):
    if not hasattr(pitch, "__len__"):
        pitch = (pitch, pitch)
    return Spec(nx, ny, pitch[0], pitch[1], lambda_, (C.c_double * 3)(*scale), offset_z,
                (C.c_double * 2)(*tilt_deg))


def _points(x, y, z, amp, phase, device):
    if device:
        arrs = [x, y, z, amp, phase]
        n = x.shape[0]
        return Points(*[ptr(a) for a in arrs], n), arrs
    arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (x, y, z, amp, phase)]
    n = len(arrs[0])
    return Points(*[ptr(a) for a in arrs], n), arrs


def spec_from_keti_config(path):
    s = Spec()
    rc = load().oph_spec_from_keti_config(path.encode(), C.byref(s))
    if rc != OPH_OK:
        raise OphError(rc, f"cannot parse {path}")
    return s


def _check(rc, what):
    if rc != OPH_OK:
        raise OphError(rc, what)


def save_bmp8(path, pixels):
    """8-bit palette BMP, byte-identical to the reference's ophSaveFileBmp."""
    a = np.ascontiguousarray(pixels, np.uint8)
    _check(load().oph_save_bmp8(ptr(a), a.shape[1], a.shape[0], path.encode()), f"cannot write {path}")


def save_bmp24(path, rgb):
    """24-bit BMP of packed RGB888 rows (h x 3w bytes or h x w x 3)."""
    a = np.ascontiguousarray(rgb, np.uint8)
    w = a.shape[1] if a.ndim == 3 else a.shape[1] // 3
    _check(load().oph_save_bmp24(ptr(a), w, a.shape[0], path.encode()), f"cannot write {path}")


def load_bmp_gray8(path):
    lib = load()
    w, h = C.c_int(), C.c_int()
    _check(lib.oph_load_bmp_gray8(path.encode(), C.byref(w), C.byref(h), None, 0), f"cannot read {path}")
    out = np.zeros((h.value, w.value), np.uint8)
    _check(lib.oph_load_bmp_gray8(path.encode(), C.byref(w), C.byref(h), ptr(out), out.size), f"cannot read {path}")
    return out


def _load_points3(fn, path):
    n = C.c_int64()
    _check(fn(path.encode(), None, 0, C.byref(n)), f"cannot load {path}")
    out = np.zeros((n.value, 3), np.float64)
    _check(fn(path.encode(), ptr(out), n.value, C.byref(n)), f"cannot load {path}")
    return out


def load_xyz(path):
    """(n, 3) float64 x, y, z of a WRP xyz text file."""
    return _load_points3(load().oph_load_xyz, path)


def load_cwo_points(path):
    """(n, 3) float64 of a CWO .3d (int32) or .3df (float32) point file."""
    return _load_points3(load().oph_load_cwo_points, path)


def dm_config_load(path):
    c = DmConfig()
    _check(load().oph_dm_config_load(path.encode(), C.byref(c)), f"cannot parse {path}")
    return c


def load_keti_points(path):
    """Returns (n, 5) float64: x, y, z, phase, amplitude (file column order)."""
    lib = load()
    n = C.c_int64()
    rc = lib.oph_load_keti_points(path.encode(), None, 0, C.byref(n))
    if rc != OPH_OK:
        raise OphError(rc, f"cannot load {path}")
    out = np.zeros((n.value, 5), np.float64)
    rc = lib.oph_load_keti_points(path.encode(), ptr(out), n.value, C.byref(n))
    if rc != OPH_OK:
        raise OphError(rc, f"cannot load {path}")
    return out
