"""ctypes bindings of the CPU oracle (oph_oracle.c) and of the compiled
reference KETI sources (oracle/_ref/libketi_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the CPU restatement the GPU path is compared with.  The product package never imports
this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboph_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libketi_ref.so")


def build(ref=True):
    """Compile the oracle (and the reference KETI library when the reference
    tree exists) with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all" if ref else LIB])


class _Spec(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("pitch_x", C.c_double),
                ("pitch_y", C.c_double), ("lambda_", C.c_double),
                ("scale", C.c_double * 3), ("offset_z", C.c_double),
                ("tilt_deg", C.c_double * 2)]


class _KetiSpecF32(C.Structure):
    _fields_ = [("scale", C.c_float * 3), ("offset_depth", C.c_float),
                ("pitch", C.c_float * 2), ("nx", C.c_int), ("ny", C.c_int),
                ("lambda_", C.c_float), ("tilt", C.c_float * 2)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        _lib = C.CDLL(LIB)
        P = C.c_void_p
        _lib.orc_rs.argtypes = [C.POINTER(_Spec), C.c_int64, P, P, P, P, P, P, C.c_int]
        _lib.orc_keti_fringe_f32.argtypes = [C.POINTER(_KetiSpecF32), C.c_int, P, P, P, C.c_int]
        _lib.orc_normalize_u8.argtypes = [P, P, C.c_int, C.c_int]
        _lib.orc_scale_u8.argtypes = [P, P, C.c_int64]
        _lib.orc_dm_reconstruct.argtypes = [P, C.c_int, C.c_int] + [C.c_double] * 9 + [P, C.c_int]
        _lib.orc_fresnel.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int64,
                                     P, P, P, P, P, P, P, C.c_int64, P, C.c_int]
        _lib.orc_huygens.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int64, P, P, P, P, P, P, C.c_int]
        _lib.orc_wrp.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int64, P, P, P, P, P, P, C.c_int]
        _lib.orc_fft2d.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        _lib.orc_asm.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int, C.c_int]
        _lib.orc_asm_transfer.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_int, P]
        _lib.orc_dm_encode.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, P,
                                       C.c_int]
        _lib.orc_depth_levels.argtypes = [C.c_double, C.c_double, C.c_int, P]
        _lib.orc_depth_index.argtypes = [C.c_int, P, C.c_int, C.c_int, C.c_double,
                                         C.c_double, C.c_int, P]
        _lib.orc_depthmap.argtypes = [C.c_int, C.c_int, P, P, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_int, C.c_int, P, P, P, C.c_int]
    return _lib


def ref_lib():
    """The reference KETI library, or None when it was not built (it is built
    only where the reference tree exists)."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = C.CDLL(REF_LIB)
        P = C.c_void_p
        _ref.ref_keti_load.argtypes = [C.c_char_p, P, C.c_int]
        _ref.ref_keti_spec.argtypes = [C.c_char_p, P]
        _ref.ref_keti_fringe.argtypes = [P, C.c_int, P, P]
        _ref.ref_keti_normalize.argtypes = [P, P, C.c_int, C.c_int]
    return _ref


def _sizes(*ns):
    for n in ns:
        if n < 1:
            raise ValueError(f"transform sizes must be positive, got {n}")


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def rs(spec, x, y, z, amp, phase=None, nthreads=0):
    """FP64 complex RS field, (ny, nx) complex128. spec: dict with nx, ny,
    pitch (px, py), lambda_, scale (3), offset_z, tilt_deg (2)."""
    pitch = spec["pitch"] if hasattr(spec["pitch"], "__len__") else (spec["pitch"],) * 2
    s = _Spec(spec["nx"], spec["ny"], pitch[0], pitch[1], spec["lambda_"],
              (C.c_double * 3)(*spec.get("scale", (1, 1, 1))), spec.get("offset_z", 0.0),
              (C.c_double * 2)(*spec.get("tilt_deg", (0, 0))))
    x, y, z, amp, phase = map(_d, (x, y, z, amp, phase))
    out = np.zeros((spec["ny"], spec["nx"]), np.complex128)
    lib().orc_rs(C.byref(s), len(x), _p(x), _p(y), _p(z), _p(amp), _p(phase), _p(out), nthreads)
    return out


def keti_fringe_f32(spec13, pts5, dst=None, nthreads=0):
    """FP32 reference-faithful fringe. spec13 / pts5 use ref_keti_* layouts."""
    sp = np.asarray(spec13, np.float32)
    s = _KetiSpecF32((C.c_float * 3)(*sp[0:3]), sp[3], (C.c_float * 2)(*sp[4:6]),
                     int(sp[6]), int(sp[7]), sp[8], (C.c_float * 2)(*sp[9:11]))
    pts = np.ascontiguousarray(pts5, np.float32)
    xyz = np.ascontiguousarray(pts[:, 0:3])
    amp = np.ascontiguousarray(pts[:, 3])
    if dst is None:
        dst = np.zeros((int(sp[7]), int(sp[6])), np.float32)
    lib().orc_keti_fringe_f32(C.byref(s), len(pts), _p(xyz), _p(amp), _p(dst), nthreads)
    return dst


def normalize_u8(src):
    src = np.ascontiguousarray(src, np.float32)
    dst = np.zeros(src.shape, np.uint8)
    lib().orc_normalize_u8(_p(src), _p(dst), src.shape[1], src.shape[0])
    return dst


def scale_u8(src):
    """Write_Result_image 8-bit scaling of a double array (no flip)."""
    src = np.ascontiguousarray(src, np.float64)
    dst = np.zeros(src.shape, np.uint8)
    lib().orc_scale_u8(_p(src), _p(dst), src.size)
    return dst


def dm_reconstruct(fringe, ppx, ppy, lam, field_lens, eye_length, pupil, cx, cy, focus, nthreads=0):
    """Eye-model reconstruction magnitude (ny x nx doubles) of a real fringe."""
    f = np.ascontiguousarray(fringe, np.float64)
    _sizes(*f.shape)
    out = np.zeros(f.shape, np.float64)
    lib().orc_dm_reconstruct(_p(f), f.shape[1], f.shape[0], ppx, ppy, lam, field_lens, eye_length, pupil,
                             cx, cy, focus, _p(out), nthreads)
    return out


def fresnel(nx, ny, pitch, lam, x, y, z, amp, phase=None, rows=None, cols=None, nthreads=0):
    x, y, z, amp, phase = map(_d, (x, y, z, amp, phase))
    if rows is not None:
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        out = np.zeros(len(rows), np.complex128)
        npix = len(rows)
    else:
        out = np.zeros((ny, nx), np.complex128)
        npix = 0
    lib().orc_fresnel(nx, ny, pitch, lam, len(x), _p(x), _p(y), _p(z), _p(amp), _p(phase),
                      _p(rows), _p(cols), npix, _p(out), nthreads)
    return out


def huygens(nx, ny, pitch, lam, z, x, y, zz, amp=None, phase=None, ph=0.0, nthreads=0):
    """FP64 cwoPLS::Huygens field with amplitude and phase, (ny, nx) complex128."""
    x, y, zz, amp, phase = map(_d, (x, y, zz, amp, phase))
    out = np.zeros((ny, nx), np.complex128)
    lib().orc_huygens(nx, ny, pitch, lam, z, ph, len(x), _p(x), _p(y), _p(zz), _p(amp), _p(phase), _p(out),
                      nthreads)
    return out


def wrp(nx, ny, pitch, lam, z_wrp, x, y, z, amp=None, mode=0, phase=None, nthreads=0):
    x, y, z, amp, phase = map(_d, (x, y, z, amp, phase))
    out = np.zeros((ny, nx), np.complex128)
    lib().orc_wrp(nx, ny, pitch, lam, z_wrp, mode, len(x), _p(x), _p(y), _p(z), _p(amp), _p(phase),
                  _p(out), nthreads)
    return out


def fft2d(a, sign=-1, normalize=False, nthreads=0):
    a = np.array(a, np.complex128, order="C", copy=True)
    _sizes(*a.shape)
    lib().orc_fft2d(_p(a), a.shape[1], a.shape[0], sign, int(normalize), nthreads)
    return a


def asm(field, px, py, lam, z, pad=2, band_limit=False, nthreads=0):
    f = np.array(field, np.complex128, order="C", copy=True)
    _sizes(f.shape[0] * pad, f.shape[1] * pad)
    lib().orc_asm(_p(f), f.shape[1], f.shape[0], px, py, lam, z, pad, int(band_limit), nthreads)
    return f


def asm_transfer(Npx, Npy, px, py, lam, z, band_limit=False):
    """Angular-spectrum transfer on an Npy x Npx grid in FFT order."""
    H = np.zeros((Npy, Npx), np.complex128)
    lib().orc_asm_transfer(Npx, Npy, px, py, lam, z, int(band_limit), _p(H))
    return H


def dm_encode(spectrum, w, h, py, nthreads=0):
    """Symmetrisation encode of a centred-order spectrum (Ny x Nx) to an
    h x w real fringe (scene window centred in the grid)."""
    U = np.array(spectrum, np.complex128, order="C", copy=True)
    _sizes(*U.shape)
    fr = np.zeros((h, w), np.float64)
    lib().orc_dm_encode(_p(U), U.shape[1], U.shape[0], w, h, py, _p(fr), nthreads)
    return fr


def depth_levels(near, far, L):
    out = np.zeros(L, np.float64)
    lib().orc_depth_levels(near, far, L, _p(out))
    return out


def depth_index(dep8, quant_mode, Q, near, far, L):
    d = np.ascontiguousarray(dep8, np.uint8)
    idx = np.zeros(d.shape, np.int32)
    lib().orc_depth_index(d.size, _p(d), quant_mode, Q, near, far, L, _p(idx))
    return idx


def depthmap(img8, dep8, Nx, Ny, px, py, lam, near, far, L, quant_mode=0, Q=256,
             field_lens=0.0, rho=(1.0, 0.0), layers=(1, 0), want_spectrum=False,
             want_fringe=True, want_field=False, nthreads=0):
    """FP64 depth-map hologram of one colour.  Returns (spectrum, fringe) or,
    with want_field, (spectrum, fringe, field); entries not asked for are None."""
    img8 = np.ascontiguousarray(img8, np.uint8)
    dep8 = np.ascontiguousarray(dep8, np.uint8)
    h, w = img8.shape
    _sizes(Nx, Ny)
    if w > Nx or h > Ny:
        raise ValueError("scene larger than the transform grid")
    spec = np.zeros((Ny, Nx), np.complex128) if want_spectrum else None
    fr = np.zeros((h, w), np.float64) if want_fringe else None
    fld = np.zeros((h, w), np.complex128) if want_field else None
    lib().orc_depthmap(w, h, _p(img8), _p(dep8), Nx, Ny, px, py, lam, near, far, L,
                       quant_mode, Q, field_lens, rho[0], rho[1], layers[0], layers[1],
                       _p(spec), _p(fr), _p(fld), nthreads)
    if want_field:
        return spec, fr, fld
    return spec, fr
