"""CPU: pins of the depth-map oracle (the reference's DepthMap CPU path,
restated; FFTW and Qt are absent so the reference itself cannot run here)."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O


def test_depth_levels_drift():
    """Repeated addition (HologramGenerator.cpp:396-401): 90-110 mm in 256
    levels drops the last level in the reference's while loop; the oracle
    continues the same sequence to 256 entries."""
    near, far = 0.09, 0.11
    lev = O.depth_levels(near, far, 256)
    dstep = (far - near) / 255
    v, seq = near, []
    while v <= far:
        seq.append(v)
        v += dstep
    assert len(seq) == 255
    assert np.array_equal(lev[:255], np.array(seq))
    assert lev[255] == seq[-1] + dstep and abs(lev[255] - far) < 1e-15


def test_depth_index_modes():
    d = np.arange(256, dtype=np.uint8)
    assert np.array_equal(O.depth_index(d, 0, 256, 0.2, 0.7, 256), 256 - d.astype(np.int32))
    # requantised to 16 levels: every u8 depth lands in exactly one level and
    # the level index falls as depth rises (d = 255 is the near plane)
    idx = O.depth_index(d, 1, 256, 0.09, 0.11, 16)
    assert idx.min() >= 1 and idx.max() <= 16
    assert np.all(np.diff(idx) <= 0) and idx[-1] == 1 and idx[0] == 16


def _single_pixel_spectrum(Nx, Ny, h, w, r0, c0, val, px, py, lam, zl):
    """Closed form of one layer holding one pixel, centred-order spectrum,
    evaluated without an FFT: fftshift moves index n to (n - N/2) mod N, the
    DFT of a delta is a phase ramp, the output shift reads X[(t + N/2) mod N]."""
    oy, ox = (Ny - h) // 2, (Nx - w) // 2
    i0, j0 = r0 + oy, c0 + ox
    ip, jp = (i0 - Ny // 2) % Ny, (j0 - Nx // 2) % Nx
    t = np.arange(Ny)[:, None]
    s = np.arange(Nx)[None, :]
    ky, kx = (t + Ny // 2) % Ny, (s + Nx // 2) % Nx
    X = val * np.exp(-2j * np.pi * (ky * ip / Ny + kx * jp / Nx))
    fx = -1 / (2 * px) + s / (Nx * px)
    fy = 1 / (2 * py) - 1 / (Ny * py) - t / (Ny * py)
    k = 2 * np.pi / lam
    H = np.exp(1j * k * (-zl) * np.sqrt(1 - (lam * fx) ** 2 - (lam * fy) ** 2))
    return np.exp(1j * k * zl) * H * X


@pytest.mark.parametrize("pad", [False, True])
def test_depthmap_single_pixel_spectrum(pad):
    h, w = 6, 10
    Ny, Nx = (16, 32) if pad else (h, w)
    img = np.zeros((h, w), np.uint8)
    dep = np.zeros((h, w), np.uint8)
    img[2, 7] = 200
    dep[2, 7] = 250                 # mode 0: layer 256 - 250 = 6
    px, py, lam, near, far = 8e-6, 8e-6, 532e-9, 0.09, 0.11
    if not pad:
        # unpadded (the reference's own case): scene = grid, power-of-two sides
        h, w, Ny, Nx = 8, 16, 8, 16
        img = np.zeros((h, w), np.uint8); dep = np.zeros((h, w), np.uint8)
        img[2, 7], dep[2, 7] = 200, 250
    spec, _ = O.depthmap(img, dep, Nx, Ny, px, py, lam, near, far, 256, want_spectrum=True,
                         want_fringe=False)
    zl = O.depth_levels(near, far, 256)[5]
    ref = _single_pixel_spectrum(Nx, Ny, h, w, 2, 7, 200 / 255, px, py, lam, zl)
    assert rel_l2(spec, ref) < 1e-8  # the closed form above keeps k z ~ 1e6 rad unreduced


def test_depthmap_layer_linearity_and_field():
    rng = np.random.default_rng(4)
    h, w = 12, 20
    img = rng.integers(0, 256, (h, w)).astype(np.uint8)
    dep = rng.integers(0, 256, (h, w)).astype(np.uint8)
    args = (32, 16, 8e-6, 8e-6, 633e-9, 0.09, 0.11, 16)
    kw = dict(quant_mode=1, want_spectrum=True, want_field=True)
    s_all, f_all, u_all = O.depthmap(img, dep, *args, **kw)
    s1, f1, u1 = O.depthmap(img, dep, *args, layers=(1, 7), **kw)
    s2, f2, u2 = O.depthmap(img, dep, *args, layers=(8, 16), **kw)
    assert rel_l2(s1 + s2, s_all) < 1e-12
    assert rel_l2(f1 + f2, f_all) < 1e-12
    assert rel_l2(u1 + u2, u_all) < 1e-12
    # the field is the normalised inverse of the spectrum: Parseval over the
    # whole grid bounds the scene-window energy
    assert np.sum(abs(u_all) ** 2) <= np.sum(abs(s_all) ** 2) / (32 * 16) * (1 + 1e-12)
    # a zero-intensity scene gives nothing (alpha = [img > 0])
    z, fz = O.depthmap(np.zeros_like(img), dep, *args, quant_mode=1)
    assert np.all(fz == 0)
