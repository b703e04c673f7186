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


def _kept_rows(Ny):
    """Spectrum rows the symmetrisation encode keeps, 0-based: 1-based rows
    c - floor(c/2) .. that + c - 1 with c = floor(Ny/2)
    (HologramGenerator.cpp:494-496, shifted by -1 at :511-514)."""
    c = Ny // 2
    y1 = c - c // 2
    return y1 - 1, y1 + c - 2


def _encode_closed_form(S, w, h, py):
    """Fringe of a centred-order spectrum S as an explicit inverse-DFT sum over
    the kept rows (no FFT): the normalised backward transform between the two
    fftshifts takes bin (y0, x0) to exp(+i 2 pi ((x0 - Nx/2)(c + ox + Nx/2)/Nx
    + (y0 - Ny/2)(r + oy + Ny/2)/Ny)) / (Nx Ny) at scene pixel (r, c), then
    the shift phase exp(i 2 pi yy/(4 py)), yy = h py/2 - py r - py
    (HologramGenerator_CPU.cpp:376-390), and Re."""
    Ny, Nx = S.shape
    oy, ox = (Ny - h) // 2, (Nx - w) // 2
    y1, y2 = _kept_rows(Ny)
    y0 = np.arange(y1, y2 + 1)
    x0 = np.arange(Nx)
    r = np.arange(h)
    c = np.arange(w)
    Ey = np.exp(2j * np.pi * np.outer(r + oy + Ny // 2, y0 - Ny // 2) / Ny)
    Ex = np.exp(2j * np.pi * np.outer(x0 - Nx // 2, c + ox + Nx // 2) / Nx)
    v = Ey @ S[y1:y2 + 1, :] @ Ex / (Nx * Ny)
    yy = h * py / 2 - py * r - py
    return (v * np.exp(2j * np.pi * yy / (4 * py))[:, None]).real


@pytest.mark.parametrize("Ny,Nx,h,w", [(16, 8, 16, 8), (16, 32, 10, 20), (32, 16, 24, 12)])
def test_encode_edges_closed_form(Ny, Nx, h, w):
    """Encode of a hand-built spectrum: complex bins on both sides of both crop
    edges (the outer ones must vanish), at random columns."""
    rng = np.random.default_rng(Ny + Nx + h)
    py = 8e-6
    y1, y2 = _kept_rows(Ny)
    S = np.zeros((Ny, Nx), np.complex128)
    for row in (y1 - 1, y1, (y1 + y2) // 2, y2, y2 + 1):
        for col in rng.choice(Nx, 2, replace=False):
            S[row, col] = rng.normal() + 1j * rng.normal()
    fr = O.dm_encode(S, w, h, py)
    ref = _encode_closed_form(S, w, h, py)
    assert np.abs(fr - ref).max() < 1e-12 * np.abs(ref).max()
    # the outer bins alone encode to zero
    S_out = np.zeros_like(S)
    S_out[[y1 - 1, y2 + 1]] = S[[y1 - 1, y2 + 1]]
    assert np.abs(O.dm_encode(S_out, w, h, py)).max() < 1e-15


@pytest.mark.parametrize("field_lens", [0.0, 0.4])
@pytest.mark.parametrize("pad", [False, True])
def test_depthmap_single_pixel_fringe(pad, field_lens):
    """Whole pipeline of a one-pixel scene, spectrum and encode both in closed
    form; with a field lens the layer sits at z = -f d/(d - f)
    (HologramGenerator.cpp:432)."""
    h, w = (8, 16) if not pad else (6, 10)
    Ny, Nx = (h, w) if not pad else (16, 32)
    img = np.zeros((h, w), np.uint8)
    dep = np.zeros((h, w), np.uint8)
    img[3, 5], dep[3, 5] = 180, 247   # mode 0: layer 256 - 247 = 9
    px, py, lam, near, far = 8e-6, 8e-6, 473e-9, 0.09, 0.11
    _, fr = O.depthmap(img, dep, Nx, Ny, px, py, lam, near, far, 256, field_lens=field_lens)
    d = O.depth_levels(near, far, 256)[8]
    zl = -field_lens * d / (d - field_lens) if field_lens > 0 else d
    S = _single_pixel_spectrum(Nx, Ny, h, w, 3, 5, 180 / 255, px, py, lam, zl)
    ref = _encode_closed_form(S, w, h, py)
    assert np.abs(fr - ref).max() < 1e-9 * np.abs(ref).max()
