"""CPU: pins of the oracle's FFT, angular-spectrum propagation and WRP
accumulation by mathematics (the reference's FFTW/cuFFT and CWO kernels are
third-party or closed; see the oracle header)."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O


def test_fft_vs_numpy_and_dft():
    rng = np.random.default_rng(1)
    for ny, nx in [(32, 64), (128, 16), (256, 512)]:
        a = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
        assert rel_l2(O.fft2d(a, -1), np.fft.fft2(a)) < 1e-13
        assert rel_l2(O.fft2d(a, +1, normalize=True), np.fft.ifft2(a)) < 1e-13
    # brute-force O(N^2) DFT, independent of numpy's FFT
    a = rng.standard_normal((8, 16)) + 1j * rng.standard_normal((8, 16))
    ky, kx = np.arange(8)[:, None, None, None], np.arange(16)[None, :, None, None]
    y, x = np.arange(8)[None, None, :, None], np.arange(16)[None, None, None, :]
    W = np.exp(-2j * np.pi * (ky * y / 8 + kx * x / 16))
    dft = (W * a[None, None]).sum(axis=(2, 3))
    assert rel_l2(O.fft2d(a, -1), dft) < 1e-13
    # Parseval
    b = O.fft2d(a, -1)
    assert abs(np.sum(abs(b) ** 2) / a.size - np.sum(abs(a) ** 2)) < 1e-9


def test_fft_mixed_radix_vs_numpy_and_dft():
    """Lengths with factors 3, 5 and larger primes (the shipped depth-map
    grid is 11520 x 2160 = 2^8 3^2 5 x 2^4 3^3 5): numpy's FFT, a direct
    O(N^2) DFT, and the full-size transform against numpy on one axis."""
    rng = np.random.default_rng(3)
    for ny, nx in [(6, 10), (15, 12), (45, 30), (135, 18), (7, 11), (540, 720), (2160, 8)]:
        a = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
        assert rel_l2(O.fft2d(a, -1), np.fft.fft2(a)) < 1e-13, (ny, nx)
        assert rel_l2(O.fft2d(a, +1, normalize=True), np.fft.ifft2(a)) < 1e-13, (ny, nx)
    ny, nx = 9, 15
    a = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    ky, kx = np.arange(ny)[:, None, None, None], np.arange(nx)[None, :, None, None]
    y, x = np.arange(ny)[None, None, :, None], np.arange(nx)[None, None, None, :]
    dft = (np.exp(-2j * np.pi * (ky * y / ny + kx * x / nx)) * a[None, None]).sum(axis=(2, 3))
    assert rel_l2(O.fft2d(a, -1), dft) < 1e-13
    a = rng.standard_normal((2, 11520)) + 1j * rng.standard_normal((2, 11520))
    assert rel_l2(O.fft2d(a, -1), np.fft.fft2(a)) < 1e-13


def test_asm_plane_wave_eigenfunction():
    """Without padding a grid-periodic plane wave is an eigenfunction: the
    output is H(fx, fy) times the input, H = exp(i 2 pi z sqrt(1/l^2 - f^2))."""
    n, p, lam, z = 64, 8e-6, 532e-9, 0.05
    kx0, ky0 = 5, -3
    yy, xx = np.mgrid[0:n, 0:n]
    u = np.exp(2j * np.pi * (kx0 * xx + ky0 * yy) / n)
    out = O.asm(u, p, p, lam, z, pad=1)
    fx, fy = kx0 / (n * p), ky0 / (n * p)
    H = np.exp(2j * np.pi * z * np.sqrt(1 / lam ** 2 - fx ** 2 - fy ** 2))
    assert np.abs(out - H * u).max() < 1e-9


def test_asm_forward_back_identity_and_evanescent():
    rng = np.random.default_rng(2)
    u = rng.standard_normal((64, 32)) + 1j * rng.standard_normal((64, 32))
    v = O.asm(O.asm(u, 8e-6, 8e-6, 532e-9, 0.03, pad=1), 8e-6, 8e-6, 532e-9, -0.03, pad=1)
    assert rel_l2(v, u) < 1e-12
    # pitch below lambda/2: high frequencies are evanescent and removed
    w = O.asm(u, 0.2e-6, 0.2e-6, 532e-9, 1e-6, pad=1)
    spec = np.fft.fft2(w)
    f = np.fft.fftfreq(64, 0.2e-6)[:, None] ** 2 + np.fft.fftfreq(32, 0.2e-6)[None, :] ** 2
    assert np.abs(spec[f >= 1 / 532e-9 ** 2]).max() < 1e-9


def test_asm_spherical_wave_propagation():
    """Padded ASM carries the sampled spherical wave exp(ikr)/r of a point at
    z1 to the field of the same point at z1 + dz (central region)."""
    n, p, lam, z1, dz = 256, 8e-6, 532e-9, 0.1, 0.004
    c = (np.arange(n) - n / 2 + 0.5) * p
    X, Y = np.meshgrid(c, c)

    def sph(z):
        r = np.sqrt(X ** 2 + Y ** 2 + z ** 2)
        t = (r / lam) % 1.0
        return np.exp(2j * np.pi * t) / r
    out = O.asm(sph(z1), p, p, lam, dz, pad=2)
    ref = sph(z1 + dz)
    m = slice(n // 4, 3 * n // 4)
    assert rel_l2(out[m, m], ref[m, m]) < 2e-2
    # a wrong sign of z would go the other way
    bad = O.asm(sph(z1), p, p, lam, -dz, pad=2)
    assert rel_l2(bad[m, m], ref[m, m]) > 0.5


def test_wrp_single_point_closed_form():
    n, p, lam = 64, 8e-6, 532e-9
    x0, y0, z0, zw = 3.3e-5, -4.1e-5, -0.103, -0.1
    out = O.wrp(n, n, p, lam, zw, [x0], [y0], [z0])
    dz = zw - z0
    w = int(abs(dz) * np.tan(np.arcsin(lam / (2 * p))) / p)
    assert w == 12
    tx, ty = int(x0 / p) + n // 2, int(y0 / p) + n // 2
    nz = np.argwhere(out != 0)
    assert nz[:, 0].min() == ty - w and nz[:, 0].max() == ty + w - 1
    assert nz[:, 1].min() == tx - w and nz[:, 1].max() == tx + w - 1
    L = np.longdouble
    wy, wx = np.mgrid[-w:w, -w:w]
    r = np.sqrt((wx * L(p)) ** 2 + (wy * L(p)) ** 2 + L(dz) ** 2)  # dz > 0: sign +
    t = (r / L(lam)) % 1
    ref = np.exp(2j * np.pi * t.astype(np.float64)) / (r.astype(np.float64) + 0.05)
    assert np.abs(out[ty - w:ty + w, tx - w:tx + w] - ref).max() < 1e-9  # |v| ~ 20


def test_wrp_modes_sign_and_clipping():
    n, p, lam = 48, 8e-6, 532e-9
    # mode 1 half-width of the reference, WRP.cpp:111
    dz = 0.002
    w1 = int(abs(lam * dz / p / p / 2 + 0.5)) * 2 - 1
    out = O.wrp(n, n, p, lam, 0.0, [0.0], [0.0], [-dz], mode=1)
    nz = np.argwhere(out != 0)
    assert nz[:, 0].max() - nz[:, 0].min() + 1 == min(2 * w1, n)
    # negative dz: r is negative -> exp(-ik|r|)/(0.05 - |r|)
    o2 = O.wrp(n, n, p, lam, 0.0, [0.0], [0.0], [+0.001])
    v = o2[n // 2, n // 2]
    t = (-0.001 / lam) % 1
    assert abs(v - np.exp(2j * np.pi * t) / (0.05 - 0.001)) < 1e-9
    # a point near the corner is clipped, not wrapped
    o3 = O.wrp(n, n, p, lam, 0.0, [-(n // 2 - 2) * p], [-(n // 2 - 2) * p], [-0.004])
    nz = np.argwhere(o3 != 0)
    assert nz.min() == 0 and nz[:, 0].max() < n // 2


def test_wrp_linearity():
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-1e-4, 1e-4, (2, 30))
    z = rng.uniform(-0.106, -0.101, 30)
    a = rng.uniform(0, 1, 30)
    f = O.wrp(64, 64, 8e-6, 532e-9, -0.1, x, y, z, a)
    g = O.wrp(64, 64, 8e-6, 532e-9, -0.1, x[:11], y[:11], z[:11], a[:11]) + \
        O.wrp(64, 64, 8e-6, 532e-9, -0.1, x[11:], y[11:], z[11:], a[11:])
    assert rel_l2(f, g) < 1e-13


def _bl_cutoff(lam, df, z):
    """Matsushima's band limit on one frequency axis with the other at zero
    (K. Matsushima, T. Shimobaba, Opt. Express 17, 19662 (2009), eq. 13):
    |f| <= 1/(lam sqrt((2 df z)^2 + 1)), df the grid's frequency step."""
    return 1.0 / (lam * np.sqrt((2 * df * z) ** 2 + 1))


@pytest.mark.parametrize("Npx,Npy,px,py,z", [(64, 128, 8e-6, 6e-6, 0.016),
                                              (256, 128, 8e-6, 8e-6, 0.1)])
def test_asm_band_limit_axis_cutoff(Npx, Npy, px, py, z):
    lam = 532e-9
    H = O.asm_transfer(Npx, Npy, px, py, lam, z, band_limit=True)
    H0 = O.asm_transfer(Npx, Npy, px, py, lam, z, band_limit=False)
    for n, p, line, line0 in ((Npx, px, H[0, :], H0[0, :]), (Npy, py, H[:, 0], H0[:, 0])):
        df = 1.0 / (n * p)
        f = np.fft.fftfreq(n, p)
        cut = _bl_cutoff(lam, df, z)
        inside = np.abs(f) <= cut
        # the cutoff falls between bins (no rounding question) and mid-band,
        # where a threshold off by 2x would keep a different set of bins
        assert np.min(np.abs(np.abs(f) - cut)) > 0.05 * df
        assert 4 <= inside.sum() <= n - 4
        for scale in (0.5, 2.0):
            cut2 = 1.0 / (lam * np.sqrt((df * z / scale) ** 2 * 4 + 1))
            assert not np.array_equal(np.abs(f) <= cut2, inside)
        assert np.array_equal(line != 0, inside)
        assert np.array_equal(line[inside], line0[inside])
        ref = np.exp(2j * np.pi * z * np.sqrt(1 / lam ** 2 - f ** 2))
        assert np.abs(line0 - ref).max() < 1e-9


def test_asm_band_limited_plane_waves():
    """Through orc_asm itself (no padding): a plane wave just inside the
    band-limit cutoff passes with H, one just outside is removed."""
    n, p, lam, z = 64, 8e-6, 532e-9, 0.016
    df = 1.0 / (n * p)
    kc = int(_bl_cutoff(lam, df, z) / df)
    yy, xx = np.mgrid[0:n, 0:n]
    for kx0, ky0, passes in ((kc, 0, True), (kc + 1, 0, False), (0, -kc, True),
                             (0, -kc - 1, False)):
        u = np.exp(2j * np.pi * (kx0 * xx + ky0 * yy) / n)
        out = O.asm(u, p, p, lam, z, pad=1, band_limit=True)
        if passes:
            fx, fy = kx0 * df, ky0 * df
            Hv = np.exp(2j * np.pi * z * np.sqrt(1 / lam ** 2 - fx ** 2 - fy ** 2))
            assert np.abs(out - Hv * u).max() < 1e-9
        else:
            assert np.abs(out).max() < 1e-12


@pytest.mark.parametrize("band_limit", [False, True])
def test_asm_padded_composition(band_limit):
    """Padded ASM = crop(IFFT(H . FFT(embed))) with the embedding offset
    (Npx/2 - nx/2, Npy/2 - ny/2) of CWO::Diffract (cwo.cpp:2235-2239), built
    here with numpy's FFT and the transfer alone."""
    rng = np.random.default_rng(7)
    ny, nx, p, lam, z = 16, 32, 8e-6, 633e-9, 0.02
    u = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    Npy, Npx = 2 * ny, 2 * nx
    g = np.zeros((Npy, Npx), np.complex128)
    oy, ox = Npy // 2 - ny // 2, Npx // 2 - nx // 2
    g[oy:oy + ny, ox:ox + nx] = u
    H = O.asm_transfer(Npx, Npy, p, p, lam, z, band_limit)
    ref = np.fft.ifft2(H * np.fft.fft2(g))[oy:oy + ny, ox:ox + nx]
    assert rel_l2(O.asm(u, p, p, lam, z, pad=2, band_limit=band_limit), ref) < 1e-12


def test_asm_paraxial_limit_fresnel_tf():
    """A beam of low spatial frequencies: the angular spectrum agrees with the
    Fresnel transfer function H = exp(ikz) exp(-i pi lam z (fx^2 + fy^2))
    (WM/FresnelPropagation.m:7-14; Goodman, Fourier Optics, eq. 4-21), the
    difference being the phase pi z lam^3 f^4 / 4 ~ 1e-7 rad here."""
    n, p, lam, z = 128, 8e-6, 532e-9, 0.02
    c = (np.arange(n) - n / 2) * p
    X, Y = np.meshgrid(c, c)
    u = np.exp(-(X ** 2 + Y ** 2) / (20 * p) ** 2) * np.exp(2j * np.pi * 3 * X / (n * p))
    out = O.asm(u, p, p, lam, z, pad=1)
    f = np.fft.fftfreq(n, p)
    FX, FY = np.meshgrid(f, f)
    HF = np.exp(2j * np.pi * ((z / lam) % 1)) * np.exp(-1j * np.pi * lam * z * (FX ** 2 + FY ** 2))
    ref = np.fft.ifft2(np.fft.fft2(u) * HF)
    assert rel_l2(out, ref) < 1e-5
    assert rel_l2(out, u * np.exp(2j * np.pi * ((z / lam) % 1))) > 0.1  # it did diffract
    bad = O.asm(u, p, p, lam, -z, pad=1)
    assert rel_l2(bad, ref) > 0.5


def test_wrp_mode1_vs_reference_output():
    """The reference example's own output (W/openwrp/openwrp.cpp:15-18):
    SingleWRP(0.5 mm) of object.txt on a 515^2 plane at 8 um, 532 nm, saved
    as an arg image (tests/golden/make_wrp_golden.py).  The reference sums in
    FP32 (cosf of a float wave number times r, ~1e-3 rad per term) with a
    racing OpenMP AddPixel, and CWO's arg writer is closed; its quantisation
    was read off the file (zero pixels are 127: floor).  Measured: 99.19 %
    of pixels equal, 99.989 % within one level."""
    import os
    from conftest import GOLDEN, wrp_table_agreement
    g = np.load(os.path.join(GOLDEN, "wrp_table.npz"), allow_pickle=False)
    xyz = g["xyz"].astype(np.float64)
    lam, p, n, zw = (float(v) for v in g["params"])
    U = O.wrp(int(n), int(n), p, lam, zw, xyz[:, 0], xyz[:, 1], xyz[:, 2], mode=1)
    exact, near = wrp_table_agreement(U, g["gray"])
    assert exact >= 0.99 and near >= 0.9995, (exact, near)
    # mode 0 (the grating-angle window) or a flipped plane is far off
    U0 = O.wrp(int(n), int(n), p, lam, zw, xyz[:, 0], xyz[:, 1], xyz[:, 2], mode=0)
    assert wrp_table_agreement(U0, g["gray"])[0] < 0.8
    assert wrp_table_agreement(U[::-1], g["gray"])[0] < 0.8   # 70 % of pixels are empty in both


def test_wrp_mode2_is_windowed_fresnel():
    """cwoPLS::WRP (mode 2) is the Fresnel point field restricted to each
    point's window: for points on the pixel grid (pitch 2^-17 m, so x/p is
    exact) the plane equals the sum of orc_fresnel's single-point fields at
    distance dz = z_wrp - z_n, each masked to [-w, w)^2 around its pixel with
    w = (int)(|dz| lambda/(2p)/p) (cwoPLS.cpp:557)."""
    n, p, lam, zw = 64, 2.0 ** -17, 532e-9, 0.0
    pts = [(3, -5, -0.004, 0.7), (-10, 8, -0.0061, 1.0), (20, 20, -0.0035, 0.4)]
    x = np.array([q[0] * p for q in pts])
    y = np.array([q[1] * p for q in pts])
    z = np.array([q[2] for q in pts])
    a = np.array([q[3] for q in pts])
    out = O.wrp(n, n, p, lam, zw, x, y, z, a, mode=2)
    ref = np.zeros((n, n), np.complex128)
    for k in range(len(pts)):
        d = zw - z[k]
        w = int(abs(d) * (lam / (2 * p)) / p)
        assert 3 <= w < 40
        f = O.fresnel(n, n, p, lam, x[k:k + 1], y[k:k + 1], [d], a[k:k + 1])
        tx, ty = pts[k][0] + n // 2, pts[k][1] + n // 2
        m = np.zeros((n, n), bool)
        m[max(ty - w, 0):ty + w, max(tx - w, 0):tx + w] = True
        ref += np.where(m, f, 0)
    assert np.abs(out - ref).max() < 1e-9
    # no 1/(r + 0.05) and a different phase from mode 0
    assert rel_l2(O.wrp(n, n, p, lam, zw, x, y, z, a, mode=0), out) > 0.5


def _mround(v):
    """MATLAB round: half away from zero"""
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def _matlab_wrp(obj, lam, p, res, z_wrp, phases):
    """Hologram_WRP.m:16-36 transcribed (OpenHolo_WRP/WRP_matlab_ver 1.0):
    1-based index arithmetic, the temp array of res + N and the crop,
    meshgrid as MATLAB's, the script's rand*2*pi replaced by `phases`."""
    k = 2 * np.pi / lam
    z = z_wrp - obj[:, 2]
    N = _mround(np.abs(lam * z / (p ** 2) / 2) + 0.5) * 2 - 1
    Nx = _mround(obj[:, 0] / p) + (res - 1) / 2
    Ny = _mround(obj[:, 1] / p) + (res - 1) / 2
    H = np.zeros((res, res), np.complex128)
    for o in range(len(obj)):
        n = int(N[o])
        v = np.arange(-(n - 1) / 2, (n - 1) / 2 + 1) * p
        y_run, x_run = np.meshgrid(v, v)
        r = np.sign(z[o]) * np.sqrt(z[o] ** 2 + y_run ** 2 + x_run ** 2)
        sub = np.exp(1j * phases[o]) * np.exp(1j * k * r) / r
        temp = np.zeros((res + n, res + n), np.complex128)
        a, b = int(Nx[o]), int(Ny[o])
        temp[a - 1:a - 1 + n, b - 1:b - 1 + n] = sub
        H += temp[(n + 1) // 2 - 1:res + (n - 1) // 2, (n + 1) // 2 - 1:res + (n - 1) // 2]
    return H


def test_wrp_mode3_matches_matlab_statement():
    """Mode 3 against the MATLAB script's own construction (square N x N
    window, 1/r, per-point phase, x along the script's rows): the oracle's
    plane is the transpose of the script's matrix.  Odd resolution as in the
    script (1025), points of both signs of dz, some windows clipped."""
    rng = np.random.default_rng(8)
    res, p, lam, zw = 41, 3.9e-6, 532e-9, 0.5e-3
    n = 9
    obj = np.stack([rng.uniform(-12, 12, n) * p, rng.uniform(-12, 12, n) * p,
                    rng.uniform(0.2e-3, 0.9e-3, n)], axis=1)
    obj[0, 2] = 0.1e-3      # dz > 0
    obj[1, :2] = 0.0        # on the axis
    ph = rng.uniform(0, 2 * np.pi, n)
    H = _matlab_wrp(obj, lam, p, res, zw, ph)
    U = O.wrp(res, res, p, lam, zw, obj[:, 0], obj[:, 1], obj[:, 2], mode=3, phase=ph)
    assert rel_l2(U, H.T) < 1e-10
    # the transposition, the phase and the 1/r amplitude each matter
    assert rel_l2(U, H) > 0.1
    assert rel_l2(O.wrp(res, res, p, lam, zw, obj[:, 0], obj[:, 1], obj[:, 2], mode=3), H.T) > 0.1
    # one point: the window is N = 2 floor(|lambda dz / (2 p^2)|) + 1 pixels wide
    one = O.wrp(res, res, p, lam, zw, [0.0], [0.0], [0.0])  # mode 0 window for contrast
    m3 = O.wrp(res, res, p, lam, zw, [0.0], [0.0], [0.0], mode=3)
    w = int(np.floor(lam * zw / (2 * p * p)))
    cols = np.nonzero(np.abs(m3).sum(axis=0))[0]
    assert cols.min() == (res - 1) // 2 - 1 - w and cols.max() == (res - 1) // 2 - 1 + w
    assert np.abs(one).sum() > 0
    # dz = 0 contributes nothing
    assert np.all(O.wrp(res, res, p, lam, zw, [0.0], [0.0], [zw], mode=3) == 0)
