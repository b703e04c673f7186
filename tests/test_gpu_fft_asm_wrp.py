"""GPU parity of the FFT, the padded angular-spectrum propagation and the WRP
accumulation against the FP64 oracle (and numpy's FFT for the transform)."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
from openhologram_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ny,nx", [(16, 64), (32, 32), (256, 128), (1024, 512), (16, 4096), (2048, 4096),
                                   (4096, 4096), (64, 8192)])
def test_fft2d_vs_numpy(ctx, ny, nx):
    print(f"case: fft2d {ny}x{nx}", flush=True)
    rng = np.random.default_rng(ny + nx)
    a = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    f = ctx.fft2d(a, -1)
    ref = np.fft.fft2(a.astype(np.complex128))
    err = rel_l2(f, ref)
    print("forward rel-L2", err)
    assert err < 2e-6
    b = ctx.fft2d(f, +1, normalize=True)
    assert rel_l2(b, a) < 4e-6


def test_fft2d_delta_and_linearity(ctx):
    a = np.zeros((64, 128), np.complex64)
    a[3, 5] = 1.0
    f = ctx.fft2d(a)
    ky, kx = np.mgrid[0:64, 0:128]
    assert np.abs(f - np.exp(-2j * np.pi * (3 * ky / 64 + 5 * kx / 128))).max() < 1e-5


@pytest.mark.parametrize("ny,nx,pitch,lam,z,band", [
    (128, 256, 8e-6, 532e-9, 0.05, False),
    (256, 256, 8e-6, 633e-9, 0.1, True),
    (64, 128, 1e-6, 532e-9, 0.002, False),     # lambda/p large: FP64 transfer path
])
def test_asm_vs_oracle(ctx, ny, nx, pitch, lam, z, band):
    print(f"case: asm {ny}x{nx} p={pitch} z={z} band={band}", flush=True)
    rng = np.random.default_rng(7)
    u = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    out = ctx.asm(u, pitch, lam, z, pad=2, band_limit=band)
    ref = O.asm(u.astype(np.complex128), pitch, pitch, lam, z, pad=2, band_limit=band)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5


def test_asm_pad1_and_errors(ctx):
    rng = np.random.default_rng(8)
    u = (rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))).astype(np.complex64)
    out = ctx.asm(u, 8e-6, 532e-9, 0.02, pad=1)
    assert rel_l2(out, O.asm(u, 8e-6, 8e-6, 532e-9, 0.02, pad=1)) < 1e-5
    import openhologram_amd as oph
    with pytest.raises(oph.OphError):
        ctx.asm(np.zeros((48, 64), np.complex64), 8e-6, 532e-9, 0.02, pad=2)   # 96 not a power of two
    with pytest.raises(oph.OphError):
        ctx.asm(u, 8e-6, 532e-9, 0.02, pad=3)


@pytest.mark.parametrize("ny,nx,pad", [(32, 4096, 1), (16, 2048, 2)])
def test_asm_4096_rows(ctx, ny, nx, pad):
    """4096-point row transforms (first and last stages in registers): a
    field filling the grid (pad 1) and the half-width data window (pad 2)."""
    print(f"case: asm {ny}x{nx} pad {pad}", flush=True)
    rng = np.random.default_rng(nx + pad)
    u = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    out = ctx.asm(u, 8e-6, 532e-9, 0.03, pad=pad)
    ref = O.asm(u.astype(np.complex128), 8e-6, 8e-6, 532e-9, 0.03, pad=pad)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5


@pytest.mark.parametrize("nx,pitch,lam,z,band", [
    (64, 8e-6, 633e-9, 0.1, True),          # band limit
    (32, 1e-6, 532e-9, 0.002, False),       # FP64 transfer
    (32, 1e-6, 532e-9, 0.002, True),        # FP64 transfer and band limit
    (32, 0.35e-6, 532e-9, 0.0005, False),   # evanescent corners (transfer zero)
])
def test_asm_4096_columns(ctx, nx, pitch, lam, z, band):
    """2048-row fields: 4096-point column transforms with the transfer applied
    between the forward and inverse butterflies in registers, in every
    transfer form."""
    print(f"case: asm 2048x{nx} p={pitch} z={z} band={band}", flush=True)
    rng = np.random.default_rng(nx + int(band))
    u = (rng.standard_normal((2048, nx)) + 1j * rng.standard_normal((2048, nx))).astype(np.complex64)
    out = ctx.asm(u, pitch, lam, z, pad=2, band_limit=band)
    ref = O.asm(u.astype(np.complex128), pitch, pitch, lam, z, pad=2, band_limit=band)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5


@pytest.mark.parametrize("band", [False, True])
def test_asm_4096_columns_full_height(ctx, band):
    """4096-row field without padding: 4096-point columns whose data fill the
    whole column (the general register-fused column path, not the
    middle-half one of twofold padding)."""
    print(f"case: asm 4096x16 pad 1 band={band}", flush=True)
    rng = np.random.default_rng(40 + int(band))
    u = (rng.standard_normal((4096, 16)) + 1j * rng.standard_normal((4096, 16))).astype(np.complex64)
    out = ctx.asm(u, 8e-6, 532e-9, 0.05, pad=1, band_limit=band)
    ref = O.asm(u.astype(np.complex128), 8e-6, 8e-6, 532e-9, 0.05, pad=1, band_limit=band)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5


def test_asm_many_rows_per_workgroup(ctx):
    """2048^2 field on the 4096^2 grid: the persistent row passes take
    several rows per workgroup (next row prefetched across the transform)
    and the column pass runs its full grid; against the FP64 oracle."""
    print("case: asm 2048x2048 pad 2", flush=True)
    rng = np.random.default_rng(21)
    u = (rng.standard_normal((2048, 2048)) + 1j * rng.standard_normal((2048, 2048))).astype(np.complex64)
    out = ctx.asm(u, 8e-6, 532e-9, 0.1, pad=2)
    ref = O.asm(u.astype(np.complex128), 8e-6, 8e-6, 532e-9, 0.1, pad=2)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5
    # rows far apart in the grid (the first and last rows of each workgroup's walk)
    for rows in (np.s_[:64], np.s_[-64:], np.s_[1000:1100]):
        assert rel_l2(out[rows], ref[rows]) < 1e-5


def test_wrp_many_points_and_linearity(ctx):
    """300,000 points: binning chunks of two points per thread; against the
    oracle, and the sum of two halves (accumulate) equals the whole."""
    print("case: wrp 128x128 300k points", flush=True)
    x, y, z, a, zw = scenes.wrp_points(300_000, seed=9)
    x, y = x * 0.05, y * 0.05
    out = ctx.wrp(128, 128, 8e-6, 532e-9, zw, x, y, z, a)
    ref = O.wrp(128, 128, 8e-6, 532e-9, zw, x, y, z, a)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5
    h = 150_000
    part = ctx.wrp(128, 128, 8e-6, 532e-9, zw, x[:h], y[:h], z[:h], a[:h])
    both = ctx.wrp(128, 128, 8e-6, 532e-9, zw, x[h:], y[h:], z[h:], a[h:], out=part, accumulate=True)
    assert rel_l2(both, out) < 1e-6
    # ~75,000 list entries per tile: many segments, each ending on a binning
    # chunk boundary, so a second run is bit-identical
    assert np.array_equal(out, ctx.wrp(128, 128, 8e-6, 532e-9, zw, x, y, z, a))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_wrp_vs_oracle(ctx, mode):
    print(f"case: wrp 256x256 mode {mode}", flush=True)
    x, y, z, a, zw = scenes.wrp_points(3000, seed=5)
    x, y = x * 0.1, y * 0.1   # keep most windows inside a 2 mm plane
    out = ctx.wrp(256, 256, 8e-6, 532e-9, zw, x, y, z, a, mode=mode)
    ref = O.wrp(256, 256, 8e-6, 532e-9, zw, x, y, z, a, mode=mode)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5
    # deterministic: a second call is bit-identical
    out2 = ctx.wrp(256, 256, 8e-6, 532e-9, zw, x, y, z, a, mode=mode)
    assert np.array_equal(out, out2)


def test_wrp_large_plane(ctx):
    """4096 x 2048 plane (2048 tiles, more than four per binning thread):
    the tile-offset scan's general path, against the oracle, deterministic."""
    print("case: wrp 4096x2048 2000 points", flush=True)
    x, y, z, a, zw = scenes.wrp_points(2000, seed=12)
    y = y * 0.5
    out = ctx.wrp(4096, 2048, 8e-6, 532e-9, zw, x, y, z, a)
    ref = O.wrp(4096, 2048, 8e-6, 532e-9, zw, x, y, z, a)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5
    assert np.array_equal(out, ctx.wrp(4096, 2048, 8e-6, 532e-9, zw, x, y, z, a))


def test_wrp_edge_cases(ctx):
    # empty set, a point outside the plane, unit amplitude, accumulate, ragged size
    e = np.zeros(0)
    assert np.all(ctx.wrp(64, 64, 8e-6, 532e-9, 0.0, e, e, e) == 0)
    x, y, z = np.array([1.0, 2e-5]), np.array([1.0, -3e-5]), np.array([-0.003, 0.002])
    out = ctx.wrp(100, 70, 8e-6, 532e-9, 0.0, x, y, z)
    ref = O.wrp(100, 70, 8e-6, 532e-9, 0.0, x, y, z)
    assert rel_l2(out, ref) < 1e-5
    acc = ctx.wrp(100, 70, 8e-6, 532e-9, 0.0, x, y, z, out=out.copy(), accumulate=True)
    assert rel_l2(acc, 2 * ref) < 1e-5


def test_wrp_hologram_small(ctx):
    """WRP + padded ASM at reduced size against the oracle chain."""
    print("case: wrp_hologram 512 P=5000", flush=True)
    x, y, z, a, zw = scenes.wrp_points(5000, seed=6)
    x, y = x * 0.2, y * 0.2
    d = -zw
    out = ctx.wrp_hologram(512, 8e-6, 532e-9, zw, d, x, y, z, a, pad=2)
    w = O.wrp(512, 512, 8e-6, 532e-9, zw, x, y, z, a)
    ref = O.asm(w, 8e-6, 8e-6, 532e-9, d, pad=2)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5


def test_wrp_hologram_100k_2048sq_full(ctx):
    """The full WRP workload: 100,000 points, 2048^2 WRP, 100 mm ASM on a
    4096^2 padded grid, complex64 out; rel-L2 vs the FP64 oracle <= 1e-5
    (the plain path's bar)."""
    print("case: wrp_hologram 100k points 2048^2 full", flush=True)
    x, y, z, a, zw = scenes.wrp_points(100_000, seed=2)
    out = ctx.wrp_hologram(2048, 8e-6, 532e-9, zw, -zw, x, y, z, a, pad=2)
    w = O.wrp(2048, 2048, 8e-6, 532e-9, zw, x, y, z, a)
    ref = O.asm(w, 8e-6, 8e-6, 532e-9, -zw, pad=2)
    err = rel_l2(out, ref)
    print("rel-L2 wrp_hologram 100k", err)
    assert err < 1e-5


def _mode1_tiles(nx, ny, p, lam, zw, x, y, z):
    """List entries the binning needs for mode-1 windows (tiles of 64)."""
    dz = zw - z
    w = (np.abs(lam * dz / p / p / 2 + 0.5)).astype(np.int64) * 2 - 1
    tx = np.trunc(x / p).astype(np.int64) + nx // 2
    ty = np.trunc(y / p).astype(np.int64) + ny // 2
    def span(c, n):
        lo, hi = np.maximum(c - w, 0), np.minimum(c + w - 1, n - 1)
        return np.where(lo <= hi, hi // 64 - lo // 64 + 1, 0)
    return int(np.sum(span(tx, nx) * span(ty, ny)))


def test_wrp_overflow_retry_with_accumulate(ctx):
    """Mode-1 windows at |dz| ~ 31 mm (half-width ~ 256 px) need more list
    entries than the first capacity (24 per point + 65536): the binning is
    redone at the exact size, and an accumulating call adds the whole plane
    exactly once."""
    print("case: wrp overflow retry 1024x1024 mode 1", flush=True)
    rng = np.random.default_rng(31)
    n = 1500
    x, y = rng.uniform(-1.5e-3, 1.5e-3, (2, n))
    z = rng.uniform(-0.0315, -0.0305, n)
    a = rng.uniform(0.2, 1.0, n)
    assert _mode1_tiles(1024, 1024, 8e-6, 532e-9, 0.0, x, y, z) > 24 * n + 65536
    base = (rng.standard_normal((1024, 1024)) + 1j * rng.standard_normal((1024, 1024))).astype(np.complex64)
    # a fresh context: its list buffer starts at the first capacity
    from openhologram_amd import Context
    with Context(0) as c2:
        out = c2.wrp(1024, 1024, 8e-6, 532e-9, 0.0, x, y, z, a, mode=1, out=base.copy(), accumulate=True)
    ref = O.wrp(1024, 1024, 8e-6, 532e-9, 0.0, x, y, z, a, mode=1)
    err = rel_l2(out - base, ref)
    print("rel-L2", err)
    assert err < 1e-5


def test_wrp_async_flags(ctx):
    """OPH_ASYNC needs device pointers; an asynchronous overflow leaves the
    output unwritten and is reported once by oph_sync; an asynchronous WRP
    followed by other asynchronous work and a sync reports nothing."""
    from openhologram_amd import OphError
    x = np.array([1e-5, -2e-5])
    y = np.array([0.0, 3e-5])
    z = np.array([-0.002, -0.003])
    with pytest.raises(OphError):
        ctx.wrp(64, 64, 8e-6, 532e-9, 0.0, x, y, z, asynchronous=True)
    dx, dy, dz = ctx.to_device(x), ctx.to_device(y), ctx.to_device(z)
    out = ctx.empty((64, 64), np.complex64)
    ctx.wrp(64, 64, 8e-6, 532e-9, 0.0, dx, dy, dz, out=out, device=True, asynchronous=True)
    fr = ctx.to_device(np.random.default_rng(1).standard_normal((64, 64)).astype(np.float32))
    u8 = ctx.empty((64, 64), np.uint8)
    ctx.normalize_u8(fr, device=True, dst=u8)
    ctx.sync()
    assert rel_l2(out.numpy(), O.wrp(64, 64, 8e-6, 532e-9, 0.0, x, y, z)) < 1e-5
    # overflow: the windows of the retry test, asynchronously, in a fresh
    # context (its list buffer starts at the first capacity)
    print("case: wrp async overflow 1024x1024 mode 1", flush=True)
    from openhologram_amd import Context
    rng = np.random.default_rng(31)
    n = 1500
    px, py = rng.uniform(-1.5e-3, 1.5e-3, (2, n))
    pz = rng.uniform(-0.0315, -0.0305, n)
    assert _mode1_tiles(1024, 1024, 8e-6, 532e-9, 0.0, px, py, pz) > 24 * n + 65536
    with Context(0) as c2:
        big = c2.to_device(np.full((1024, 1024), 3 + 4j, np.complex64))
        c2.wrp(1024, 1024, 8e-6, 532e-9, 0.0, c2.to_device(px), c2.to_device(py), c2.to_device(pz), mode=1,
               out=big, device=True, asynchronous=True, accumulate=True)
        with pytest.raises(OphError) as e:
            c2.sync()
        assert "overflow" in str(e.value)
        assert np.all(big.numpy() == np.complex64(3 + 4j))
        c2.sync()   # reported once
        del big


def test_wrp_reference_table(ctx):
    """oph_wrp mode 1 against the reference example's own output,
    W/openwrp/wrp_table.bmp (SingleWRP(0.5 mm) of object.txt, 515^2 at
    8 um, 532 nm), read as an arg image; the bar and its basis are those of
    the oracle's pin (tests/test_oracle_fft_asm_wrp.py)."""
    import os
    from conftest import GOLDEN, wrp_table_agreement
    print("case: wrp reference table 515x515", flush=True)
    g = np.load(os.path.join(GOLDEN, "wrp_table.npz"), allow_pickle=False)
    xyz = g["xyz"].astype(np.float64)
    lam, p, n, zw = (float(v) for v in g["params"])
    U = ctx.wrp(int(n), int(n), p, lam, zw, xyz[:, 0], xyz[:, 1], xyz[:, 2], mode=1)
    exact, near = wrp_table_agreement(U, g["gray"])
    print("agreement exact", exact, "within 1", near)
    assert exact >= 0.99 and near >= 0.9995
    ref = O.wrp(int(n), int(n), p, lam, zw, xyz[:, 0], xyz[:, 1], xyz[:, 2], mode=1)
    assert rel_l2(U, ref) < 1e-5


def test_wrp_mode2_fp64_path(ctx):
    """Mode 2 (cwoPLS::WRP) with windows whose quadratic phase passes 16
    turns (|dz| ~ 40 mm): the per-pixel FP64 path, against the oracle."""
    print("case: wrp mode 2 large dz 512x512", flush=True)
    rng = np.random.default_rng(41)
    n = 200
    x, y = rng.uniform(-1e-3, 1e-3, (2, n))
    z = rng.uniform(-0.045, -0.035, n)
    a = rng.uniform(0.2, 1.0, n)
    out = ctx.wrp(512, 512, 8e-6, 532e-9, 0.0, x, y, z, a, mode=2)
    ref = O.wrp(512, 512, 8e-6, 532e-9, 0.0, x, y, z, a, mode=2)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5


@pytest.mark.parametrize("res,pitch,lam", [
    (1025, 3.9e-6, 532e-9),   # the MATLAB script's plane: FP64 per-pixel path (corner u > 2e-3)
    (1001, 8e-6, 473e-9),     # small corner angle: the series path with amplitude 1/r
])
def test_wrp_mode3_matlab_vs_oracle(ctx, res, pitch, lam):
    """Mode 3 (the MATLAB statement: N x N window, 1/r, per-point phase,
    rounded centre on an odd plane) against the oracle, which is pinned by
    a transcription of the script; both signs of dz; deterministic."""
    print(f"case: wrp mode 3 {res} p={pitch} lam={lam}", flush=True)
    rng = np.random.default_rng(17)
    n = 2000
    zw = 0.5e-3
    x, y = rng.uniform(-400, 400, (2, n)) * pitch
    z = rng.uniform(-1.0e-3, 0.3e-3, n)
    a = rng.uniform(0.2, 1.0, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    out = ctx.wrp(res, res, pitch, lam, zw, x, y, z, a, mode=3, phase=ph)
    ref = O.wrp(res, res, pitch, lam, zw, x, y, z, a, mode=3, phase=ph)
    err = rel_l2(out, ref)
    print("rel-L2", err)
    assert err < 1e-5
    assert np.array_equal(out, ctx.wrp(res, res, pitch, lam, zw, x, y, z, a, mode=3, phase=ph))
    # the phase is read by mode 3 only
    assert np.array_equal(ctx.wrp(128, 128, pitch, lam, zw, x * 0.1, y * 0.1, z, a, mode=1, phase=ph),
                          ctx.wrp(128, 128, pitch, lam, zw, x * 0.1, y * 0.1, z, a, mode=1))
