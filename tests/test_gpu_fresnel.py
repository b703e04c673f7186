"""GPU parity of the Fresnel point-cloud field: the plain VALU kernel (rel-L2
1e-5) and the MFMA contraction per factor precision against the FP64 oracle."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
from openhologram_amd import scenes

pytestmark = pytest.mark.gpu
# rel-L2 bounds per method: rounding of both chirp factors to 8 (BF16) or 11
# (FP16) significant bits, or to ~16 / ~22 with the hi + lo split
BOUND = {"valu": 1e-5, "bf16": 6e-3, "bf16_split": 1e-4, "fp16": 1e-3, "fp16_split": 1e-5}


@pytest.mark.parametrize("method", list(BOUND))
def test_fresnel_small_vs_oracle(ctx, method):
    print(f"case: fresnel 200x136 P=3000 {method}", flush=True)
    x, y, p, z, a, ph = scenes.fresnel_points(3000, planes=8, seed=7, half=0.4e-3)
    U = ctx.fresnel(200, 136, 8e-6, 532e-9, x, y, z, a, ph, method=method)
    ref = O.fresnel(200, 136, 8e-6, 532e-9, x, y, z, a, ph)
    err = rel_l2(U, ref)
    print(method, "rel-L2", err)
    assert err < BOUND[method]


def _ragged_planes(seed):
    """4000 points on 5 planes with ragged counts (3000, 600, 366, 33, 1): each
    plane's last 32-point chunk is partly padding, one plane is a single point."""
    rng = np.random.default_rng(seed)
    counts = [3000, 600, 366, 33, 1]
    zp = [0.1, 0.104, 0.1013, 0.11, 0.1077]
    z = np.concatenate([np.full(c, zz) for c, zz in zip(counts, zp)])
    rng.shuffle(z)
    n = len(z)
    x = rng.uniform(-0.4e-3, 0.4e-3, n)
    y = rng.uniform(-0.4e-3, 0.4e-3, n)
    a = rng.uniform(0, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    return x, y, z, a, ph


@pytest.mark.parametrize("method", ["fp16", "fp16_split", "bf16", "bf16_split"])
def test_fresnel_plane_form_vs_general(ctx, method, monkeypatch):
    """Points on a few planes take the plane-factored contraction (grouping on
    the device, per-plane chirps applied at each plane's end); the general
    contraction (OPH_FRESNEL_PLANES=0) and the FP64 oracle agree with it to
    the method's bound, on a 300 x 200 field (partial tiles on both axes)."""
    print(f"case: fresnel ragged planes 300x200 {method}", flush=True)
    x, y, z, a, ph = _ragged_planes(11)
    ref = O.fresnel(300, 200, 8e-6, 532e-9, x, y, z, a, ph)
    U = ctx.fresnel(300, 200, 8e-6, 532e-9, x, y, z, a, ph, method=method)
    U2 = ctx.fresnel(300, 200, 8e-6, 532e-9, x, y, z, a, ph, method=method)
    assert np.array_equal(U, U2)  # deterministic
    monkeypatch.setenv("OPH_FRESNEL_PLANES", "0")
    G = ctx.fresnel(300, 200, 8e-6, 532e-9, x, y, z, a, ph, method=method)
    monkeypatch.delenv("OPH_FRESNEL_PLANES")
    e1, e2 = rel_l2(U, ref), rel_l2(G, ref)
    print(method, "planes rel-L2", e1, "general rel-L2", e2)
    assert e1 < BOUND[method] and e2 < BOUND[method]


def test_fresnel_distinct_depths_general_form(ctx):
    """Every point on its own depth: the device keeps the general contraction
    (too few points per plane for the factored one); parity as usual."""
    print("case: fresnel distinct depths 160x96", flush=True)
    rng = np.random.default_rng(5)
    n = 3000
    x = rng.uniform(-0.3e-3, 0.3e-3, n)
    y = rng.uniform(-0.3e-3, 0.3e-3, n)
    z = rng.uniform(0.1, 0.12, n)
    a = rng.uniform(0, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    ref = O.fresnel(160, 96, 8e-6, 532e-9, x, y, z, a, ph)
    for m in ("fp16", "fp16_split"):
        err = rel_l2(ctx.fresnel(160, 96, 8e-6, 532e-9, x, y, z, a, ph, method=m), ref)
        print(m, "rel-L2", err)
        assert err < BOUND[m]


def test_fresnel_edge_cases(ctx):
    e = np.zeros(0)
    for m in ("valu", "fp16_split"):
        assert np.all(ctx.fresnel(64, 64, 8e-6, 532e-9, e, e, e, method=m) == 0)
    # a single point: |U| = amplitude everywhere
    U = ctx.fresnel(130, 70, 8e-6, 532e-9, [1e-5], [-2e-5], [0.1], [0.7], method="fp16_split")
    assert np.allclose(np.abs(U), 0.7, atol=1e-4)


# Bars for 1e6 points on a 4096^2 plane: the plain path at 1e-5; the MFMA
# forms at the 1e-3 contraction target (FP16) or their factor-rounding level.  The points
# lie on 64 planes, so the plane-factored form runs: its per-plane partial
# sums keep the FP32 accumulation error of the 1e6-term sum small (the split
# forms measure ~2e-6..4e-6; BF16 unsplit ~2e-3).  The general form on the
# same points accumulates all 1e6 terms in one FP32 sum (split forms
# ~6e-5..8e-5): FR1E6_GENERAL_BOUND.
FR1E6_BOUND = {"valu": 1e-5, "bf16": 6e-3, "bf16_split": 2e-5, "fp16": 1e-3, "fp16_split": 1e-5}
FR1E6_GENERAL_BOUND = {"fp16": 1e-3, "fp16_split": 2e-4}


def test_fresnel_1e6pts_4096sq_sampled(ctx):
    """1,000,000 points in 64 planes, 4096^2 at 8 um: the VALU kernel and
    every MFMA precision against the FP64 oracle on 4096 sampled pixels."""
    print("case: fresnel 1e6 points 4096^2 sampled", flush=True)
    x, y, p, z, a, ph = scenes.fresnel_points(1_000_000, planes=64, seed=4)
    rng = np.random.default_rng(9)
    rows, cols = rng.integers(0, 4096, (2, 4096))
    ref = O.fresnel(4096, 4096, 8e-6, 532e-9, x, y, z, a, ph, rows=rows, cols=cols)
    d = [ctx.to_device(v) for v in (x, y, z, a, ph)]
    out = ctx.empty((4096, 4096), np.complex64)
    for m in ("valu", "bf16", "bf16_split", "fp16", "fp16_split"):
        ctx.fresnel(4096, 4096, 8e-6, 532e-9, *d, method=m, out=out, device=True)
        U = out.numpy()[rows, cols]
        err = rel_l2(U, ref)
        print(m, "1e6 points sampled rel-L2", err, flush=True)
        assert err < FR1E6_BOUND[m]
    # the general contraction on the same points (the plane-factored form is
    # the default here: 64 planes of ~15,600 points)
    import os
    os.environ["OPH_FRESNEL_PLANES"] = "0"
    try:
        for m in ("fp16", "fp16_split"):
            ctx.fresnel(4096, 4096, 8e-6, 532e-9, *d, method=m, out=out, device=True)
            err = rel_l2(out.numpy()[rows, cols], ref)
            print(m, "general form 1e6 points sampled rel-L2", err, flush=True)
            assert err < FR1E6_GENERAL_BOUND[m]
    finally:
        del os.environ["OPH_FRESNEL_PLANES"]
