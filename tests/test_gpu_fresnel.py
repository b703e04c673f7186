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


def test_fresnel_edge_cases(ctx):
    e = np.zeros(0)
    for m in ("valu", "fp16_split"):
        assert np.all(ctx.fresnel(64, 64, 8e-6, 532e-9, e, e, e, method=m) == 0)
    # a single point: |U| = amplitude everywhere
    U = ctx.fresnel(130, 70, 8e-6, 532e-9, [1e-5], [-2e-5], [0.1], [0.7], method="fp16_split")
    assert np.allclose(np.abs(U), 0.7, atol=1e-4)


# config-4 bars: the plain path at the north-star 1e-5; the MFMA forms at the
# 1e-3 contraction target (FP16) or their factor-rounding level (the split
# forms measure ~6e-5..8e-5 over 1e6 points; BF16 unsplit ~1.9e-3)
CONFIG4_BOUND = {"valu": 1e-5, "bf16": 6e-3, "bf16_split": 2e-4, "fp16": 1e-3, "fp16_split": 2e-4}


def test_fresnel_config4_sampled(ctx):
    """1,000,000 points in 64 planes, 4096^2 at 8 um: the VALU kernel and
    every MFMA precision against the FP64 oracle on 4096 sampled pixels."""
    print("case: fresnel config4 sampled", flush=True)
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
        print(m, "config4 sampled rel-L2", err, flush=True)
        assert err < CONFIG4_BOUND[m]
