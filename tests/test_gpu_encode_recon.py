"""GPU: 8-bit encoders bit-exact against the oracle (and, for ophNormalize,
against the reference's own output), and the eye-model reconstruction
against the FP64 oracle."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ny,nx", [(1, 7), (13, 17), (768, 1024), (2160, 4096)])
def test_normalize_u8_bit_exact(ctx, ny, nx):
    print(f"case: normalize {ny}x{nx}", flush=True)
    rng = np.random.default_rng(nx + ny)
    src = (rng.normal(size=(ny, nx)) * rng.uniform(1e-3, 1e3)).astype(np.float32)
    assert np.array_equal(ctx.normalize_u8(src), O.normalize_u8(src))


def test_normalize_u8_reference_fixture(ctx, golden_keti):
    """The committed fixture holds the reference ophNormalize output of the
    reference fringe."""
    print("case: normalize KETI fixture", flush=True)
    name = golden_keti["case_names"][0]
    src = golden_keti[f"fringe_{name}"]
    assert np.array_equal(ctx.normalize_u8(src), golden_keti["normalized_u8"])


def test_encoders_constant_input(ctx):
    print("case: constant input", flush=True)
    assert not ctx.normalize_u8(np.full((4, 5), 2.5, np.float32)).any()
    assert not ctx.scale_u8(np.full(9, -1.0)).any()


@pytest.mark.parametrize("n", [1, 255, 257, 1_000_003])
def test_scale_u8_bit_exact(ctx, n):
    print(f"case: scale n={n}", flush=True)
    rng = np.random.default_rng(n)
    src = rng.normal(size=n) * 1e4
    if n == 1:
        src = np.array([3.0])
    assert np.array_equal(ctx.scale_u8(src), O.scale_u8(src))


def test_scale_u8_depthmap_fringe(ctx):
    print("case: scale of a depth-map fringe", flush=True)
    from openhologram_amd import scenes
    img, dep = scenes.depth_scene(256, 128, seed=5, colors=1)
    fr = ctx.depthmap(256, 128, 512, 256, 8e-6, 0.09, 0.11, 256, [532e-9], img, dep)[0]
    assert np.array_equal(ctx.scale_u8(fr), O.scale_u8(fr))


@pytest.mark.parametrize("focus,center,pupil", [(5e-3, (3, -2), 1.0), (0.0, (0, 0), 1.0),
                                                (2e-3, (-5, 1), 2.5e-4)])
def test_reconstruct_vs_oracle(ctx, focus, center, pupil):
    """Depth-map fringe 256 x 128 through the eye model; the eye parameters
    put the anti-aliasing and pupil masks inside the pupil plane."""
    print(f"case: reconstruct focus={focus} center={center} pupil={pupil}", flush=True)
    from openhologram_amd import scenes
    nx, ny, p, lam = 256, 128, 8e-6, 532e-9
    img, dep = scenes.depth_scene(nx, ny, seed=6, colors=1)
    fr = ctx.depthmap(nx, ny, 2 * nx, 2 * ny, p, 0.09, 0.11, 256, [lam], img, dep)[0]
    f = 0.25 * nx * p * p / lam
    pex, pey = lam * f / (nx * p), lam * f / (ny * p)
    c = (center[0] * pex, center[1] * pey)
    ours = ctx.dm_reconstruct(fr, p, lam, f, 50e-3, pupil, c, focus)
    ref = O.dm_reconstruct(fr, p, p, lam, f, 50e-3, pupil, c[0], c[1], focus, nthreads=8)
    e = rel_l2(ours, ref)
    print("rel-L2", e)
    assert ref.max() > 0 and e < 1e-5


def test_reconstruct_rejects_bad_sizes(ctx):
    from openhologram_amd.lib import OphError
    with pytest.raises(OphError):
        ctx.dm_reconstruct(np.zeros((100, 128)), 8e-6, 532e-9, 0.1, 0.05, 0.006)
