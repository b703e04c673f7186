"""GPU parity of the exact RS point-cloud field (oph_pc_rs) against the FP64
oracle (tolerance rel-L2 1e-5 on the complex field) and the reference's
FP32 fringe fixtures."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
import openhologram_amd as oph
from openhologram_amd import scenes

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _s64(spec):
    return dict(nx=spec["nx"], ny=spec["ny"], pitch=spec["pitch"], lambda_=spec["lambda_"],
                scale=spec.get("scale", (1, 1, 1)), offset_z=spec.get("offset_z", 0.0),
                tilt_deg=spec.get("tilt_deg", (0, 0)))


def test_rs_1000pts_1024sq_full_size(ctx):
    """1000 seeded points, 1024^2 at 8 um, 532 nm, complex128 out."""
    print("case: rs 1024x1024 P=1000", flush=True)
    x, y, z, a, ph = scenes.rs_points(1000, seed=1)
    spec = dict(scenes.RS_SPEC)
    U = ctx.pc_rs(spec, x, y, z, a, ph)
    ref = O.rs(_s64(spec), x, y, z, a, ph)
    err = rel_l2(U, ref)
    print("rel-L2", err)
    assert err < TOL


@pytest.mark.parametrize("nx,ny,tilt", [(100, 37, (0, 0)), (64, 64, (22.5, 22.5)), (130, 70, (3, -7))])
def test_rs_ragged_and_tilt(ctx, nx, ny, tilt):
    print(f"case: rs ragged {nx}x{ny} tilt {tilt}", flush=True)
    x, y, z, a, ph = scenes.rs_points(300, seed=7)
    spec = dict(nx=nx, ny=ny, pitch=(8e-6, 6e-6), lambda_=633e-9, tilt_deg=tilt)
    U = ctx.pc_rs(spec, x, y, z, a, ph)
    assert rel_l2(U, O.rs(_s64(spec), x, y, z, a, ph)) < TOL


def test_rs_near_field_exact_path(ctx):
    """Points 0.5-2 mm from the plane with wide lateral spread take the FP64
    per-pixel path (second-order expansion out of range)."""
    print("case: rs near field", flush=True)
    g = scenes.SplitMix64(11)
    n = 200
    x, y = g.uniform(n, -2e-3, 2e-3), g.uniform(n, -2e-3, 2e-3)
    z = g.uniform(n, 0.5e-3, 2e-3)
    a = g.uniform(n)
    spec = dict(nx=256, ny=192, pitch=8e-6, lambda_=532e-9)
    U = ctx.pc_rs(spec, x, y, z, a)
    assert rel_l2(U, O.rs(_s64(spec), x, y, z, a)) < TOL


def test_rs_keti_fringe_fixtures(ctx, golden_keti):
    """Float fringe mode vs the reference fixtures (the reference's own FP32
    noise bounds the match) and vs the FP64 oracle (tight)."""
    pts = golden_keti["points"]
    for name in golden_keti["case_names"]:
        print("case: keti fixture", name, flush=True)
        s = golden_keti[f"spec_{name}"]
        spec = dict(nx=int(s[6]), ny=int(s[7]), pitch=(float(s[4]), float(s[5])),
                    lambda_=float(s[8]), scale=tuple(float(v) for v in s[0:3]),
                    offset_z=float(s[3]), tilt_deg=(float(s[9]), float(s[10])))
        fr = ctx.pc_rs(spec, pts[:, 0], pts[:, 1], pts[:, 2], pts[:, 3], fmt=oph.OUT_FRINGE)
        assert fr.dtype == np.float32
        ref64 = O.rs(_s64(spec), pts[:, 0], pts[:, 1], pts[:, 2], pts[:, 3]).real
        assert rel_l2(fr, ref64) < 1e-5
        assert rel_l2(fr, golden_keti[f"fringe_{name}"]) < 0.1


def test_rs_accumulate_device_and_c64(ctx):
    print("case: rs accumulate / device / c64", flush=True)
    x, y, z, a, ph = scenes.rs_points(500, seed=3)
    spec = dict(nx=128, ny=96, pitch=8e-6, lambda_=532e-9)
    ref = O.rs(_s64(spec), x, y, z, a, ph)
    out = ctx.pc_rs(spec, x, y, z, a, ph)
    ctx.pc_rs(spec, x, y, z, a, ph, out=out, accumulate=True)
    assert rel_l2(out, 2 * ref) < TOL
    d = [ctx.to_device(v) for v in (x, y, z, a, ph)]
    dout = ctx.pc_rs(spec, *d, fmt=oph.OUT_C64, device=True)
    U = dout.numpy()
    assert U.dtype == np.complex64 and rel_l2(U, ref) < 2e-5


def test_rs_empty_point_set(ctx):
    spec = dict(nx=64, ny=64, pitch=8e-6, lambda_=532e-9)
    e = np.zeros(0)
    U = ctx.pc_rs(spec, e, e, e, e)
    assert np.all(U == 0)


def test_rs_invalid_args(ctx):
    with pytest.raises(oph.OphError):
        ctx.pc_rs(dict(nx=0, ny=64, pitch=8e-6, lambda_=532e-9), [0.0], [0.0], [0.1], [1.0])
    with pytest.raises(oph.OphError):
        ctx.pc_rs(dict(nx=64, ny=64, pitch=8e-6, lambda_=-1.0), [0.0], [0.0], [0.1], [1.0])


@pytest.mark.parametrize("nx,ny,z", [(256, 192, 0.1), (130, 70, 0.02)])
def test_huygens_vs_oracle(ctx, nx, ny, z):
    """cwoPLS::Huygens (1/r amplitude, CWO grid) against the FP64 oracle,
    including points close enough for the per-pixel FP64 path."""
    print(f"case: huygens {nx}x{ny} z {z}", flush=True)
    rng = np.random.default_rng(nx)
    n = 700
    x, y = rng.uniform(-5e-4, 5e-4, (2, n))
    zz = rng.uniform(-0.005, 0.005, n)
    a, ph = rng.uniform(0, 1, n), rng.uniform(0, 2 * np.pi, n)
    U = ctx.huygens(nx, ny, 8e-6, 532e-9, z, x, y, zz, a, ph, ph=0.25)
    ref = O.huygens(nx, ny, 8e-6, 532e-9, z, x, y, zz, a, ph, ph=0.25)
    err = rel_l2(U, ref)
    print("rel-L2", err)
    assert err < 1e-5
    V = ctx.huygens(nx, ny, 8e-6, 532e-9, z, x, y, zz, a, ph, ph=0.25, fmt=oph.OUT_C64, out=U.astype(np.complex64),
                    accumulate=True)
    assert rel_l2(V, 2 * ref) < 1e-5
