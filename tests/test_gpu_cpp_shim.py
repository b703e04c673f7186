"""GPU: the reference-shaped C++ header driven from C++ (tests/cpp/
shim_check, built by build()): KETI fringe, WRP object, depth-map generator
object end to end from a config file and BMP inputs, compared with the same
library called from Python and with the reference fixture."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
import openhologram_amd as oph
from openhologram_amd import lib as L, scenes

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "cpp", "shim_check")


def _config(tmp):
    src = open(os.path.join(GOLDEN, "dm_config_openholo.txt")).read()
    for a, b in [("source\t", f"{tmp}/src\t"), ("result\t", f"{tmp}/result\t"), ("3840*3\t", "128\t"),
                 ("2160\t\t\t\t\t// SLM_PIXEL_NUMBER_Y", "64\t// SLM_PIXEL_NUMBER_Y"),
                 ("(25.4e-3)/2250*(1/3)", "8e-6"), ("(25.4e-3)/2250\t", "8e-6\t"),
                 ("8\t\t\t\t\t\t// VAR_INFO.step_num", "1\t// VAR_INFO.step_num")]:
        assert a in src, a
        src = src.replace(a, b, 1)
    open(os.path.join(tmp, "dm_config.txt"), "w").write(src)


def test_cpp_shim_generators(ctx, tmp_path, golden_keti):
    print("case: C++ shim generators", flush=True)
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "src"))
    _config(tmp)
    img, dep = scenes.depth_scene(128, 64, seed=11, colors=3)
    rgb = np.stack([img[0], img[1], img[2]], axis=-1)
    L.save_bmp24(os.path.join(tmp, "src", "RGB_D0.bmp"), rgb)
    L.save_bmp8(os.path.join(tmp, "src", "D_D0.bmp"), dep[::-1])   # bottom-up storage
    r = subprocess.run([EXE, "gpu", GOLDEN, tmp], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = dict(l.partition(" ")[::2] for l in r.stdout.splitlines())
    assert [l.split()[2:] for l in r.stdout.splitlines() if l.startswith("wrp ")] == [["0", "0"], ["0", "0"]]

    # KETI fringe: identical to the library called from Python on the same
    # float inputs, and within FP32 noise of the reference's own fringe
    fr = np.fromfile(os.path.join(tmp, "fringe.f32"), np.float32).reshape(64, 96)
    pts = golden_keti["points"].astype(np.float64)
    s = golden_keti["spec_example_96x64"].astype(np.float64)
    spec = dict(nx=96, ny=64, pitch=(s[4], s[5]), lambda_=s[8], scale=tuple(s[0:3]), offset_z=s[3],
                tilt_deg=(s[9], s[10]))
    py = ctx.pc_rs(spec, pts[:, 0], pts[:, 1], pts[:, 2], pts[:, 3], fmt=oph.OUT_FRINGE)
    assert np.array_equal(fr, py)
    assert rel_l2(fr, golden_keti["fringe_example_96x64"]) < 0.1
    assert float(out["gpu_entry_maxdiff"]) == 0.0
    u8 = np.fromfile(os.path.join(tmp, "fringe.u8"), np.uint8).reshape(64, 96)
    assert np.array_equal(u8, ctx.normalize_u8(fr))

    # WRP object: plane, propagation (power-of-two side and embedded side)
    xyz = np.array([[1e-4, 2e-4, -1e-3], [-1e-4, 0, -1.5e-3]])
    for n in (64, 50):
        w = np.fromfile(os.path.join(tmp, f"wrp{n}.c64"), np.complex64).reshape(n, n)
        want = ctx.wrp(n, n, np.float32(8e-6), np.float32(532e-9), 0.0, xyz[:, 0], xyz[:, 1], xyz[:, 2], mode=1)
        assert np.array_equal(w, want)
        p = np.fromfile(os.path.join(tmp, f"prop{n}.c64"), np.complex64).reshape(n, n)
        m = 64
        big = np.zeros((m, m), np.complex64)
        o = m // 2 - n // 2
        big[o:o + n, o:o + n] = want
        big = ctx.asm(big, np.float32(8e-6), np.float32(532e-9), np.float32(0.01), pad=2)
        assert np.array_equal(p, big[o:o + n, o:o + n])

    # depth-map generator object: fringe of the config, the result image,
    # the reconstruction of the configured focus step
    cfg = L.dm_config_load(os.path.join(tmp, "dm_config.txt"))
    gray = L.load_bmp_gray8(os.path.join(tmp, "src", "RGB_D0.bmp"))
    f = np.fromfile(os.path.join(tmp, "dm_fringe.f64")).reshape(64, 128)
    want = ctx.depthmap(128, 64, 128, 64, 8e-6, cfg.near_depthmap, cfg.far_depthmap, 256, [cfg.lambda_],
                        gray[None], dep, quant_mode=1, field_lens=cfg.field_lens)[0]
    assert np.array_equal(f, want)
    res = L.load_bmp_gray8(os.path.join(tmp, "result", "2250ppi_3840_UD_flat0.bmp"))
    assert res.shape == (64, 128 // 3)
    sim = np.fromfile(os.path.join(tmp, "dm_sim.f64")).reshape(64, 128)
    focus = (cfg.sim_from + cfg.sim_to) / 2
    ref = ctx.dm_reconstruct(f, 8e-6, cfg.lambda_, cfg.field_lens, cfg.eye_length, cfg.eye_pupil_diameter,
                             (cfg.eye_center_xy[0], cfg.eye_center_xy[1]), focus)
    assert np.array_equal(sim, ref)
    assert out["dm_error"] == "''"
