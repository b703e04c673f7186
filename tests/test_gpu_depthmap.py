"""GPU parity of the depth-map layer hologram (oph_depthmap) against the FP64
oracle restatement of the reference's CPU path."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
from openhologram_amd import scenes

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _oracle_rgb(img, dep, Nx, Ny, p, lams, near, far, L, out="fringe", **kw):
    res = []
    for c, lam in enumerate(lams):
        r = O.depthmap(img[c], dep, Nx, Ny, p, p, lam, near, far, L, want_fringe=(out != "field"),
                       want_field=(out == "field"), **kw)
        res.append(r[2] if out == "field" else r[1])
    return np.stack(res)


@pytest.mark.parametrize("w,h,Nx,Ny,L,mode", [
    (32, 16, 32, 16, 256, 0),      # the reference's unpadded case
    (40, 24, 64, 32, 16, 1),       # padded, requantised to 16 levels
    (100, 60, 256, 128, 64, 1),
])
def test_depthmap_rgb_vs_oracle(ctx, w, h, Nx, Ny, L, mode):
    print(f"case: depthmap {w}x{h} grid {Nx}x{Ny} L={L} mode={mode}", flush=True)
    img, dep = scenes.depth_scene(w, h, seed=3)
    lams = (638e-9, 532e-9, 473e-9)
    fr = ctx.depthmap(w, h, Nx, Ny, 8e-6, 0.09, 0.11, L, lams, img, dep, quant_mode=mode)
    ref = _oracle_rgb(img, dep, Nx, Ny, 8e-6, lams, 0.09, 0.11, L, quant_mode=mode)
    for c in range(3):
        err = rel_l2(fr[c], ref[c])
        print("colour", c, "rel-L2", err)
        assert err < TOL


def test_depthmap_field_fringe32_lens_and_phase(ctx):
    """Complex hologram-field output, float fringe, a field lens and a
    random-phase factor against the oracle."""
    print("case: depthmap field/lens/rho", flush=True)
    img, dep = scenes.depth_scene(48, 32, seed=5, colors=1)
    kw = dict(quant_mode=1, field_lens=1.0, rho=(np.cos(0.7), np.sin(0.7)))
    u = ctx.depthmap(48, 32, 64, 64, 8e-6, 0.2, 0.7, 32, [522e-9], img, dep, out="field", **kw)
    ref = _oracle_rgb(img, dep, 64, 64, 8e-6, [522e-9], 0.2, 0.7, 32, out="field", **kw)
    assert rel_l2(u, ref) < TOL
    f32 = ctx.depthmap(48, 32, 64, 64, 8e-6, 0.2, 0.7, 32, [522e-9], img, dep, out="fringe32", **kw)
    reff = _oracle_rgb(img, dep, 64, 64, 8e-6, [522e-9], 0.2, 0.7, 32, **kw)
    assert f32.dtype == np.float32 and rel_l2(f32, reff) < TOL


def test_depthmap_layer_range_and_empty(ctx):
    img, dep = scenes.depth_scene(40, 24, seed=6, colors=1)
    args = (40, 24, 64, 32, 8e-6, 0.09, 0.11, 16, [532e-9], img, dep)
    full = ctx.depthmap(*args, quant_mode=1)
    a = ctx.depthmap(*args, quant_mode=1, layers=(1, 9))
    b = ctx.depthmap(*args, quant_mode=1, layers=(10, 16))
    assert rel_l2(a + b, full) < 1e-6
    z = ctx.depthmap(40, 24, 64, 32, 8e-6, 0.09, 0.11, 16, [532e-9], np.zeros_like(img), dep, quant_mode=1)
    assert np.all(z == 0)


def test_depthmap_rgb_1920x1080_full_size(ctx):
    """The RGB workload at full size: 1920x1080 scene, 4096x2048 grid, 256
    layers x 3 colours.  The oracle cannot run all 768 layers in seconds, so
    (a) four populated layers are compared with the oracle at full size and
    (b) linearity over a split of the layer range holds at full size."""
    print("case: depthmap RGB 1920x1080 on 4096x2048 full", flush=True)
    img, dep = scenes.depth_scene(1920, 1080, seed=3)
    lams = (638e-9, 532e-9, 473e-9)
    args = (1920, 1080, 4096, 2048, 8e-6, 0.09, 0.11, 256, lams, img, dep)
    full = ctx.depthmap(*args, out="fringe32")
    lo = ctx.depthmap(*args, out="fringe32", layers=(1, 128))
    hi = ctx.depthmap(*args, out="fringe32", layers=(129, 256))
    err_lin = rel_l2(lo + hi, full)
    print("linearity rel-L2", err_lin)
    assert err_lin < 1e-5
    idx = 256 - dep.astype(np.int32)
    busy = np.bincount(idx.ravel(), minlength=257)
    l0 = int(np.argmax(busy))
    part = ctx.depthmap(*args, layers=(l0, l0 + 3))
    ref = O.depthmap(img[1], dep, 4096, 2048, 8e-6, 8e-6, lams[1], 0.09, 0.11, 256,
                     layers=(l0, l0 + 3), nthreads=0)[1]
    err = rel_l2(part[1], ref)
    print("4-layer full-size rel-L2", err)
    assert err < TOL


# ---- grids that are not powers of two (mixed-radix transforms) --------------------

@pytest.mark.parametrize("w,h,Nx,Ny,L,mode,lens", [
    (720, 540, 720, 540, 16, 1, 0.0),      # unpadded, 2^4 3^2 5 x 2^2 3^3 5
    (600, 500, 720, 540, 32, 1, 1.0),      # scene inside the grid, field lens
    (1920, 1080, 1920, 1080, 8, 1, 0.0),   # HD scene transformed unpadded
])
def test_depthmap_mixed_radix_vs_oracle(ctx, w, h, Nx, Ny, L, mode, lens):
    print(f"case: depthmap mixed {w}x{h} grid {Nx}x{Ny} L={L} lens={lens}", flush=True)
    img, dep = scenes.depth_scene(w, h, seed=11, colors=1)
    kw = dict(quant_mode=mode, field_lens=lens)
    fr = ctx.depthmap(w, h, Nx, Ny, 8e-6, 0.09, 0.11, L, [532e-9], img, dep, **kw)[0]
    ref = O.depthmap(img[0], dep, Nx, Ny, 8e-6, 8e-6, 532e-9, 0.09, 0.11, L, **kw)[1]
    err = rel_l2(fr, ref)
    print("rel-L2", err)
    assert err < TOL
    u = ctx.depthmap(w, h, Nx, Ny, 8e-6, 0.09, 0.11, L, [532e-9], img, dep, out="field", **kw)[0]
    uref = O.depthmap(img[0], dep, Nx, Ny, 8e-6, 8e-6, 532e-9, 0.09, 0.11, L, want_fringe=False,
                      want_field=True, **kw)[2]
    assert rel_l2(u, uref) < TOL


def _shipped():
    from openhologram_amd import lib as L
    import os
    from conftest import GOLDEN
    return L.dm_config_load(os.path.join(GOLDEN, "dm_config_openholo.txt"))


def test_depthmap_shipped_config_11520x2160(ctx):
    """The reference's own configuration (DM/bin/config_openholo.txt: 11520 x
    2160 SLM, pitches 25.4 mm/2250/3 x 25.4 mm/2250, 522 nm, field lens 1 m,
    depth 0.2..0.7 m requantised to 256 levels, no padding) end to end.  The
    scene fills the SLM as the reference's resampled input does.  (a) A scene
    with three depth values against the oracle at full size; (b) a smooth
    depth map over all levels: linearity over a split of the layer range."""
    c = _shipped()
    w, h = c.pn[0], c.pn[1]
    assert (w, h) == (11520, 2160)
    lam = c.lambda_
    print("case: depthmap shipped configuration, 3 levels", flush=True)
    img, _ = scenes.depth_scene(w, h, seed=21, colors=1, gaussians=2)
    dep = np.full((h, w), 40, np.uint8)
    dep[:, w // 3:2 * w // 3] = 128
    dep[h // 2:, :] = 215
    kw = dict(quant_mode=c.flag_change_depth_quantization, field_lens=c.field_lens)
    args = (w, h, w, h, (c.pp[0], c.pp[1]), c.near_depthmap, c.far_depthmap, c.num_of_depth, [lam])
    fr = ctx.depthmap(*args, img, dep, **kw)[0]
    ref = O.depthmap(img[0], dep, w, h, c.pp[0], c.pp[1], lam, c.near_depthmap, c.far_depthmap,
                     c.num_of_depth, **kw)[1]
    err = rel_l2(fr, ref)
    print("3-level rel-L2", err, flush=True)
    assert err < TOL
    print("case: depthmap shipped config all levels", flush=True)
    img, dep = scenes.depth_scene(w, h, seed=22, colors=1)
    full = ctx.depthmap(*args, img, dep, out="fringe32", **kw)[0]
    lo = ctx.depthmap(*args, img, dep, out="fringe32", layers=(1, 128), **kw)[0]
    hi = ctx.depthmap(*args, img, dep, out="fringe32", layers=(129, 256), **kw)[0]
    err_lin = rel_l2(lo + hi, full)
    print("linearity rel-L2", err_lin)
    assert err_lin < 1e-5 and np.all(np.isfinite(full)) and np.abs(full).max() > 0


def test_depthmap_unsupported_grid_reports(ctx):
    """A side that is neither a power of two nor one of the general lengths
    is refused with OPH_ENOTSUP, not computed."""
    from openhologram_amd.lib import OphError
    img, dep = scenes.depth_scene(100, 60, seed=1, colors=1)
    with pytest.raises(OphError):
        ctx.depthmap(100, 60, 100, 60, 8e-6, 0.09, 0.11, 8, [532e-9], img, dep)
    with pytest.raises(OphError):
        ctx.depthmap(100, 60, 720, 64, 8e-6, 0.09, 0.11, 8, [532e-9], img, dep)
