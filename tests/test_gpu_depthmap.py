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


def test_depthmap_config3_full_size(ctx):
    """The RGB workload at full size: 1920x1080 scene, 4096x2048 grid, 256
    layers x 3 colours.  The oracle cannot run all 768 layers in seconds, so
    (a) four populated layers are compared with the oracle at full size and
    (b) linearity over a split of the layer range holds at full size."""
    print("case: depthmap config3 full", flush=True)
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
