"""Learned correction of a coarse depth-map hologram.

A cheap hologram is the depth-map field computed with few depth layers (the
depth map requantised to, e.g., 16 levels); the full one uses all 256.  A
small residual CNN (oph_cnn_*) maps the first toward the second.  It is
trained only on pairs this method generates; the full-layer method stays the
reference: the reconstruction PSNR of the corrected hologram at sampled
depths is checked against the full-layer hologram's, and below the threshold
the frame falls back to the full-layer result.

Everything here drives liboph_hip.so through the ctypes layer; the
depth-map fields come from oph_depthmap (OPH_DM_OUT_FIELD), the
reconstructions from oph_asm_propagate.
"""
import os

import numpy as np

from . import scenes
from .lib import Cnn

LAMBDA = 532e-9
PITCH = 8e-6
NEAR, FAR = 90e-3, 110e-3


def field(ctx, img, dep, n, layers, pad=2):
    """Complex hologram field (n x n, complex64) of a 1-colour scene with
    `layers` depth levels (256: the u8 depth as is; fewer: requantised)."""
    N = n * pad
    if layers == 256:
        return ctx.depthmap(n, n, N, N, PITCH, NEAR, FAR, 256, [LAMBDA], img, dep, out="field")[0]
    return ctx.depthmap(n, n, N, N, PITCH, NEAR, FAR, layers, [LAMBDA], img, dep, out="field",
                        quant_mode=1)[0]


def _scene(args):
    n, s = args
    return scenes.depth_scene(n, n, seed=int(s), colors=1)


def make_scenes(seeds, n, workers=None):
    """Seeded one-colour depth-map scenes, generated on host worker processes."""
    seeds = [int(s) for s in seeds]
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(seeds) < 8:
        return [_scene((n, s)) for s in seeds]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_scene, [(n, s) for s in seeds], chunksize=4)


def make_pairs(ctx, seeds, n, coarse_layers=16, workers=None):
    """(coarse, full) complex fields for the seeded scenes."""
    sc = make_scenes(seeds, n, workers)
    xs = np.empty((len(sc), n, n), np.complex64)
    ts = np.empty((len(sc), n, n), np.complex64)
    for i, (img, dep) in enumerate(sc):
        xs[i] = field(ctx, img, dep, n, coarse_layers)
        ts[i] = field(ctx, img, dep, n, 256)
    return xs, ts


def to_channels(u, scale):
    """complex (B, H, W) -> NHWC float32 (B, H, W, 2) divided by scale"""
    return np.stack([u.real, u.imag], axis=-1).astype(np.float32) / np.float32(scale)


def from_channels(x, scale):
    return (x[..., 0] + 1j * x[..., 1]).astype(np.complex64) * np.float32(scale)


def reconstruct_intensity(ctx, u, z):
    """|field propagated by +z|^2 (the depth-map generator propagates each
    layer by -z_l), normalised to its maximum."""
    r = ctx.asm(u, PITCH, LAMBDA, z, pad=2)
    i = np.abs(r) ** 2
    return i / max(float(i.max()), 1e-30)


def psnr(a, b):
    mse = float(np.mean((a - b) ** 2))
    return 99.0 if mse <= 0 else 10.0 * np.log10(1.0 / mse)


def sampled_depths(k=3):
    """k depths at the quartiles of the depth range"""
    return [NEAR + (FAR - NEAR) * (i + 1) / (k + 1) for i in range(k)]


def recon_psnr(ctx, u, ref, depths):
    """minimum over the sampled depths of PSNR(recon(u), recon(ref))"""
    return min(psnr(reconstruct_intensity(ctx, u, z), reconstruct_intensity(ctx, ref, z)) for z in depths)


def train(ctx, net, xs, ts, scale, steps=300, batch=8, patch=128, lr=2e-3, seed=0, log=None):
    """Adam on random patches of the training pairs (complex fields, cut and
    converted per batch); returns the loss history."""
    rng = np.random.default_rng(seed)
    n, H, W = xs.shape
    hist = []
    for it in range(steps):
        idx = rng.integers(0, n, batch)
        ys = rng.integers(0, H - patch + 1, batch)
        xs_ = rng.integers(0, W - patch + 1, batch)
        xb = to_channels(np.stack([xs[i, y:y + patch, x:x + patch] for i, y, x in zip(idx, ys, xs_)]), scale)
        tb = to_channels(np.stack([ts[i, y:y + patch, x:x + patch] for i, y, x in zip(idx, ys, xs_)]), scale)
        rate = lr * 0.5 * (1 + np.cos(np.pi * it / steps))   # cosine decay
        hist.append(net.train_step(xb, tb, rate))
        if log and (it % max(1, steps // 10) == 0 or it == steps - 1):
            log(f"step {it} loss {hist[-1]:.4e}")
    return hist


def correct(ctx, net, coarse, scale):
    """CNN correction of one coarse field (H, W) complex -> complex64"""
    y = net.forward(to_channels(coarse[None], scale), bf16_acts=True)
    return from_channels(y[0], scale)


def generate(ctx, net, img, dep, n, scale, threshold_db=35.0, audit=True, coarse_layers=16):
    """The learned-correction pipeline for one frame: coarse field, CNN
    correction, and (audit=True) the reconstruction check against the
    full-layer method at 3 sampled depths with fallback below threshold.
    Returns (field, info)."""
    c = field(ctx, img, dep, n, coarse_layers)
    u = correct(ctx, net, c, scale)
    info = {"fallback": False}
    if audit:
        full = field(ctx, img, dep, n, 256)
        p = recon_psnr(ctx, u, full, sampled_depths())
        info["psnr_db"] = p
        if p < threshold_db:
            info["fallback"] = True
            return full, info
    return u, info
