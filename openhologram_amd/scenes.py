"""Seeded synthetic inputs of the five benchmark workloads.

A portable SplitMix64 stream (uniforms = (u64 >> 11) * 2**-53) so that the
GPU path, the CPU oracle and any later run see bit-identical doubles without
depending on numpy's generators.
"""
import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


class SplitMix64:
    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def u64(self, n):
        # state_i = seed + (i+1) * golden; output = mix(state_i)
        with np.errstate(over="ignore"):
            g = np.uint64(0x9E3779B97F4A7C15)
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * g
            self.state = self.state + np.uint64(n) * g
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
        return lo + (hi - lo) * u

    def integers(self, n, k):
        return (self.u64(n) % np.uint64(k)).astype(np.int64)


def rs_points(n=1000, seed=1):
    """Plain RS workload: n points uniform in a 10 mm cube centred 100 mm from
    the hologram, amplitude U[0,1), initial phase U[0, 2 pi)."""
    g = SplitMix64(seed)
    x = g.uniform(n, -5e-3, 5e-3)
    y = g.uniform(n, -5e-3, 5e-3)
    z = g.uniform(n, 95e-3, 105e-3)
    a = g.uniform(n)
    ph = g.uniform(n, 0.0, 2 * np.pi)
    return x, y, z, a, ph


RS_SPEC = dict(nx=1024, ny=1024, pitch=8e-6, lambda_=532e-9)


def wrp_points(n=100_000, seed=2):
    """WRP workload: n points uniform over 20 x 20 mm laterally and 5 mm in
    depth, 101-106 mm from the hologram.  Coordinates are along the
    propagation axis (light travels +z towards the hologram at z = 0), so
    z in [-106, -101] mm; amplitude U[0,1).  Returns x, y, z, a and the WRP
    plane z_wrp = max(z) + 1 mm (1 mm in front of the nearest point)."""
    g = SplitMix64(seed)
    x = g.uniform(n, -10e-3, 10e-3)
    y = g.uniform(n, -10e-3, 10e-3)
    z = g.uniform(n, -106e-3, -101e-3)
    a = g.uniform(n)
    z_wrp = float(z.max()) + 1e-3
    return x, y, z, a, z_wrp


WRP_SPEC = dict(n=2048, pitch=8e-6, lambda_=532e-9, pad=2)


def fresnel_points(n=1_000_000, planes=64, seed=4, z0=100e-3, z1=120e-3, half=15e-3):
    """Fresnel/MFMA workload: n points on `planes` depth planes
    z_p = z0 + (z1 - z0) p/(planes-1), x, y uniform over +-half, amplitude
    U[0,1), initial phase U[0, 2 pi).  Returns x, y, plane index, z, a, ph."""
    g = SplitMix64(seed)
    p = g.integers(n, planes)
    zp = z0 + (z1 - z0) * np.arange(planes) / max(planes - 1, 1)
    x = g.uniform(n, -half, half)
    y = g.uniform(n, -half, half)
    a = g.uniform(n)
    ph = g.uniform(n, 0.0, 2 * np.pi)
    return x, y, p.astype(np.int32), zp[p], a, ph


def _box_blur(u, k):
    """k x k box blur with edge padding (cumulative sums, separable)."""
    r = k // 2
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r + 1, r)
        c = np.cumsum(np.pad(u, pad, mode="edge"), axis=axis)
        n = u.shape[axis]
        hi = np.take(c, np.arange(k, k + n), axis=axis)
        lo = np.take(c, np.arange(0, n), axis=axis)
        u = (hi - lo) / k
    return u


def depth_scene(w=1920, h=1080, seed=3, colors=3, blur=9, gaussians=8):
    """Depth-map workload: depth = sum of `gaussians` seeded 2-D Gaussians
    (centre uniform over the image, sigma U[50,300] px scaled by the image
    width / 1920, weight U[0.5,1]) normalised to [0,1] and quantised to u8;
    intensity per colour = seeded U[0,1) noise box-blurred (blur x blur),
    normalised and quantised to u8.  Returns (intensity[colors, h, w] u8,
    depth[h, w] u8)."""
    g = SplitMix64(seed)
    xs = np.arange(w, dtype=np.float64)
    ys = np.arange(h, dtype=np.float64)
    d = np.zeros((h, w))
    s = w / 1920.0
    for _ in range(gaussians):
        cx, cy = g.uniform(1, 0, w)[0], g.uniform(1, 0, h)[0]
        sig = g.uniform(1, 50, 300)[0] * s
        wt = g.uniform(1, 0.5, 1.0)[0]
        gx = np.exp(-(xs - cx) ** 2 / (2 * sig * sig))
        gy = np.exp(-(ys - cy) ** 2 / (2 * sig * sig))
        d += wt * np.outer(gy, gx)
    d = (d - d.min()) / max(d.max() - d.min(), 1e-300)
    dep = np.floor(d * 255 + 0.5).astype(np.uint8)
    imgs = []
    for _ in range(colors):
        u = _box_blur(g.uniform(w * h).reshape(h, w), blur)
        u = (u - u.min()) / max(u.max() - u.min(), 1e-300)
        imgs.append(np.floor(u * 255 + 0.5).astype(np.uint8))
    return np.stack(imgs), dep


DEPTH_SPEC = dict(w=1920, h=1080, Nx=4096, Ny=2048, pitch=8e-6, near=90e-3, far=110e-3, L=256,
                  lambdas=(638e-9, 532e-9, 473e-9))
