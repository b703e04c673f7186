"""CPU: the KETI point-cloud oracle pinned against the reference's compiled
source (golden fixtures) and against geometry and higher-precision checks."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
from oracle import oracle as O


def _spec64(s):
    return dict(nx=int(s[6]), ny=int(s[7]), pitch=(float(s[4]), float(s[5])), lambda_=float(s[8]),
                scale=tuple(float(v) for v in s[0:3]), offset_z=float(s[3]),
                tilt_deg=(float(s[9]), float(s[10])))


def test_fp32_restatement_bit_exact(golden_keti):
    """The FP32 restatement reproduces the reference fringe bit for bit."""
    pts = golden_keti["points"]
    for name in golden_keti["case_names"]:
        s = golden_keti[f"spec_{name}"]
        mine = O.keti_fringe_f32(s, pts)
        ref = golden_keti[f"fringe_{name}"]
        assert np.array_equal(mine, ref), name


def test_fp32_restatement_accumulates(golden_keti):
    pts = golden_keti["points"][:50]
    s = golden_keti["spec_example_96x64"]
    one = O.keti_fringe_f32(s, pts)
    two = O.keti_fringe_f32(s, pts, dst=one.copy())
    assert np.allclose(two, 2 * one, rtol=1e-6, atol=1e-4)


def test_normalize_bit_exact(golden_keti):
    name = golden_keti["case_names"][0]
    assert np.array_equal(O.normalize_u8(golden_keti[f"fringe_{name}"]), golden_keti["normalized_u8"])


def test_ref_library_matches_fixtures(golden_keti):
    """Where the reference tree is present, its compiled library still gives
    the committed fixtures (the fixtures are not stale)."""
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built here")
    pts = np.ascontiguousarray(golden_keti["points"])
    name = golden_keti["case_names"][1]
    s = np.ascontiguousarray(golden_keti[f"spec_{name}"])
    dst = np.zeros((int(s[7]), int(s[6])), np.float32)
    ref.ref_keti_fringe(pts.ctypes.data, len(pts), s.ctypes.data, dst.ctypes.data)
    assert np.array_equal(dst, golden_keti[f"fringe_{name}"])


def test_fp64_oracle_vs_reference_fringe(golden_keti):
    """Re(U) of the FP64 field at phase 0 equals the reference FP32 fringe up
    to the reference's own FP32 phase noise (k r ~ 1e5..1e6 rad)."""
    pts = golden_keti["points"]
    for name in golden_keti["case_names"]:
        s = golden_keti[f"spec_{name}"]
        U = O.rs(_spec64(s), pts[:, 0], pts[:, 1], pts[:, 2], pts[:, 3])
        err = rel_l2(U.real, golden_keti[f"fringe_{name}"])
        assert err < 0.1, (name, err)
        # a y-flip or a missing half-pixel would be far off
        assert rel_l2(U.real[::-1], golden_keti[f"fringe_{name}"]) > 0.3


def test_fp64_single_point_long_double():
    """Phase of one point against an 80-bit long-double evaluation."""
    spec = dict(nx=33, ny=17, pitch=(8e-6, 6e-6), lambda_=532e-9, tilt_deg=(3.0, -2.0))
    x, y, z, a, ph = 1.3e-3, -0.7e-3, 0.1013, 0.37, 2.1
    U = O.rs(spec, [x], [y], [z], [a], [ph])
    L = np.longdouble
    col = np.arange(33, dtype=L)
    row = np.arange(17, dtype=L)[:, None]
    X = (col + L(0.5)) * L(8e-6) - L(33) * L(8e-6) / 2
    Y = L(17) * L(6e-6) / 2 - (row + L(0.5)) * L(6e-6)
    r = np.sqrt((X - L(x)) ** 2 + (Y - L(y)) ** 2 + L(z) ** 2)
    lam = L(532e-9)
    turns = r / lam - (X * np.sin(L(3.0) * np.pi / 180) + Y * np.sin(L(-2.0) * np.pi / 180)) / lam
    turns = turns - np.floor(turns)
    expect = a * np.exp(1j * (2 * np.pi * turns.astype(np.float64) + ph))
    assert np.abs(U - expect).max() < 1e-9


def test_fp64_symmetry_and_linearity():
    spec = dict(nx=64, ny=48, pitch=8e-6, lambda_=532e-9)
    U = O.rs(spec, [0.0], [0.0], [0.1], [1.0])
    assert np.allclose(U, U[:, ::-1], atol=1e-9)   # X(col) = -X(nx-1-col)
    assert np.allclose(U, U[::-1, :], atol=1e-9)
    rng = np.random.default_rng(5)
    x, y = rng.uniform(-1e-3, 1e-3, (2, 20))
    z = rng.uniform(0.09, 0.11, 20)
    a = rng.uniform(0, 1, 20)
    full = O.rs(spec, x, y, z, a)
    part = O.rs(spec, x[:7], y[:7], z[:7], a[:7]) + O.rs(spec, x[7:], y[7:], z[7:], a[7:])
    assert np.allclose(full, part, atol=1e-10)
    # translating a point by one pitch in +x shifts the field one column right
    s1 = O.rs(spec, [2e-4], [1e-4], [0.1], [1.0])
    s2 = O.rs(spec, [2e-4 + 8e-6], [1e-4], [0.1], [1.0])
    assert np.allclose(s2[:, 1:], s1[:, :-1], atol=1e-8)


def test_loader_files_present():
    assert os.path.exists(os.path.join(GOLDEN, "keti_TestPointCloud.dat"))
    assert os.path.exists(os.path.join(GOLDEN, "keti_TestSpec.config"))


def _huygens_brute(nx, ny, p, lam, z, ph, x, y, zz, a, phi):
    L = np.longdouble
    out = np.zeros((ny, nx), np.complex128)
    for i in range(ny):
        for j in range(nx):
            dx, dy, dz = L(j - nx // 2) * L(p) - x.astype(L), L(i - ny // 2) * L(p) - y.astype(L), zz.astype(L) + L(z)
            r = np.sqrt(dx * dx + dy * dy + dz * dz)
            t = (r / L(lam)) % 1
            out[i, j] = np.sum(a / r.astype(np.float64) * np.exp(1j * (2 * np.pi * t.astype(np.float64) + ph + phi)))
    return out


def test_huygens_brute_force_and_rs():
    """cwoPLS::Huygens (CWO/cwoPLS.cpp:316-368) restated: against a
    long-double per-pair sum, and tied to the reference-pinned RS oracle: for
    one point, U r = a exp(i(k r + phi)) is orc_rs's field on the same
    pixels (RS centres are (col + 1/2) p - Lx/2 and Ly/2 - (row + 1/2) p, so
    the point moves by -p/2 and the rows flip)."""
    rng = np.random.default_rng(6)
    n, nx, ny, p, lam = 6, 9, 7, 8e-6, 633e-9
    x, y = rng.uniform(-3e-5, 3e-5, (2, n))
    zz = rng.uniform(0.02, 0.03, n)
    a, phi = rng.uniform(0.2, 1, n), rng.uniform(0, 2 * np.pi, n)
    U = O.huygens(nx, ny, p, lam, 0.07, x, y, zz, a, phi, ph=0.3)
    assert np.abs(U - _huygens_brute(nx, ny, p, lam, 0.07, 0.3, x, y, zz, a, phi)).max() < 1e-8 * np.abs(U).max()
    # one point against orc_rs
    x1, y1, z1 = np.array([1.2e-5]), np.array([-2.5e-5]), np.array([0.1])
    H = O.huygens(32, 24, p, lam, 0.0, x1 - p / 2, y1 - p / 2, z1, [0.8], [0.4])[::-1]
    R = O.rs(dict(nx=32, ny=24, pitch=p, lambda_=lam), x1, y1, z1, [0.8], [0.4])
    Y, X = np.mgrid[0:24, 0:32]
    r = np.sqrt(((X + 0.5) * p - 16 * p - x1[0]) ** 2 + (12 * p - (Y + 0.5) * p - y1[0]) ** 2 + z1[0] ** 2)
    assert np.abs(H * r - R).max() < 1e-9
    assert np.abs(np.conj(H) * r - R).max() > 0.5
