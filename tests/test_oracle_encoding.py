"""CPU: pins of the FP64 restatement of the MATLAB encodings
(OpenHolo_Encoding/fn_encoding.m), by closed forms and hand-computed values."""
import numpy as np
import pytest

from oracle import encoding_ref as E


def _wave(ny, nx, kx, ky, amp=1.0):
    y, x = np.mgrid[0:ny, 0:nx]
    return amp * np.exp(2j * np.pi * (kx * x / nx + ky * y / ny))


@pytest.mark.parametrize("passband,kept,dropped", [
    (E.LEFT, [(-2, 0), (-16, 3)], [(-1, 0), (0, 0), (15, 2)]),
    (E.RIGHT, [(0, 0), (15, 1)], [(-1, 0), (-16, 2)]),
    (E.TOP, [(0, -2), (3, -16)], [(0, -1), (1, 0), (2, 15)]),
    (E.BOTTOM, [(0, 0), (1, 15)], [(0, -1), (2, -16)]),
])
def test_ssb_plane_waves(passband, kept, dropped):
    """A sum of plane waves: those in the pass band come through unchanged,
    the others vanish (the block zeroed from 1-based floor(N/2) also takes
    frequency -1 and DC for left/top); the result is the real part scaled to
    [0, 1]."""
    ny, nx = 32, 32
    rng = np.random.default_rng(passband)
    u_keep = sum(_wave(ny, nx, kx, ky, rng.uniform(0.5, 1) * np.exp(1j * rng.uniform(0, 6)))
                 for kx, ky in kept)
    u_drop = sum(_wave(ny, nx, kx, ky, rng.uniform(0.5, 1)) for kx, ky in dropped)
    out = E.ssb(u_keep + u_drop, passband)
    re = u_keep.real
    assert np.abs(out - (re - re.min()) / (re.max() - re.min())).max() < 1e-12


def test_ni_hand_values():
    u = np.array([[2j, 1.0, -1.0]])
    # reference amplitude 2: |2i + 2|^2 = 8, |3|^2 = 9, |1|^2 = 1
    assert np.allclose(E.ni(u), [[8 / 9, 1.0, 1 / 9]], rtol=0, atol=1e-15)


def test_p2a_and_real_hand_values():
    u = np.array([[1j, -1.0, 1.0 - 1j]])
    assert np.allclose(E.p2a(u), [[np.pi / 2 + 0.5, np.pi + 0.5, -np.pi / 4 + 0.5]], atol=1e-15)
    assert np.array_equal(E.real(u), [[0.0, -1.0, 1.0]])


def test_burckhardt_reconstructs_unit_amplitude():
    """For |U| = max|U| the three components are non-negative and
    A1 + A2 e^{i 2pi/3} + A3 e^{i 4pi/3} = e^{i(angle + pi)}; for smaller
    amplitudes the amplitude-carrying component is scaled and the other
    is not (the reference's terms as written)."""
    rng = np.random.default_rng(3)
    ph = rng.uniform(-np.pi + 1e-3, np.pi - 1e-3, (8, 16))
    u = np.exp(1j * ph)
    B = E.burckhardt(u)
    A1, A2, A3 = B[:, 0::3], B[:, 1::3], B[:, 2::3]
    assert B.min() > -1e-12
    w = np.exp(2j * np.pi / 3)
    assert np.abs(A1 + A2 * w + A3 * w * w - np.exp(1j * (ph + np.pi))).max() < 1e-12
    a = rng.uniform(0.1, 1.0, (8, 16))
    a[0, 0] = 1.0
    Ba = E.burckhardt(a * u)
    # first branch: A1 carries the amplitude, A2 does not
    t = ph + np.pi
    first = t < 2 * np.pi / 3
    assert np.allclose(Ba[:, 0::3][first], (a * A1)[first], atol=1e-12)
    assert np.allclose(Ba[:, 1::3][first], A2[first], atol=1e-12)


def test_two_phase_sum_is_field():
    """e^{i A1} + e^{i A2} = 2 (|U|/max|U|) e^{i(angle + pi)}."""
    rng = np.random.default_rng(4)
    u = rng.standard_normal((12, 10)) + 1j * rng.standard_normal((12, 10))
    T = E.two_phase(u)
    s = np.exp(1j * T[:, 0::2]) + np.exp(1j * T[:, 1::2])
    m = np.abs(u).max()
    assert np.abs(s - 2 * np.abs(u) / m * np.exp(1j * (np.angle(u) + np.pi))).max() < 1e-12
