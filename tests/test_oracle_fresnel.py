"""CPU: pins of the Fresnel point-sum oracle (cwoPLS::Fresnel restated with
amplitude, initial phase and the plane carrier)."""
import numpy as np

from conftest import rel_l2
from oracle import oracle as O


def _brute(nx, ny, p, lam, x, y, z, a, ph):
    L = np.longdouble
    out = np.zeros((ny, nx), np.complex128)
    for i in range(ny):
        for j in range(nx):
            X, Y = L(j - nx // 2) * L(p), L(i - ny // 2) * L(p)
            t = (L(z) / L(lam) + ((X - L(x)) ** 2 + (Y - L(y)) ** 2) / (2 * L(z) * L(lam))) % 1
            out[i, j] = np.sum(a * np.exp(1j * (2 * np.pi * t.astype(np.float64) + ph)))
    return out


def test_fresnel_vs_long_double_brute_force():
    rng = np.random.default_rng(1)
    n = 7
    x, y = rng.uniform(-2e-4, 2e-4, (2, n))
    z = rng.choice([0.1, 0.105, 0.11], n)
    a, ph = rng.uniform(0, 1, n), rng.uniform(0, 2 * np.pi, n)
    U = O.fresnel(9, 6, 8e-6, 532e-9, x, y, z, a, ph)
    assert np.abs(U - _brute(9, 6, 8e-6, 532e-9, x, y, z, a, ph)).max() < 1e-9


def test_fresnel_separable_plane_and_sampling():
    """All points on one plane: U = A B^T with 1-D chirp factors (the
    factorisation the MFMA form uses), evaluated independently here."""
    rng = np.random.default_rng(2)
    n, nx, ny, p, lam, z = 40, 32, 24, 8e-6, 532e-9, 0.1
    x, y = rng.uniform(-1e-4, 1e-4, (2, n))
    a, ph = rng.uniform(0, 1, n), rng.uniform(0, 2 * np.pi, n)
    U = O.fresnel(nx, ny, p, lam, x, y, np.full(n, z), a, ph)
    X = (np.arange(nx) - nx // 2) * p
    Y = (np.arange(ny) - ny // 2) * p
    w = a * np.exp(1j * (2 * np.pi * ((z / lam) % 1) + ph))
    A = w[None, :] * np.exp(2j * np.pi * (((Y[:, None] - y[None, :]) ** 2 / (2 * lam * z)) % 1))
    B = np.exp(2j * np.pi * (((X[:, None] - x[None, :]) ** 2 / (2 * lam * z)) % 1))
    assert rel_l2(U, A @ B.T) < 1e-9  # both sides reduce ~2e5 turns in FP64
    rows, cols = np.array([0, 5, 23]), np.array([31, 0, 7])
    S = O.fresnel(nx, ny, p, lam, x, y, np.full(n, z), a, ph, rows=rows, cols=cols)
    assert np.abs(S - U[rows, cols]).max() < 1e-12


def test_fresnel_paraxial_limit_of_rs():
    """Ties the Fresnel oracle to the RS oracle (itself pinned by the compiled
    reference): for points near the axis the Fresnel phase k z + k rho^2/(2z)
    is RS's k r up to k rho^4/(8 z^3) (< 3e-6 rad here).  The grids differ:
    RS pixel centres are ((col + 1/2) p - Lx/2, Ly/2 - (row + 1/2) p)
    (ophgen_pointcloud.cpp:231,235), Fresnel's are ((j - Nx/2) p, (i - Ny/2) p)
    (cwoPLS.cpp:400-401); so RS row r matches Fresnel row Ny - 1 - r with the
    points moved by -p/2 in x and y."""
    rng = np.random.default_rng(5)
    n, nx, ny, p, lam = 12, 32, 24, 8e-6, 532e-9
    x, y = rng.uniform(-1.2e-4, 1.2e-4, (2, n))
    z = rng.uniform(0.19, 0.21, n)
    a, ph = rng.uniform(0.2, 1, n), rng.uniform(0, 2 * np.pi, n)
    rs = O.rs(dict(nx=nx, ny=ny, pitch=p, lambda_=lam), x, y, z, a, ph)

    def fres(dx, dy):
        return O.fresnel(nx, ny, p, lam, x + dx, y + dy, z, a, ph)
    assert rel_l2(fres(-p / 2, -p / 2)[::-1], rs) < 1e-5
    # the row order, the phase sign and the half-pixel origin each matter
    assert rel_l2(fres(-p / 2, -p / 2), rs) > 0.1
    assert rel_l2(np.conj(fres(-p / 2, -p / 2)[::-1]), rs) > 0.2
    assert rel_l2(fres(0.0, 0.0)[::-1], rs) > 1e-3
