"""FP64 numpy restatement of the MATLAB hologram encodings of
OpenHolo_Encoding/fn_encoding.m (with fn_myfft2.m / fn_myifft2.m).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
encodings (oph_encode_field); the product package never imports it.

Each function follows the reference lines cited; MATLAB's 1-based ranges are
converted to 0-based slices.  The reference's own quirks are kept (noted
where they occur).  How each is pinned: tests/test_oracle_encoding.py.
"""
import numpy as np

SSB, NI, P2A, REAL, BURCKHARDT, TWO_PHASE = range(6)
LEFT, RIGHT, TOP, BOTTOM = range(4)


def myfft2(u):
    """fn_myfft2.m:2: fftshift(fft2(ifftshift(u)))."""
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(u)))


def myifft2(a):
    """fn_myifft2.m:2: fftshift(ifft2(ifftshift(a)))."""
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(a)))


def ssb(u, passband=LEFT):
    """Single-sideband (fn_encoding.m:29-47).  The zeroed block starts at
    1-based index floor(N/2) (:34-40), i.e. 0-based floor(N/2) - 1, so the
    'left' band also drops the DC column and the one below it."""
    ny, nx = u.shape
    a = myfft2(np.asarray(u, np.complex128))
    if passband == LEFT:
        a[:, nx // 2 - 1:] = 0
    elif passband == RIGHT:
        a[:, :nx // 2] = 0
    elif passband == TOP:
        a[ny // 2 - 1:, :] = 0
    elif passband == BOTTOM:
        a[:ny // 2, :] = 0
    re = myifft2(a).real
    pos = re - re.min()
    return pos / pos.max()


def ni(u):
    """Numerical interference with a plane reference of amplitude max|U|
    (fn_encoding.m:49-54)."""
    u = np.asarray(u, np.complex128)
    ref = np.abs(u).max()
    i = np.abs(u + ref) ** 2
    return i / i.max()


def p2a(u):
    """fn_encoding.m:57 as written: angle(U) + pi/(2 pi) = angle(U) + 0.5
    (operator precedence; the intended (angle + pi)/(2 pi) is not what the
    reference computes)."""
    return np.angle(np.asarray(u, np.complex128)) + 0.5


def real(u):
    """fn_encoding.m:20-21."""
    return np.asarray(u, np.complex128).real.copy()


def burckhardt(u):
    """Burckhardt three-component encoding (fn_encoding.m:60-83), output
    Ny x 3Nx with columns A1, A2, A3 interleaved.  As written, the second
    term of each branch (:69, :73, :77) lacks the amplitude factor, and a
    phase of exactly 2 pi (angle = pi) falls in no branch (all zero)."""
    u = np.asarray(u, np.complex128)
    ny, nx = u.shape
    u = u / np.abs(u).max()
    ph = np.angle(u) + np.pi
    am = np.abs(u)
    A1, A2, A3 = np.zeros((3, ny, nx))
    s3 = np.sqrt(3)
    m = (ph >= 0) & (ph < 2 * np.pi / 3)
    A1[m] = am[m] * (np.cos(ph[m]) + np.sin(ph[m]) / s3)
    A2[m] = 2 * np.sin(ph[m]) / s3
    m = (ph >= 2 * np.pi / 3) & (ph < 4 * np.pi / 3)
    q = ph[m] - 2 * np.pi / 3
    A2[m] = am[m] * (np.cos(q) + np.sin(q) / s3)
    A3[m] = 2 * np.sin(q) / s3
    m = (ph >= 4 * np.pi / 3) & (ph < 2 * np.pi)
    q = ph[m] - 4 * np.pi / 3
    A3[m] = am[m] * (np.cos(q) + np.sin(q) / s3)
    A1[m] = 2 * np.sin(q) / s3
    out = np.zeros((ny, 3 * nx))
    out[:, 0::3], out[:, 1::3], out[:, 2::3] = A1, A2, A3
    return out


def two_phase(u):
    """Two-phase (double-phase) encoding (fn_encoding.m:85-98): phases
    angle + pi +- acos(|U|/max|U|), output Ny x 2Nx interleaved."""
    u = np.asarray(u, np.complex128)
    ny, nx = u.shape
    u = u / np.abs(u).max()
    am = np.abs(u)
    ph = np.angle(u) + np.pi
    d = np.arccos(np.minimum(am, 1.0))
    out = np.zeros((ny, 2 * nx))
    out[:, 0::2], out[:, 1::2] = ph + d, ph - d
    return out


def encode(u, method, passband=LEFT):
    """fn_encoding.m:1-27 dispatch."""
    return {SSB: lambda: ssb(u, passband), NI: lambda: ni(u), P2A: lambda: p2a(u),
            REAL: lambda: real(u), BURCKHARDT: lambda: burckhardt(u),
            TWO_PHASE: lambda: two_phase(u)}[method]()
