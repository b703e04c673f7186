"""GPU parity of the MATLAB encodings (oph_encode_field) against the FP64
restatement of OpenHolo_Encoding/fn_encoding.m, on the same complex64 input."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import encoding_ref as E

pytestmark = pytest.mark.gpu

METHODS = {"ssb": E.SSB, "ni": E.NI, "p2a": E.P2A, "real": E.REAL, "burckhardt": E.BURCKHARDT,
           "two_phase": E.TWO_PHASE}


def _field(ny, nx, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)


def _away_from_cuts(u, cuts, eps=1e-5):
    """Pixels whose phase angle + pi is not within eps of a branch boundary
    (float rounding may put those in the neighbouring branch)."""
    ph = np.angle(u.astype(np.complex128)) + np.pi
    return np.all([np.abs(ph - c) > eps for c in cuts], axis=0)


@pytest.mark.parametrize("method", list(METHODS))
def test_encode_field_vs_oracle(ctx, method):
    print(f"case: encode_field {method} 128x64", flush=True)
    u = _field(64, 128, 11)
    out = ctx.encode_field(u, method)
    ref = E.encode(u.astype(np.complex128), METHODS[method])
    assert out.shape == ref.shape
    if method in ("burckhardt", "two_phase", "p2a"):
        k = ref.shape[1] // u.shape[1]
        ok = _away_from_cuts(u, [0.0, 2 * np.pi / 3, 4 * np.pi / 3, 2 * np.pi])
        ok = np.repeat(ok, k, axis=1)
        assert ok.mean() > 0.99
        err = rel_l2(out[ok], ref[ok])
    else:
        err = rel_l2(out, ref)
    print(method, "rel-L2", err)
    assert err < 1e-5


@pytest.mark.parametrize("passband", ["left", "right", "top", "bottom"])
def test_encode_ssb_passbands(ctx, passband):
    u = _field(32, 256, 12)
    out = ctx.encode_field(u, "ssb", passband=passband)
    ref = E.ssb(u.astype(np.complex128), {"left": E.LEFT, "right": E.RIGHT, "top": E.TOP,
                                          "bottom": E.BOTTOM}[passband])
    assert rel_l2(out, ref) < 1e-5


def test_encode_field_device_and_errors(ctx):
    from openhologram_amd import OphError
    u = _field(16, 16, 13)
    d = ctx.to_device(u)
    out = ctx.empty((16, 32), np.float32)
    ctx.encode_field(d, "two_phase", device=True, out=out)
    assert np.allclose(out.numpy(), ctx.encode_field(u, "two_phase"))
    with pytest.raises(OphError):
        ctx.encode_field(_field(16, 24, 1), "ssb")   # SSB needs power-of-two sides
    with pytest.raises(OphError):
        ctx.encode_field(u, "ssb", asynchronous=True)  # async needs device pointers
