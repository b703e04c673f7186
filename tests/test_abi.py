"""CPU: the C ABI library loads, exports every declared entry point, and its
host-side logic (file readers, argument checks without a GPU) behaves."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from openhologram_amd import lib as L


def _declared():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        names += L.header_functions(h)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version():
    assert L.load().oph_version().decode().endswith("gfx950")


def test_keti_spec_parser_matches_reference_fixture(golden_keti):
    s = L.spec_from_keti_config(os.path.join(GOLDEN, "keti_TestSpec.config"))
    ref = golden_keti["spec"]
    assert (s.nx, s.ny) == (int(ref[6]), int(ref[7]))
    got = np.array([s.scale[0], s.scale[1], s.scale[2], s.offset_z, s.pitch_x, s.pitch_y,
                    s.lambda_, s.tilt_deg[0], s.tilt_deg[1]])
    want = ref[[0, 1, 2, 3, 4, 5, 8, 9, 10]]
    # the reference stores floats; the ABI keeps the decimal value in double
    assert np.array_equal(got.astype(np.float32), want)
    assert s.lambda_ == 532e-9 and s.pitch_x == 4.62e-7


def test_keti_loader_matches_reference_loader(golden_keti):
    pts = L.load_keti_points(os.path.join(GOLDEN, "keti_TestPointCloud.dat"))
    ref = golden_keti["points"]  # x, y, z, amplitude, phase (reference struct order)
    assert pts.shape == (1453, 5)
    assert np.array_equal(pts[:, 0:3].astype(np.float32), ref[:, 0:3])
    assert np.array_equal(pts[:, 3].astype(np.float32), ref[:, 4])  # phase column
    assert np.array_equal(pts[:, 4].astype(np.float32), ref[:, 3])  # amplitude column


def test_loader_errors(tmp_path):
    with pytest.raises(L.OphError) as e:
        L.load_keti_points(str(tmp_path / "missing.dat"))
    assert e.value.code == L.OPH_EIO
    bad = tmp_path / "bad.dat"
    bad.write_text("2\n0\t1 2 3 4 5\n5\t1 2 3 4 5\n")   # index mismatch (ophgen_pointcloud.cpp:32-42)
    with pytest.raises(L.OphError):
        L.load_keti_points(str(bad))
    cfg = tmp_path / "short.config"
    cfg.write_text("".join(f"k{i} = {i}\n" for i in range(16)))  # != 17 lines
    with pytest.raises(L.OphError):
        L.spec_from_keti_config(str(cfg))


def test_null_context_is_rejected():
    lib = L.load()
    s = L.make_spec(8, 8, 8e-6, 532e-9)
    assert lib.oph_pc_rs(None, C.byref(s), None, None, 0) == L.OPH_EINVAL
    assert lib.oph_last_error(None) == b"null context"
