"""CPU: the reference-shaped C++ header (include/openholo_hip.hpp) used from
C++ -- loaders, spec file, setters, depth-map config, BMP and xyz files --
compared with the Python binding of the same library and the fixtures."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN
from openhologram_amd import lib as L

CPP = os.path.join(os.path.dirname(__file__), "cpp")
EXE = os.path.join(CPP, "shim_check")


@pytest.fixture(scope="module")
def host_run(tmp_path_factory):
    subprocess.run(["make", "-s", "-C", CPP], check=True)
    tmp = tmp_path_factory.mktemp("shim")
    out = subprocess.run([EXE, "host", GOLDEN, str(tmp)], check=True, capture_output=True, text=True).stdout
    lines = {}
    for line in out.splitlines():
        k, _, v = line.partition(" ")
        lines[k] = v.split()
    return lines, tmp


def test_shim_point_loaders(host_run):
    lines, _ = host_run
    pts = L.load_keti_points(os.path.join(GOLDEN, "keti_TestPointCloud.dat"))   # x y z phase amp
    assert lines["points"] == [str(len(pts))] * 3
    first = np.array(lines["first"], np.float32)
    last = np.array(lines["last"], np.float32)
    want = lambda r: np.array([r[0], r[1], r[2], r[4], r[3]], np.float32)   # x y z amplitude phase
    assert np.array_equal(first, want(pts[0])) and np.array_equal(last, want(pts[-1]))
    assert lines["missing"] == ["-1"]


def test_shim_spec_and_setters(host_run, golden_keti):
    lines, _ = host_run
    assert lines["spec_ok"][0] == "1" and lines["spec_ok"][2] == "circle"
    v = np.array(lines["spec"], np.float32)
    s = golden_keti["spec"]   # the reference parser's 13 values of the same file
    assert np.array_equal(v[[0, 1, 2, 3, 4, 5, 6, 7, 11, 12, 13]], s[[0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10]])
    assert np.array_equal(v[[8, 9]], s[[11, 12]])
    assert np.float32(lines["setters"][0]) == np.float32(633e-9) and lines["setters"][1:] == ["7", "9"]


def test_shim_config_bmp_xyz(host_run):
    lines, tmp = host_run
    assert lines["config_ok"] == ["1"]
    assert lines["config"][:2] == ["11520", "2160"] and lines["config"][4:] == ["256", "256"]
    px = (np.arange(64 * 4) * 7 % 256).astype(np.uint8).reshape(4, 64)
    L.save_bmp8(str(tmp / "py.bmp"), px)
    assert open(tmp / "shim.bmp", "rb").read() == open(tmp / "py.bmp", "rb").read()
    assert lines["xyz"] == ["0", "2", "-0.1", "-0.1005"]
