"""Writes the golden fixtures of tests/golden/ from the reference's own KETI
sources, compiled by oracle/Makefile into oracle/_ref/libketi_ref.so.

Run in a container that holds the reference tree:
    python tests/golden/make_golden.py

Fixtures (data only):
  keti_TestPointCloud.dat, keti_TestSpec.config  -- the reference's example
      input files (K/example_cpu/), copied byte for byte
  keti_fringe.npz  -- for each case: the float32 fringe produced by the
      reference ophGenCghPointCloud (single-threaded) on those points, the
      13-float spec it was run with, and the reference ophNormalize u8 image
      of the first case; plus the points as the reference loader parsed them.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle  # noqa: E402

REF = "/root/reference/OpenHolo_KETI/VS/Openholo_Generation_PointCloud_Library_beta_v1.0/example_cpu"


def cases(spec):
    """(name, spec13) fringe cases around the reference example spec."""
    out = []
    s = spec.copy(); s[6], s[7] = 96, 64
    out.append(("example_96x64", s))                       # tilt 22.5, 0.462 um
    s = spec.copy(); s[4] = s[5] = 8e-6; s[6], s[7] = 64, 48; s[9] = s[10] = 0.0
    s[3] = 0.1                                             # 8 um, no tilt, far field
    out.append(("pitch8um_64x48", s))
    s = spec.copy(); s[6], s[7] = 128, 40; s[9], s[10] = 10.0, -5.0; s[4], s[5] = 1e-6, 2e-6
    out.append(("aniso_128x40", s))
    return out


def main():
    oracle.build(ref=True)
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libketi_ref.so not built (reference tree absent?)")
    for src, dst in [("TestPointCloud.dat", "keti_TestPointCloud.dat"),
                     ("TestSpec.config", "keti_TestSpec.config")]:
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
    buf = np.zeros((4096, 5), np.float32)
    n = ref.ref_keti_load(os.path.join(REF, "TestPointCloud.dat").encode(), buf.ctypes.data, 4096)
    pts = buf[:n].copy()
    spec = np.zeros(13, np.float32)
    assert ref.ref_keti_spec(os.path.join(REF, "TestSpec.config").encode(), spec.ctypes.data) == 1
    data = {"points": pts, "spec": spec}
    names = []
    for name, s in cases(spec):
        dst = np.zeros((int(s[7]), int(s[6])), np.float32)
        ref.ref_keti_fringe(pts.ctypes.data, n, s.ctypes.data, dst.ctypes.data)
        data[f"fringe_{name}"] = dst
        data[f"spec_{name}"] = s
        names.append(name)
    first = data[f"fringe_{names[0]}"].copy()
    u8 = np.zeros(first.shape, np.uint8)
    ref.ref_keti_normalize(first.ctypes.data, u8.ctypes.data, first.shape[1], first.shape[0])
    data["normalized_u8"] = u8
    data["case_names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "keti_fringe.npz"), **data)
    print("wrote", os.path.join(HERE, "keti_fringe.npz"), "cases", names)


if __name__ == "__main__":
    main()
