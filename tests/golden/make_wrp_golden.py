"""Writes tests/golden/wrp_table.npz from the reference WRP example's own
input and output files (data only; nothing of the reference is run):

  object.txt     19,952 points "x y z" in metres, read as the reference's
                 readxyz does (W/openwrp/WRP.cpp:7-55: whitespace tokens
                 through stod, then stored in the float cwoObjPoint), so the
                 fixture keeps them as float32;
  wrp_table.bmp  the WRP plane that W/openwrp/openwrp.cpp:15-18 saved:
                 Setparameter(532e-9, 8e-6, 515), SingleWRP(0.5e-3),
                 SaveAsImage(CWO_SAVE_AS_ARG); a 24-bit BMP with three equal
                 channels, kept as one u8 channel with the top image row first.

Run in a container that holds the reference tree:
    python tests/golden/make_wrp_golden.py
"""
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/OpenHolo_WRP/openwrp_console_ver 1.0/openwrp"


def read_bmp24_gray(path):
    raw = open(path, "rb").read()
    assert raw[:2] == b"BM"
    off, = struct.unpack_from("<I", raw, 10)
    w, h = struct.unpack_from("<ii", raw, 18)
    bpp, = struct.unpack_from("<H", raw, 28)
    assert bpp == 24
    stride = (w * 3 + 3) & ~3
    img = np.frombuffer(raw, np.uint8, count=stride * abs(h), offset=off)
    img = img.reshape(abs(h), stride)[:, :w * 3].reshape(abs(h), w, 3)
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])
    g = img[..., 0]
    return np.ascontiguousarray(g[::-1] if h > 0 else g)   # bottom-up file -> top row first


def main():
    vals = []
    with open(os.path.join(REF, "object.txt")) as f:
        for line in f:
            vals += [float(t) for t in line.split()]
    xyz = np.array(vals, np.float64).reshape(-1, 3).astype(np.float32)
    gray = read_bmp24_gray(os.path.join(REF, "wrp_table.bmp"))
    np.savez_compressed(os.path.join(HERE, "wrp_table.npz"), xyz=xyz, gray=gray,
                        params=np.array([532e-9, 8e-6, 515, 0.5e-3], np.float32))
    print("wrote wrp_table.npz:", xyz.shape, gray.shape)


if __name__ == "__main__":
    main()
