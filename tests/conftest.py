import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def rel_l2(a, b):
    """||a - b|| / ||b||"""
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def arg_gray(U):
    """8-bit arg image as the reference example's SaveAsImage(CWO_SAVE_AS_ARG)
    reads on its own output: floor((arg + pi)/(2 pi) * 255) (the empty
    pixels, arg 0, are 127 in the file)."""
    return np.floor((np.angle(U) + np.pi) / (2 * np.pi) * 255).astype(np.int32)


def wrp_table_agreement(U, gray):
    """(fraction exact, fraction within +-1 LSB cyclically) of an arg image
    of U against the reference's wrp_table.bmp pixels."""
    d = (gray.astype(np.int32) - arg_gray(U) + 128) % 256 - 128
    return float(np.mean(d == 0)), float(np.mean(np.abs(d) <= 1))


@pytest.fixture(scope="session")
def golden_keti():
    return np.load(os.path.join(GOLDEN, "keti_fringe.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def ctx():
    from openhologram_amd import Context
    c = Context(0)
    yield c
    c.close()
