import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sal-fort_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def rel_l2(ax, ay, bx, by):
    """relative L2 error of (ax, ay) against (bx, by) over both components"""
    num = np.sum((np.asarray(ax) - bx) ** 2 + (np.asarray(ay) - by) ** 2)
    den = np.sum(np.asarray(bx) ** 2 + np.asarray(by) ** 2)
    return float(np.sqrt(num / den))


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


def fib(n):
    i = np.arange(n)
    z = 1 - (2 * i + 1) / n
    ph = i * np.pi * (3 - np.sqrt(5))
    r = np.sqrt(1 - z * z)
    return r * np.cos(ph), r * np.sin(ph), z


@pytest.fixture(scope="session")
def oracle():
    from oracle_py import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
