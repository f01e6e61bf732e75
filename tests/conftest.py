"""Shared fixtures.  GPU tests carry ``@pytest.mark.gpu``; everything else runs
on the CPU (the oracle, the golden vectors and the host side of libvct)."""
from __future__ import annotations

import json
import lzma
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 (MI355X) GPU")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True,
                       capture_output=True)
    return lib


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def golden():
    return Golden()


class Golden:
    """Accessors for the committed fixtures in tests/golden/."""

    def path(self, name):
        return os.path.join(GOLDEN, name)

    def xz(self, name, dtype, shape):
        with open(self.path(name), "rb") as f:
            return np.frombuffer(lzma.decompress(f.read()), dtype).reshape(shape).copy()

    def npz(self, name):
        return np.load(self.path(name), allow_pickle=False)

    def hashes(self):
        with open(self.path("kat_hashes.json")) as f:
            return json.load(f)

    def icru110(self):
        return self.xz("icru110_phantom.i16.xz", np.int16, (6, 512, 512))

    def kat_f_input(self):
        return self.xz("kat_f_input.f32.xz", np.float32, (800, 609))

    def sigma0_120kv(self):
        return self.xz("icru110_120kV_sigma0.f32.xz", np.float32, (800, 609))

    def synthetic_spectrum(self, name, directory):
        t = self.npz("ref_tables.npz")[f"synthetic_{name}_text"].tobytes()
        p = os.path.join(directory, f"synthetic_{name}.txt")
        with open(p, "wb") as f:
            f.write(t)
        return p


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The torch device for GPU tests; skips (never falls back) without one."""
    if not gpu_available():
        pytest.skip("no GPU visible")
    import torch
    from virtualct_amd import _lib
    _lib.lib()  # raises if libvct.so is not built
    return torch.device("cuda", 0)


def rel_max(a, b):
    """max|a-b| / max|b|"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# The measured-data loader cases and the seeded inputs that
# tests/golden/make_golden.py gave the reference: (name, class, height from
# the txt, original rate [Hz], sino_fact, rotations, width, seed).
MEASURED_CASES = [
    ("pr_f1", "PR", 800, 80.0, 1, 2, 500, 11),
    ("pr_f2", "PR", 800, 80.0, 2, 2, 582, 12),
    ("in_80", "IN", 800, 80.0, 1, 2, 500, 13),
    ("in_300", "IN", 600, 300.0, 1, 2, 500, 14),
    ("in_f2", "IN", 800, 80.0, 2, 3, 582, 15),
]


def measured_inputs(inpr, H, rate, rot, seed):
    """(air file [6][800][609], detector file contents, the rows used)."""
    rng = np.random.default_rng(seed)
    air = rng.uniform(0.5, 1.5, (6, 800, 609)).astype(np.float32)
    if inpr == "PR":
        det = rng.uniform(0.0, 1.2, (rot, 800, 609)).astype(np.float32)
        return air, det, det
    step = 5 if rate >= 200 else 1
    det = rng.uniform(0.0, 0.9, (rot * H * step, 609)).astype(np.float32)
    return air, det, det[::step].reshape(rot, H, 609).copy()
