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


# Measured files with couch motion through the reference's whole MBFPA program:
# (name, class, rotations, views per rotation, channels, seed, z [cm], couch [cm/rotation]).
# The image is 256^2 x 0.1 cm: every masked pixel's ray then meets the detector
# at channels 164-445, inside the 500 measured channels.  (The reference's npf
# interpolates channels up to its compile-time End_Channel, 608, with the
# measured width as row stride, so for larger images it reads the next row and,
# on the last rows, past its buffer.)
HELICAL_N = 256
HELICAL_CASES = [
    ("pr_rot3", "PR", 3, 800, 500, 40, 1.3, 0.8),
    ("in_rot2", "IN", 2, 800, 500, 42, 1.3, 0.8),
]


def helical_txt(template: str, inpr: str, width: int, height: int, z: float, couch: float, N: int = 512) -> str:
    """The reference's MBFPA input txt with measured files det.raw / air.raw,
    noise off, couch motion, the reconstruction start z and an N x N image."""
    t = template
    reps = [
        ("// reconstruction size (square format) \n512 ", f"// reconstruction size (square format) \n{N} "),
        ("// Energy Label (IN (for treatment beam) or PR (for CT))\nPR",
         f"// Energy Label (IN (for treatment beam) or PR (for CT))\n{inpr}"),
        ("// sinogram width \n609", f"// sinogram width \n{width}"),
        ("// sinogram height \n800", f"// sinogram height \n{height}"),
        ("// sinogram image data for input \n*", "// sinogram image data for input \ndet.raw"),
        ("// Air scan image data for input // parameters are fixed\n*",
         "// Air scan image data for input // parameters are fixed\nair.raw"),
        ("//Couch speed [cm/rotation] // 0.4=fine, 0.8=normal, 1.2=coarse\n0.0",
         f"//Couch speed [cm/rotation] // 0.4=fine, 0.8=normal, 1.2=coarse\n{couch}"),
        ("// reconstruction start z [cm]\n0.0", f"// reconstruction start z [cm]\n{z}"),
        ("0.0000537", "0.0"),
    ]
    for a, b in reps:
        assert t.count(a) == 1, a
        t = t.replace(a, b)
    return t


def fbp_in_input(seed: int = 61) -> np.ndarray:
    """A seeded "IN" sinogram [800][609]: counts in [40000, 120000) with a
    smooth bump, so that part of it falls below the reference's 50000 cut
    (read as 1) and part of the filtered values pass its [1000, 300000] window."""
    rng = np.random.default_rng(seed)
    ch = np.arange(609)[None, :]
    v = np.arange(800)[:, None]
    base = 0.3 + 0.25 * np.exp(-((ch - 304.5 - 120 * np.sin(v * 2 * np.pi / 800)) / 90.0) ** 2)
    return (1.0 / base * 20000.0 + rng.uniform(0, 8000, (800, 609))).astype(np.float32)
