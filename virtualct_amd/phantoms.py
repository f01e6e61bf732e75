"""Seeded synthetic phantoms in the reference's density format.

The reference reads ``Weight_input_*.raw``: int16 [M][N][N], partial density
of each element in mg/cm^3 (value/1000 -> g/cm^3, M/main.cpp:93-109), row
index = -y.  These generators write the same format.

* ``body_phantom``: 6 elements (H, C, N, O, P, Ca); a 400 x 300 mm
  soft-tissue ellipse (rho 1.03), two lung ellipses (rho 0.26) and 3-6 bone
  ellipses (rho 1.4-1.9, Ca .15-.22, P .07-.10) at seeded positions.
* ``calibration_phantom``: 8 elements (adds Mg, Si); a 330 mm water cylinder
  with 16 cylindrical inserts of 28 mm diameter on radii 55 and 105 mm
  (after the Gammex layout of M/phantom_creation_Gammex.cpp:246-319), rho
  0.3-1.8 and seeded elemental fractions.
"""
from __future__ import annotations

import numpy as np

SOFT = np.array([0.102, 0.143, 0.034, 0.708, 0.002, 0.011])  # H C N O P Ca


def _grid(N: int, pix_cm: float):
    c = (np.arange(N) - N / 2 + 0.5) * pix_cm
    x = c[None, :]
    y = -c[:, None]  # row index = -y
    return x, y


def _ellipse(x, y, cx, cy, ax, by, ang=0.0):
    ca, sa = np.cos(ang), np.sin(ang)
    u = (x - cx) * ca + (y - cy) * sa
    v = -(x - cx) * sa + (y - cy) * ca
    return (u / ax) ** 2 + (v / by) ** 2 <= 1.0


def _to_int16(rho_w: np.ndarray) -> np.ndarray:
    return np.round(1000.0 * rho_w).astype(np.int16)


def body_phantom(seed: int = 1000, N: int = 512, pix_cm: float | None = None) -> np.ndarray:
    """int16 [6,N,N] partial densities (mg/cm^3)."""
    pix_cm = 51.2 / N if pix_cm is None else pix_cm
    rng = np.random.Generator(np.random.PCG64(seed))
    x, y = _grid(N, pix_cm)
    rho = np.zeros((N, N))
    frac = np.zeros((6, N, N))
    body = _ellipse(x, y, 0.0, 0.0, 20.0, 15.0)
    rho[body] = 1.03
    frac[:, body] = SOFT[:, None]
    for sgn in (-1, 1):
        cx = sgn * rng.uniform(6.0, 9.0)
        lung = _ellipse(x, y, cx, rng.uniform(-1.0, 2.0), rng.uniform(4.0, 6.0), rng.uniform(6.0, 9.0),
                        rng.uniform(-0.3, 0.3)) & body
        rho[lung] = 0.26
    nb = int(rng.integers(3, 7))
    for _ in range(nb):
        r = rng.uniform(1.0, 3.0)
        cx, cy = rng.uniform(-14.0, 14.0), rng.uniform(-10.0, 10.0)
        bone = _ellipse(x, y, cx, cy, r * rng.uniform(0.7, 1.3), r * rng.uniform(0.7, 1.3),
                        rng.uniform(0, np.pi)) & body
        ca, p = rng.uniform(0.15, 0.22), rng.uniform(0.07, 0.10)
        rest = SOFT[:4] / SOFT[:4].sum() * (1.0 - ca - p)
        w = np.concatenate([rest, [p, ca]])
        rho[bone] = rng.uniform(1.4, 1.9)
        frac[:, bone] = w[:, None]
    return _to_int16(rho[None] * frac)


def calibration_phantom(seed: int = 2000, N: int = 512, pix_cm: float | None = None) -> np.ndarray:
    """int16 [8,N,N] partial densities (mg/cm^3): H C N O P Ca Mg Si."""
    pix_cm = 51.2 / N if pix_cm is None else pix_cm
    rng = np.random.Generator(np.random.PCG64(seed))
    x, y = _grid(N, pix_cm)
    dens = np.zeros((8, N, N))
    water = (x ** 2 + y ** 2) <= 16.5 ** 2
    dens[0][water] = 1.0 * 0.1119
    dens[3][water] = 1.0 * 0.8881
    k = 0
    for radius, count in ((5.5, 8), (10.5, 8)):
        for i in range(count):
            ang = 2 * np.pi * i / count + (np.pi / count if radius > 6 else 0.0)
            cx, cy = radius * np.cos(ang), radius * np.sin(ang)
            ins = ((x - cx) ** 2 + (y - cy) ** 2) <= 1.4 ** 2
            rho = rng.uniform(0.3, 1.8)
            w = rng.dirichlet(np.array([2.0, 3.0, 1.0, 6.0, 0.5, 1.0, 0.5, 0.5]))
            dens[:, ins] = (rho * w)[:, None]
            k += 1
    return _to_int16(dens)


def write_raw(path: str, phantom: np.ndarray) -> str:
    np.ascontiguousarray(phantom, np.int16).tofile(path)
    return path


def densities_f32(phantom_i16: np.ndarray) -> np.ndarray:
    """(float)s/1000 exactly as M/main.cpp:103."""
    return phantom_i16.astype(np.float32) / np.float32(1000)
