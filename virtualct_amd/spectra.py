"""X-ray spectra and attenuation tables.

* ``write_spectrum`` writes the reference's spectrum text format
  (M/Spectrum/PhaseSpace_photon_*.txt: one header line, then
  ``Elo Ehi weight`` per bin in MeV).
* ``kramers_kv`` / ``linac_6mv`` build the synthetic spectra of the
  benchmark configurations.
* ``attenuation`` runs the reference's table builder through the C ABI
  (``vct_attenuation_coefficient``, restating
  M/prior_weight_production.cpp:180-297) and returns att [M,NE], pdf [NE].
* ``closed_form_attenuation`` is the alternative mu/rho model named by the
  benchmark configurations (photoelectric ~ Z^3.8 E^-3.2, Klein-Nishina
  Compton ~ Z/A, pair production ~ Z^2 (E - 1.022 MeV)); it is an optional
  second table generator, applied identically on every path.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

from . import _lib

ELEMENTS = ("H", "C", "N", "O", "P", "Ca", "Mg", "Si")
ELEMENT_Z = (1, 6, 7, 8, 15, 20, 12, 14)
ELEMENT_A = (1, 12, 14, 16, 31, 40, 24, 28)


def write_spectrum(path: str, lo_mev, hi_mev, weights) -> str:
    lo, hi, w = (np.asarray(x, np.float64) for x in (lo_mev, hi_mev, weights))
    with open(path, "w") as f:
        f.write("1\n")
        for a, b, c in zip(lo, hi, w):
            f.write(f"{a:.4f} {b:.4f} {c:.9e}\n")
    return path


def kramers_kv(kvp: float, e_min_kev: float = 10.0, bin_kev: float = 1.0):
    """Kramers-shaped bremsstrahlung N(E) ~ (kVp - E)/E at 1-keV bin centres,
    bins from e_min to kVp (110 bins for 120 kV, 70 for 80 kV)."""
    n = int(round((kvp - e_min_kev) / bin_kev))
    lo = (e_min_kev + bin_kev * np.arange(n)) / 1000.0
    hi = lo + bin_kev / 1000.0
    e = (lo + hi) * 0.5 * 1000.0
    w = np.clip((kvp - e) / e, 0.0, None)
    return lo, hi, w


def linac_6mv(n_bins: int = 120, width_mev: float = 0.05, e0_mev: float = 0.05):
    """6 MV bremsstrahlung-like spectrum, ~(6 - E)/E over n_bins x 50 keV
    starting at 0.05 MeV (bin centres 0.075 .. 6.025 MeV)."""
    lo = e0_mev + width_mev * np.arange(n_bins)
    hi = lo + width_mev
    e = (lo + hi) * 0.5
    w = np.clip((6.0 - e) / e, 0.0, None)
    return lo, hi, w


SYNTHETIC = {
    "80kV": lambda: kramers_kv(80.0),
    "120kV": lambda: kramers_kv(120.0),
    "6MV": lambda: linac_6mv(),
}


def synthetic_spectrum_file(name: str, directory: str | None = None) -> str:
    directory = directory or tempfile.mkdtemp(prefix="vct_spec_")
    path = os.path.join(directory, f"PhaseSpace_photon_{name}_synthetic.txt")
    write_spectrum(path, *SYNTHETIC[name]())
    return path


def attenuation(spectrum_path: str, M: int = 6, nthin: int = 1, Z=None, A=None):
    """Reference table builder (host side of the C ABI): att [M,NE], pdf [NE], E [NE] (MeV)."""
    att = np.zeros(M * 1024, np.float32)
    pdf = np.zeros(1024, np.float32)
    en = np.zeros(1024, np.float64)
    ne = C.c_int(0)
    zp = ap = None
    if Z is not None:
        zp = np.ascontiguousarray(Z, np.float64)
        ap = np.ascontiguousarray(A, np.float64)
    _lib.call("vct_attenuation_coefficient", spectrum_path.encode(), M,
              C.c_void_p(zp.ctypes.data if zp is not None else 0),
              C.c_void_p(ap.ctypes.data if ap is not None else 0), nthin,
              C.c_void_p(att.ctypes.data), C.c_void_p(pdf.ctypes.data), C.byref(ne),
              C.c_void_p(en.ctypes.data))
    n = ne.value
    return att[: M * n].reshape(M, n).copy(), pdf[:n].copy(), en[:n].copy()


def closed_form_attenuation(energy_mev, M: int = 6):
    """mu/rho [M,NE] (cm^2/g) from the closed form of the benchmark
    configurations.  The photoelectric and pair constants are matched to the
    reference model for oxygen at 60 keV and 6 MeV."""
    E = np.asarray(energy_mev, np.float64)
    Z = np.asarray(ELEMENT_Z[:M], np.float64)[:, None]
    A = np.asarray(ELEMENT_A[:M], np.float64)[:, None]
    re2 = 2.81794 ** 2 * 1e-3  # classical electron radius^2 in barn-scaled units
    eps = E[None, :] / 0.511
    kn = 2 * np.pi * ((1 + eps) / eps ** 2 * (2 * (1 + eps) / (1 + 2 * eps) - np.log1p(2 * eps) / eps)
                      + np.log1p(2 * eps) / (2 * eps) - (1 + 3 * eps) / (1 + 2 * eps) ** 2)
    compton = 6.0221 * Z / A * re2 * kn
    # photoelectric ~ Z^3.8 E^-3.2 and pair ~ Z^2 (E - 1.022), per atom -> per gram
    c_pe = 6.0221 * re2 * 3.45e-6 * (1 + 0.008 * 8) * 8 ** 3 / (0.06 / 0.511) ** 3 / (8 ** 3.8 * 0.06 ** -3.2)
    photo = 6.0221 / A * c_pe * Z ** 3.8 * E[None, :] ** -3.2
    c_pp = 6.0221 * 8 / 16 * re2 * 0.2545 * (6.0 / 0.511 - 2.332) * 8 / 137 / (8 ** 2 / 16 * (6.0 - 1.022))
    pair = c_pp * Z ** 2 / A * np.clip(E[None, :] - 1.022, 0.0, None)
    return (photo + compton + pair).astype(np.float32)
