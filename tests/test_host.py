"""Host side of libvct on the CPU: the library loads and exports every symbol
declared in include/vct.h, the host table builders match the reference's
golden vectors, argument errors come back as status codes, and the
synthetic phantom/spectrum generators are deterministic."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle.oracle import Geom


def header_symbols():
    text = open(os.path.join(ROOT, "include", "vct.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vct_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from virtualct_amd import _lib
    L = _lib.lib()
    missing = [s for s in header_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert L.vct_abi_version() == 1
    assert set(header_symbols()) >= set(_lib.exported_symbols())


def test_geometry_defaults_are_reference_constants():
    from virtualct_amd import default_geometry
    g = default_geometry()
    assert (g.recon_size, g.n_channels, g.n_views) == (512, 609, 800)
    assert (g.pixel_cm, g.sid_cm, g.det_radius_cm, g.det_src_iso_cm) == (0.1, 85.0, 142.6, 85.0)
    assert (g.iso_channel, g.max_alpha, g.max_beta, g.xfactor) == (304.5, 0.326086956, 0.326086956, 1.0)


def test_attenuation_matches_reference_tables(golden, tmp_path):
    from virtualct_amd import spectra
    t = golden.npz("ref_tables.npz")
    for name, path in [("shipped_120kV", golden.path("spectrum_120kV.txt")),
                       ("shipped_6MV", golden.path("spectrum_6MV.txt"))] + \
            [(f"synthetic_{n}", golden.synthetic_spectrum(n, str(tmp_path))) for n in ("80kV", "120kV", "6MV")]:
        att, pdf, _ = spectra.attenuation(path, 6)
        np.testing.assert_array_equal(att, t[name + "_att"], err_msg=name)
        np.testing.assert_array_equal(pdf, t[name + "_pdf"], err_msg=name)


def test_attenuation_eight_elements_matches_oracle(oracle, golden):
    from virtualct_amd import spectra
    for f in ("spectrum_120kV.txt", "spectrum_6MV.txt"):
        att, pdf, e = spectra.attenuation(golden.path(f), 8)
        a2, p2, e2 = oracle.attenuation(golden.path(f), 8)
        np.testing.assert_array_equal(att, a2)
        np.testing.assert_array_equal(pdf, p2)
        np.testing.assert_array_equal(e, e2)
    # thinning by 2 keeps the normalisation and halves the bins
    att2, pdf2, _ = spectra.attenuation(golden.path("spectrum_120kV.txt"), 6, nthin=2)
    a3, p3, _ = oracle.attenuation(golden.path("spectrum_120kV.txt"), 6, nthin=2)
    assert att2.shape == (6, 50)
    np.testing.assert_array_equal(att2, a3)
    np.testing.assert_array_equal(pdf2, p3)


def test_const_map_matches_reference_rows(golden):
    from virtualct_amd import _lib, default_geometry
    c = golden.npz("ref_constmap_rows.npz")
    g = default_geometry()
    for filt in (1, 2):
        Cm = np.zeros((609, 609), np.float64)
        _lib.call("vct_const_map", C.byref(g), filt, C.c_void_p(Cm.ctypes.data))
        np.testing.assert_array_equal(Cm[c["rows"]], c[f"filter{filt}"])


def test_status_codes_not_exits(tmp_path):
    from virtualct_amd import _lib, default_geometry
    L = _lib.lib()
    att = np.zeros(8 * 1024, np.float32)
    pdf = np.zeros(1024, np.float32)
    ne = C.c_int()
    st = L.vct_attenuation_coefficient(str(tmp_path / "missing.txt").encode(), 6, None, None, 1,
                                       att.ctypes.data, pdf.ctypes.data, C.byref(ne), None)
    assert st == 4  # VCT_ERR_IO
    assert b"missing.txt" in L.vct_last_error()
    st = L.vct_attenuation_coefficient(b"x", 0, None, None, 1, att.ctypes.data, pdf.ctypes.data, C.byref(ne), None)
    assert st == 1
    g = default_geometry()
    Cm = np.zeros((609, 609))
    assert L.vct_const_map(C.byref(g), 3, Cm.ctypes.data) == 1
    g.recon_size = 1
    assert L.vct_const_map(C.byref(g), 2, Cm.ctypes.data) == 1
    with pytest.raises(_lib.VctError):
        _lib.call("vct_const_map", C.byref(g), 2, C.c_void_p(Cm.ctypes.data))


def test_init_without_gpu_reports_error():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is visible")
    from virtualct_amd import _lib
    ctx = C.c_void_p()
    st = _lib.lib().vct_init(0, C.byref(ctx))
    assert st != 0 and not ctx.value


def test_phantoms_deterministic_and_in_format():
    from virtualct_amd import phantoms
    a = phantoms.body_phantom(1000)
    b = phantoms.body_phantom(1000)
    c = phantoms.body_phantom(1001)
    assert a.dtype == np.int16 and a.shape == (6, 512, 512)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # soft tissue: rho 1.03 with the ICRU fractions
    centre = a[:, 256, 256].astype(np.float64) / 1000
    assert centre.sum() == pytest.approx(1.03, abs=5e-3) or centre.sum() == pytest.approx(0.26, abs=5e-3) \
        or centre.sum() > 1.3
    cal = phantoms.calibration_phantom(2000)
    assert cal.shape == (8, 512, 512)
    water = cal[:, 256 - 150, 256].astype(np.float64) / 1000
    assert water[0] == pytest.approx(0.112, abs=1e-3) and water[3] == pytest.approx(0.888, abs=1e-3)
    for N in (256, 384, 768):
        p = phantoms.body_phantom(5, N=N, pix_cm=50.0 / N)
        assert p.shape == (6, N, N)


def test_spectrum_writer_roundtrip(tmp_path, oracle):
    from virtualct_amd import spectra
    for name, nbins in (("80kV", 70), ("120kV", 110), ("6MV", 120)):
        path = spectra.synthetic_spectrum_file(name, str(tmp_path))
        att, pdf, e = spectra.attenuation(path, 6)
        assert att.shape == (6, nbins)
        assert pdf.sum() == pytest.approx(1.0, abs=1e-6)
        a2, p2, _ = oracle.attenuation(path, 6)
        np.testing.assert_array_equal(att, a2)
    cf = spectra.closed_form_attenuation(np.array([0.06, 6.0]), 8)
    assert cf.shape == (8, 2) and np.all(cf > 0)


def test_batched_row_layout():
    """The batched kernels' rows (8*B floats): materials 0-3 of phantom b at
    4b, 4-5 at 4B+2b, 6-7 at 6B+2b; B = 1 keeps the plain [8]; interleave /
    deinterleave and to_planes / from_planes invert each other."""
    import torch
    from virtualct_amd import batch
    R, B = 3, 5
    nat = torch.arange(R * B * 8, dtype=torch.float32).view(R, B, 8)
    flat = batch.to_planes(nat).view(R, 8 * B)
    for r in range(R):
        for b in range(B):
            for m in range(8):
                pos = 4 * b + m if m < 4 else (4 * B + 2 * b + m - 4 if m < 6 else 6 * B + 2 * b + m - 6)
                assert flat[r, pos] == nat[r, b, m]
    assert torch.equal(batch.from_planes(batch.to_planes(nat)), nat)
    one = torch.randn(7, 1, 8)
    assert torch.equal(batch.to_planes(one), one)
    dens = torch.rand(B, 6, 4, 4)
    rows = batch.interleave(dens)
    assert torch.all(batch.from_planes(rows)[..., 6:] == 0)
    assert torch.equal(batch.deinterleave(rows, 6, 4), dens)


def test_tuner_keeps_the_fastest_and_caches_it(tmp_path):
    """The tuner times every candidate in one untimed pass and three
    interleaved timed passes, keeps the one with the best time, writes it to
    its cache file and answers later asks (and new processes) from there."""
    from virtualct_amd.autotune import Tuner
    path = str(tmp_path / "tune.json")
    t = Tuner(path)
    calls = []

    def time_fn(cfg):
        calls.append(cfg["x"])
        return {1: 3.0, 2: 1.0, 3: 2.0}[cfg["x"]]
    best = t.best(("k", 1), lambda: [{"x": 1}, {"x": 2}, {"x": 3}], time_fn)
    assert best == {"x": 2} and calls == [1, 2, 3] * 4
    assert t.best(("k", 1), lambda: [{"x": 1}], time_fn) == {"x": 2} and calls == [1, 2, 3] * 4
    # a candidate timed slow while clocks ramp up still wins on its best pass
    seq = {1: [9.0, 9.0, 1.5, 1.5], 2: [2.0, 2.0, 2.0, 2.0]}
    ramp = lambda cfg: seq[cfg["x"]].pop(0)
    assert t.best(("ramp",), lambda: [{"x": 1}, {"x": 2}], ramp) == {"x": 1}
    assert Tuner(path).best(("k", 1), lambda: [{"x": 1}], time_fn) == {"x": 2}
    assert t.best(("k", 2), lambda: [{"x": 3}, {"x": 1}], time_fn) == {"x": 3}
