"""Pin the CPU oracle (oracle/vct_oracle.c) before trusting it.

* known-answer outputs shipped with the reference (KAT-M, KAT-F): bit exact;
* golden vectors written by the reference compiled from source
  (tests/golden/make_golden.py): tables, per-ray segment lists, segment
  counters, small-geometry sinograms and FBP images, filter-matrix rows;
* what the mathematics fixes where those leave the FP64 paths open:
  analytic chords of a uniform disk, Beer-Lambert for a single bin,
  finite differences of the FP64 forward model for the FP64 adjoint,
  published Philox4x32-10 vectors.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np
import pytest

from conftest import HELICAL_CASES, HELICAL_N, MEASURED_CASES, fbp_in_input, measured_inputs, rel_l2, rel_max
from oracle.oracle import Geom

SMALL = {"n64": (64, 0.7, 80), "n128": (128, 0.35, 160)}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_philox_published_vectors(oracle):
    # Random123 known-answer vectors for philox4x32_10
    assert list(oracle.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_attenuation_tables_match_reference(oracle, golden, tmp_path):
    t = golden.npz("ref_tables.npz")
    for name, path in [("shipped_120kV", golden.path("spectrum_120kV.txt")),
                       ("shipped_6MV", golden.path("spectrum_6MV.txt"))] + \
            [(f"synthetic_{n}", golden.synthetic_spectrum(n, str(tmp_path))) for n in ("80kV", "120kV", "6MV")]:
        att, pdf, _ = oracle.attenuation(path, 6)
        np.testing.assert_array_equal(att, t[name + "_att"], err_msg=name)
        np.testing.assert_array_equal(pdf, t[name + "_pdf"], err_msg=name)


def test_attenuation_closed_form_terms(oracle, golden):
    # E >> 1 MeV: pair production is present for Z > 1; E < 1.19 MeV it is zero.
    att, pdf, e = oracle.attenuation(golden.path("spectrum_6MV.txt"), 6)
    assert pdf.sum() == pytest.approx(1.0, abs=1e-6)
    assert np.all(np.diff(att[3][e < 1.0]) < 0)  # mu/rho(O) falls with E below 1 MeV
    # Compton per electron is the same for all elements: mu/rho ~ Z/A at 1 MeV
    i = int(np.argmin(np.abs(e - 1.0)))
    ratio = att[:, i] / (np.array([1, 6, 7, 8, 15, 20]) / np.array([1, 12, 14, 16, 31, 40]))
    assert ratio.max() / ratio.min() < 1.02


def test_const_map_rows_match_reference(oracle, golden):
    c = golden.npz("ref_constmap_rows.npz")
    for filt in (1, 2):
        Cm = oracle.const_map(Geom(), filt)
        np.testing.assert_array_equal(Cm[c["rows"]], c[f"filter{filt}"])


def test_segment_lists_match_reference(oracle, golden):
    s = golden.npz("ref_segments.npz")
    g = Geom()
    tables = (oracle.view_tables(g), oracle.channel_table(g))
    for i, ray in enumerate(s["rays"]):
        v, c = divmod(int(ray), g.W)
        vox, rr, iks, idx = oracle.trace_ray(g, v, c, tables)
        assert iks == s["count"][i]
        ref_vox = s["vox"][i, :iks]
        ref_len = s["length"][i, :iks]
        got_vox = np.zeros(iks, np.int32)
        got_len = np.zeros(iks, np.float32)
        got_vox[idx - 1] = vox
        got_len[idx - 1] = rr.astype(np.float32)
        np.testing.assert_array_equal(got_vox, ref_vox, err_msg=f"ray {ray}")
        np.testing.assert_array_equal(got_len, ref_len, err_msg=f"ray {ray}")


def test_segment_counts_full_geometry(oracle):
    cnt, total = oracle.segment_counts(Geom())
    assert total == 282_824_404  # sum of the reference's reprojection_sum
    assert cnt.max() == 1015
    assert int((cnt == 0).sum()) == 1344


@pytest.mark.parametrize("case", list(SMALL))
def test_small_sinograms_match_reference(oracle, golden, case):
    t = golden.npz("ref_small.npz")
    N, pix, V = SMALL[case]
    g = Geom(N=N, pix=pix, V=V)
    dens = t[case + "_phantom"].astype(np.float32) / np.float32(1000)
    raw = oracle.forward_ref(g, dens, t["att"], t["pdf"])
    np.testing.assert_array_equal(raw, t[case + "_raw"])
    cnt, _ = oracle.segment_counts(g)
    np.testing.assert_array_equal(cnt, t[case + "_counts"])
    norm = (raw / raw.max()).astype(np.float32)
    img = oracle.fbp_ref(g, norm, 2)
    np.testing.assert_array_equal(img, t[case + "_img"])
    f64 = oracle.forward_f64(g, dens, t["att"], t["pdf"])
    assert rel_max(f64, raw) < 1e-6


def test_kat_f_bit_exact(oracle, golden):
    img = oracle.fbp_ref(Geom(), golden.kat_f_input(), 2)
    assert sha(img) == golden.hashes()["kat_f_output_sha256"]


def test_kat_m_bit_exact(oracle, golden):
    """120 kV ICRU110: normalised noise-free sinogram equals the reference's,
    and with the glibc noise stream it equals the shipped file bit for bit."""
    g = Geom()
    dens = golden.icru110().astype(np.float32) / np.float32(1000)
    att, pdf, _ = oracle.attenuation(golden.path("spectrum_120kV.txt"), 6)
    raw = oracle.forward_ref(g, dens, att, pdf)
    norm, mx = oracle.normalize(raw)
    np.testing.assert_array_equal(norm, golden.sigma0_120kv())
    noisy = oracle.noise_glibc(norm, golden.hashes()["kat_m_sigma"], order=0)
    assert sha(noisy) == golden.hashes()["kat_m_output_sha256"]


def test_fbp_f64_close_to_reference(oracle, golden):
    g = Geom()
    p = golden.kat_f_input()
    ref = oracle.fbp_ref(g, p, 2)
    f64 = oracle.fbp_f64(g, p, 2)
    assert rel_max(ref, f64) < 5e-6
    assert rel_l2(ref, f64) < 2e-6


def test_uniform_disk_chords(oracle):
    """M=1, one bin, mu=1: -ln I is the chord of the disk within the
    voxelisation error, and 1 for rays that miss it."""
    N, pix = 128, 0.35
    g = Geom(N=N, pix=pix, V=40)
    c = (np.arange(N) - N / 2 + 0.5) * pix
    R = 15.0
    disk = ((c[None, :] ** 2 + c[:, None] ** 2) <= R * R).astype(np.float32)
    att = np.ones((1, 1), np.float32)
    pdf = np.ones(1, np.float32)
    I = oracle.forward_f64(g, disk[None], att, pdf)
    line = -np.log(I)
    cosT, sinT = oracle.view_tables(g)
    px0, py0 = oracle.channel_table(g)
    v = np.arange(g.V)[:, None]
    sx, sy = g.sid * sinT[v], g.sid * cosT[v]
    px = px0[None, :] * cosT[v] + py0[None, :] * sinT[v]
    py = -px0[None, :] * sinT[v] + py0[None, :] * cosT[v]
    dx, dy = px - sx, py - sy
    dist = np.abs(sx * dy - sy * dx) / np.hypot(dx, dy)  # distance of the ray from the centre
    chord = 2 * np.sqrt(np.clip(R * R - dist ** 2, 0, None))
    inner = dist < R - 2 * pix
    assert np.abs(line[inner] - chord[inner]).max() < 2.5 * pix
    assert np.abs(line[dist > R + 2 * pix]).max() == 0.0
    assert abs(np.mean(line[inner] - chord[inner])) < 0.05 * pix * 10


def test_single_bin_beer_lambert(oracle):
    """One energy bin: I = pdf * exp(-sum_m mu_m L_m) exactly (FP64 path)."""
    g = Geom(N=64, pix=0.7, V=20)
    rng = np.random.default_rng(3)
    dens = rng.uniform(0, 1, (3, 64, 64)).astype(np.float32)
    att = np.array([[0.2], [0.1], [0.05]], np.float32)
    pdf = np.array([0.5], np.float32)
    I, L = oracle.forward_f64(g, dens, att, pdf, want_L=True)
    expect = 0.5 * np.exp(-(L * att[:, 0].astype(np.float64)).sum(-1))
    np.testing.assert_allclose(I, expect, rtol=1e-14, atol=0)
    # linearity of the path integrals in the densities
    _, L2 = oracle.forward_f64(g, 2 * dens, att, pdf, want_L=True)
    np.testing.assert_allclose(L2, 2 * L, rtol=1e-12, atol=1e-12)


def test_adjoint_f64_finite_differences(oracle):
    """d/drho of sum_rays gI * I(rho) by central differences of the FP64
    forward model equals the FP64 adjoint."""
    g = Geom(N=32, pix=1.0 * 0.7, V=16)
    rng = np.random.default_rng(5)
    dens = rng.uniform(0.1, 1.0, (2, 32, 32)).astype(np.float32)
    att = np.array([[0.3, 0.2, 0.1], [0.15, 0.12, 0.1]], np.float32)
    pdf = np.array([0.2, 0.5, 0.3], np.float32)
    gI = rng.normal(size=(g.V, g.W))
    grad = oracle.adjoint_f64(g, dens, att, pdf, gI)
    h = 1e-3
    for (m, y, x) in [(0, 16, 16), (1, 10, 20), (0, 5, 27), (1, 25, 7)]:
        dp = dens.copy(); dp[m, y, x] += h
        dm = dens.copy(); dm[m, y, x] -= h
        fd = ((gI * oracle.forward_f64(g, dp, att, pdf)).sum() - (gI * oracle.forward_f64(g, dm, att, pdf)).sum())
        fd /= (np.float64(dp[m, y, x]) - np.float64(dm[m, y, x]))
        assert fd == pytest.approx(grad[m, y, x], rel=1e-5, abs=1e-9)


def test_noise_order_matters(oracle, golden):
    """The shipped noise stream fixes the draw order (log term first); the
    other order differs, so the KAT pins it."""
    x = np.full(16, 0.5, np.float32)
    a = oracle.noise_glibc(x, 0.01, order=0)
    b = oracle.noise_glibc(x, 0.01, order=1)
    assert not np.array_equal(a, b)
    assert np.array_equal(oracle.noise_glibc(x, 0.0, 0), x)


@pytest.mark.parametrize("alpha,beta", [(0.01, 2.0), (0.01, 3.7)])
def test_beam_hardening_matches_reference_tool(oracle, golden, alpha, beta):
    out, mn0, mn1 = oracle.beam_hardening(golden.sigma0_120kv(), alpha, beta)
    key = f"bhc_sigma0_a{alpha}_b{beta}"
    assert sha(out) == golden.hashes()[key + "_sha256"]
    assert f"{mn0:f} {mn1:f}" == golden.hashes()[key + "_stdout"]


NPF_SMALL = [("n64", 64, 0.7, 80), ("n128", 128, 0.35, 160)]
NPF_WINDOWS = [(0.4, 0.8), (1.3, 0.8), (2.0, 1.2)]


def npf_input(V, seed):
    # the inputs tests/golden/make_golden.py gave the reference
    return np.random.default_rng(seed).uniform(0.05, 1.0, (V, 609)).astype(np.float32)


def test_lengthproj_matches_reference_small(oracle, golden):
    t = golden.npz("ref_npf.npz")
    for k, (name, N, pix, V) in enumerate(NPF_SMALL):
        g = Geom(N=N, pix=pix, V=V)
        proj = npf_input(V, 100 + k)
        for z, couch in NPF_WINDOWS:
            np.testing.assert_array_equal(oracle.lengthproj_ref(g, proj, z, couch), t[f"{name}_z{z}_c{couch}"])


def test_lengthproj_matches_reference_full(oracle, golden):
    """npf of the noise-free 120 kV sinogram at 512 x 512 / 800 x 609, three
    view windows: bit-identical to the reference (SHA-256 of the image)."""
    sino = golden.sigma0_120kv()
    h = golden.hashes()
    for z, couch in NPF_WINDOWS:
        npf = oracle.lengthproj_ref(Geom(), sino, z, couch)
        assert sha(npf) == h[f"npf_sigma0_z{z}_c{couch}_sha256"]


def test_lengthproj_window_must_start_at_view_zero(oracle):
    g = Geom(N=64, pix=0.7, V=80)
    proj = npf_input(80, 1)
    assert oracle.lengthproj_ref(g, proj, 0.0, 0.8) is None   # starts at view -40
    assert oracle.lengthproj_ref(g, proj, 0.3, 0.0) is None   # no couch motion: 0/0
    assert oracle.lengthproj_ref(g, proj, 0.4, 0.8) is not None


def test_lengthproj_closed_form(oracle):
    """proj[j][c] = c makes every sample the channel position itself, so
    npf = sum_views chord * channel with the channel from the ray angle,
    written here from the geometry alone: theta = atan(|x'| / (sid - y'))
    in the frame where the source sits on the y' axis, channel = iso -/+
    theta / (max_alpha / 304) on the side of x'; chord = pix * sec of the
    ray's angle to the major axis when |dx| > |dy|, else pix (the
    reference's y-branch reuses the near end, M/lengthproj_TOMO.cpp:98)."""
    g = Geom(N=64, pix=0.7, V=80)
    V, N = g.V, g.N
    proj = np.tile(np.arange(609, dtype=np.float32), (V, 1))
    z, couch = 1.3, 0.8
    got = oracle.lengthproj_ref(g, proj, z, couch).astype(np.float64)
    start = int(z / (couch / V) - V // 2 + 1e-8)
    views = (start % V + np.arange(V)) % V
    beta = (np.arange(V) * 360.0 / V) * np.pi / 180.0
    c, s = np.cos(beta[views]), np.sin(beta[views])
    jy, jx = np.mgrid[0:N, 0:N]
    ry = N * g.pix / 2 - (jy + 0.5) * g.pix
    rx = -N * g.pix / 2 + (jx + 0.5) * g.pix
    mask = oracle.mask(N).astype(bool)
    sx, sy = g.sid * s[:, None, None], g.sid * c[:, None, None]
    ddx, ddy = rx[None] - sx, ry[None] - sy
    chord = np.where(np.abs(ddx) > np.abs(ddy), g.pix * np.sqrt(1 + (ddy / ddx) ** 2), g.pix)
    xl = rx[None] * c[:, None, None] - ry[None] * s[:, None, None]
    yl = rx[None] * s[:, None, None] + ry[None] * c[:, None, None]
    theta = np.arctan(np.abs(xl) / (g.sid - yl))
    ch = np.where(xl > 0, 304.5 - theta / (g.max_alpha / 304), 304.5 + theta / (g.max_alpha / 304))
    expect = (chord * np.clip(ch, 0, 608)).sum(0) * mask
    assert np.all(got[~mask] == 0)
    assert rel_max(got, expect) < 2e-6
    # a uniform sinogram gives the summed chord lengths
    ones = oracle.lengthproj_f64(g, np.ones((V, 609), np.float32), z, couch)
    np.testing.assert_allclose(ones, chord.sum(0) * mask, rtol=1e-12, atol=0)


# ------------------------------------------------------------ measured data

@pytest.mark.parametrize("case", MEASURED_CASES, ids=[c[0] for c in MEASURED_CASES])
def test_measured_loader_matches_reference(oracle, golden, case):
    """Air-scan normalisation (PR and IN, sino_fact 1 and 2, 80 and 300 Hz):
    volume and mask bit-identical to the reference's loadProjectionData
    (SHA-256 from tests/golden/make_golden.py)."""
    name, inpr, H_txt, rate, nskip, rot, W, seed = case
    H = H_txt // nskip
    air, _, raw = measured_inputs(inpr, H, rate, rot, seed)
    vol, mask, mx = oracle.load_measured(int(inpr == "IN"), rot, H, W, nskip, air, raw)
    h = golden.hashes()
    assert sha(vol) == h[f"measured_{name}_volume_sha256"]
    assert sha(mask) == h[f"measured_{name}_mask_sha256"]
    if inpr == "IN":
        # values below 0.1 are zeroed, so the largest value survives; only
        # view 0 of each rotation holds data
        assert mx == float(vol.max()) and 0.1 < mx <= 1.0
        assert not vol.reshape(rot, H, W)[:, 1:].any() and not mask.reshape(rot, H, W)[:, 1:].any()


def test_measured_loader_rejects_out_of_bounds_shapes(oracle):
    air = np.ones((6, 800, 609), np.float32)
    raw = np.ones((1, 800, 609), np.float32)
    assert oracle.load_measured(0, 1, 800, 583, 1, air, raw) is None   # 583 + 27 > 609 channels
    assert oracle.load_measured(0, 1, 800 // 3, 500, 3, air, raw) is None   # 267 kept rows > 266
    assert oracle.load_measured(1, 1, 801, 500, 1, air, np.ones((1, 801, 609), np.float32)) is None
    # air scans of 5 average to 1 + 1 + 1 + 1 + 1 = 5 exactly, so a PR value
    # is the detector value / 5, correctly rounded, at the kept rows
    air5 = np.full((6, 800, 609), 5.0, np.float32)
    raw = np.random.default_rng(4).uniform(0, 2, (1, 800, 609)).astype(np.float32)
    vol, mask, _ = oracle.load_measured(0, 1, 400, 582, 2, air5, raw)
    np.testing.assert_array_equal(vol, raw[0, ::2, 27:609] / np.float32(5))
    assert np.all(mask == 1)


def test_helical_window_closed_forms(oracle):
    """One rotation: the window is the first H/2*2 views.  Rotations that are
    constant c_t: with two, only view 0 blends (weight 1/2, c0/2 + c1/2), all
    other views keep c0; with three, views n < H/2 are c0 + w (c1 - c0),
    w = 1/2 - n/H, and views n >= H/2 keep c0."""
    H, W = 10, 7
    rng = np.random.default_rng(3)
    one = rng.uniform(0, 1, (H, W)).astype(np.float32)
    np.testing.assert_array_equal(oracle.helical_window(one, 1, H, W), one)
    c = np.array([0.25, 0.75, 0.5], np.float32)
    for rot in (2, 3):
        vol = np.repeat(c[:rot], H * W).reshape(rot * H, W)
        win = oracle.helical_window(vol, rot, H, W).astype(np.float64)
        for n in range(H):
            w = 0.5 * abs(n - H / 2) / (H / 2)
            if rot == 2:
                expect = (c[0] + c[1]) / 2 if n == 0 else c[0]
            else:
                expect = c[0] + w * (c[1] - c[0]) if n < H // 2 else c[0]
            np.testing.assert_allclose(win[n], expect, rtol=2 ** -23, atol=0)


@pytest.mark.parametrize("case", HELICAL_CASES, ids=[c[0] for c in HELICAL_CASES])
def test_helical_npf_matches_reference_program(oracle, golden, case):
    """Loader, helical window and npf of the oracle on seeded measured files
    with couch motion (500 of the detector's 609 channels) reproduce
    npf_float.raw of the reference's whole MBFPA program bit for bit
    (SHA-256 written by make_golden.py from oracle/_ref/mvct_mbfpa_ref):
    pins the helical blend of M/main.cpp:352-371 and the npf's use of the
    detector's channel pitch for a narrower window."""
    name, inpr, rot, H, W, seed, z, couch = case
    air, det, raw = measured_inputs(inpr, H, 80.0, rot, seed)
    vol, _, _ = oracle.load_measured(int(inpr == "IN"), rot, H, W, 1, air, raw)
    win = oracle.helical_window(vol, rot, H, W)
    npf = oracle.lengthproj_ref(Geom(N=HELICAL_N, pix=0.1), win, z, couch)
    assert npf is not None and np.count_nonzero(npf) > 1000
    assert hashlib.sha256(npf.tobytes()).hexdigest() == golden.hashes()[f"helical_{name}_npf_sha256"]
    # a window as wide as the detector gives another image: the pitch is the detector's
    wide = oracle.lengthproj_ref(Geom(N=HELICAL_N, pix=0.1, W=W), win, z, couch)
    assert not np.array_equal(wide, npf)


def test_fbp_in_mode_matches_reference_program(oracle, golden):
    """The FBP program's "IN" path: values below 50000 read as 1
    (F/projectionDataTOMO.cpp:274-292), then the filter keeps values in
    [1000, 300000] (F/filterSinogramTOMO.cpp:250-255).  The oracle equals the
    reference program's image bit for bit (SHA-256 from oracle/_ref/mvct_fbp_ref)."""
    p = fbp_in_input()
    cut = np.where(p < 50000.0, np.float32(1.0), p)
    img = oracle.fbp_ref(Geom(), cut, 2, in_mode=True)
    assert np.count_nonzero(img) > 100000
    assert hashlib.sha256(img.tobytes()).hexdigest() == golden.hashes()["fbp_in_mode_sha256"]
