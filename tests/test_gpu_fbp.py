"""GPU parity of fan-beam FBP (MFMA filter + pixel-driven backprojection)
against the oracle, the reference's golden images and the FP64 oracle."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2, rel_max
from oracle.oracle import Geom

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _ops():
    from virtualct_amd import ops
    return ops


def _geom(N=512, pix=0.1, V=800):
    from virtualct_amd import default_geometry
    return default_geometry(recon_size=N, pixel_cm=pix, n_views=V)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _exact_plan(geom):
    """A plan whose single-sinogram backprojection adds each pixel's views in
    the reference's order (not the quarter-turn orbits)."""
    from virtualct_amd import _lib
    plan = _ops().Plan(geom)
    plan.set_option(_lib.OPT_EXACT_BACKPROJECT, 1)
    return plan


@pytest.mark.parametrize("filt", [2, 1])
def test_filter_matches_oracle(gpu, golden, oracle, filt):
    ops = _ops()
    p = golden.kat_f_input()
    plan = ops.plan_for(_geom())
    fl = ops.filter_sinogram(plan, _t(p, gpu), filt)[0].cpu().numpy()
    ref = oracle.filter_ref(Geom(), p, filt)
    f64 = oracle.filter_f64(Geom(), p, filt)
    assert rel_max(fl, ref) < TOL
    assert rel_max(fl, f64) < TOL
    assert rel_l2(fl, f64) < TOL


def test_filter_in_mode_threshold(gpu, golden, oracle):
    ops = _ops()
    p = golden.kat_f_input()
    plan = ops.plan_for(_geom())
    fl = ops.filter_sinogram(plan, _t(p, gpu), 2, in_mode=True)[0].cpu().numpy()
    ref = oracle.filter_ref(Geom(), p, 2, in_mode=True)
    plain = oracle.filter_ref(Geom(), p, 2)
    # outside the thresholds both are exactly zero; elsewhere they agree
    near = (np.abs(plain - 1000.0) < 1.0) | (np.abs(plain - 300000.0) < 10.0)
    assert np.array_equal(fl[~near] == 0, ref[~near] == 0)
    both = (fl != 0) & (ref != 0)
    assert rel_max(fl[both], ref[both]) < TOL


def test_backprojection_matches_oracle(gpu, golden, oracle):
    """Same filtered input: the exact-order BP reproduces the reference's FP64
    geometry and float accumulation order (atan from the device math library)."""
    ops = _ops()
    fl = oracle.filter_ref(Geom(), golden.kat_f_input(), 2)
    plan = _exact_plan(_geom())
    img = ops.backproject(plan, _t(fl, gpu))[0].cpu().numpy()
    ref = oracle.backproject_ref(Geom(), fl)
    assert rel_max(img, ref) < 1e-6
    assert np.count_nonzero(img != ref) < 0.01 * img.size
    mask = oracle.mask(512).astype(bool)
    assert np.all(img[~mask] == 0)
    plan.close()


@pytest.mark.parametrize("N,pix,V", [(512, 0.1, 800), (128, 0.6, 160), (64, 0.7, 2100), (96, 0.5, 84)])
def test_quarter_turn_backprojection_matches_oracle(gpu, golden, oracle, N, pix, V):
    """The default single-sinogram BP shares each pixel-view geometry over the
    4 pixels of a quarter-turn orbit (views a multiple of 4, N even): within
    1e-5 of the reference-order image and of FP64 (512^2 on the KAT-F input;
    a wide field with the library atan; 2100 views; 84 views, 96^2), zero
    outside the mask, and within 1e-5 of the exact-order kernel (the float
    sums run in another order: measured 1.2e-6 at 800 views, 2.2e-6 at 2100)."""
    ops = _ops()
    g = Geom(N=N, pix=pix, V=V)
    if (N, V) == (512, 800):
        fl = oracle.filter_ref(Geom(), golden.kat_f_input(), 2)
    else:
        fl = np.random.default_rng(N + V).normal(size=(V, 609)).astype(np.float32) * 1e3
    img = ops.backproject(ops.plan_for(_geom(N, pix, V)), _t(fl, gpu))[0].cpu().numpy()
    ref = oracle.backproject_ref(g, fl)
    f64 = oracle.backproject_f64(g, fl.astype(np.float64))
    assert rel_max(img, ref) < TOL
    assert rel_max(img, f64) < TOL
    assert rel_l2(img, f64) < 1e-6
    mask = oracle.mask(N).astype(bool)
    assert np.all(img[~mask] == 0)
    plan = _exact_plan(_geom(N, pix, V))
    exact = ops.backproject(plan, _t(fl, gpu))[0].cpu().numpy()
    assert rel_max(img, exact) < TOL
    plan.close()


def test_kat_f_end_to_end(gpu, golden, oracle):
    ops = _ops()
    p = golden.kat_f_input()
    plan = ops.plan_for(_geom())
    img = ops.fbp(plan, _t(p, gpu), 2)[0].cpu().numpy()
    ref = oracle.fbp_ref(Geom(), p, 2)  # bit-identical to the shipped KAT-F image
    f64 = oracle.fbp_f64(Geom(), p, 2)
    assert rel_max(img, ref) < TOL
    assert rel_l2(img, ref) < TOL
    assert rel_max(img, f64) < TOL


@pytest.mark.parametrize("case,N,pix,V", [("n64", 64, 0.7, 80), ("n128", 128, 0.35, 160)])
def test_small_fbp_matches_reference(gpu, golden, case, N, pix, V):
    ops = _ops()
    t = golden.npz("ref_small.npz")
    raw = t[case + "_raw"]
    norm = (raw / raw.max()).astype(np.float32)
    plan = ops.plan_for(_geom(N, pix, V))
    fl = ops.filter_sinogram(plan, _t(norm, gpu), 2)[0].cpu().numpy()
    assert rel_max(fl, t[case + "_fltd"]) < TOL
    img = ops.backproject(plan, _t(fl, gpu))[0].cpu().numpy()
    assert rel_max(img, t[case + "_img"]) < TOL


def test_batched_fbp_equals_single(gpu, golden):
    import torch
    ops = _ops()
    p = golden.kat_f_input()
    plan = _exact_plan(_geom())
    stack = torch.stack([_t(p, gpu), _t(p * np.float32(0.9), gpu), _t(np.sqrt(p), gpu)])
    many = ops.fbp(plan, stack, 2)
    for i in range(3):
        one = ops.fbp(plan, stack[i], 2)[0]
        assert torch.equal(many[i], one)


@pytest.mark.parametrize("W,V,S", [(609, 800, 3), (100, 90, 3), (1000, 64, 2), (64, 33, 1)])
@pytest.mark.parametrize("filt", [2, 1])
def test_filter_shapes_match_oracle(gpu, oracle, W, V, S, filt):
    """The split-K MFMA filter over stacked sinograms whose row count is not a
    multiple of 32, for channel counts where the edge-renormalised rows fall
    in different column tiles (Toeplitz tiles read the generator from LDS,
    edge tiles the packed matrix), against the reference-order filter and FP64."""
    from virtualct_amd import default_geometry
    ops = _ops()
    geom = default_geometry(recon_size=128, pixel_cm=0.35, n_views=V, n_channels=W, iso_channel=(W - 1) / 2.0)
    plan = ops.plan_for(geom)
    rng = np.random.default_rng(W + V + S)
    p = rng.uniform(0.02, 1.0, size=(S, V, W)).astype(np.float32)
    fl = ops.filter_sinogram(plan, _t(p, gpu), filt).cpu().numpy()
    g = Geom(N=128, pix=0.35, V=V, W=W, iso_channel=(W - 1) / 2.0)
    for s in range(S):
        assert rel_max(fl[s], oracle.filter_ref(g, p[s], filt)) < TOL, s
        assert rel_max(fl[s], oracle.filter_f64(g, p[s], filt)) < TOL, s
    # the same with a caller-owned workspace
    work = ops.filter_workspace(plan, S)
    fl2 = ops.filter_sinogram(plan, _t(p, gpu), filt, work=work).cpu().numpy()
    assert np.array_equal(fl, fl2)


def test_filter_scratch_growth_keeps_captured_graph_valid(gpu):
    """A captured forward + FBP run owns its filter workspace, and the context
    scratch only grows (a larger eager call retires the old buffer instead of
    freeing it): replays after a larger eager FBP still equal the eager run."""
    import torch
    from virtualct_amd import batch, phantoms
    ops = _ops()
    att = torch.tensor([[[0.2, 0.1]] * 6], device=gpu)
    pdf = torch.tensor([[0.6, 0.4]], device=gpu)
    N, pix = 64, 0.7
    ph = torch.from_numpy(phantoms.body_phantom(41, N=N, pix_cm=pix))[None].to(gpu)
    run = batch.BatchForwardFBP([(_geom(N, pix, 160), ph)], att, pdf, sigma=0.0, blocked=False)
    _, imgs = run.run()
    eager = [i.clone() for i in imgs]
    run.capture()
    plan = ops.plan_for(_geom())
    big = torch.rand((24, 800, 609), device=gpu) * 0.5 + 0.25
    many = ops.filter_sinogram(plan, big, 2)  # 24 sinograms: the context scratch grows
    for _ in range(2):
        _, i2 = run.replay()
        torch.cuda.synchronize()
        for a, b in zip(i2, eager):
            assert torch.equal(a, b)
    assert torch.equal(many[5], ops.filter_sinogram(plan, big[5], 2)[0])


@pytest.mark.parametrize("N,pix,V,S", [(512, 0.1, 800, 16), (512, 0.1, 800, 48),
                                        (512, 0.1, 800, 5), (512, 0.1, 200, 2), (512, 0.1, 200, 3),
                                        (512, 0.1, 200, 10), (512, 0.1, 200, 17), (512, 0.1, 200, 35),
                                        (512, 0.15, 200, 6),
                                        (128, 0.6, 160, 6), (64, 0.7, 90, 3)])
def test_multi_sinogram_backprojection_equals_single(gpu, oracle, N, pix, V, S):
    """S sinograms of one geometry share each lane's pixel-view geometry (up to
    16 per lane, in even slices: 16 -> one slice of 16, 48 -> 3 x 16, 17 -> 9 + 8,
    35 -> 12 + 12 + 11); every
    image equals the single-sinogram backprojection
    bit for bit.  512 x 0.15 cm and 128 x 0.6 cm are wide fields (library atan
    fallback); at 128^2 and 64^2 the launch keeps 2^18 lanes by running one
    sinogram per lane."""
    import torch
    ops = _ops()
    rng = np.random.default_rng(40 + S)
    fl = rng.normal(size=(S, V, 609)).astype(np.float32) * 1e3
    plan = _exact_plan(_geom(N, pix, V))
    many = ops.backproject(plan, _t(fl, gpu))
    for i in range(S):
        one = ops.backproject(plan, _t(fl[i], gpu))[0]
        assert torch.equal(many[i], one), i
    if N <= 128:
        ref = oracle.backproject_ref(Geom(N=N, pix=pix, V=V), fl[S - 1])
        assert rel_max(many[S - 1].cpu().numpy(), ref) < 1e-5
    plan.close()


def test_mbfpa_then_fbp_recovers_mu_map(gpu, golden):
    """The pipeline MBFPA -> FBP gives a mu map in 1/cm in the same
    orientation as the phantom (correlation with sum rho_m mu_m(60 keV))."""
    ops = _ops()
    from virtualct_amd import spectra
    att, pdf, e = spectra.attenuation(golden.path("spectrum_120kV.txt"), 6)
    plan = ops.plan_for(_geom())
    ph = golden.icru110()
    sino = ops.mbfpa_forward(plan, _t(ph, gpu), _t(att, gpu), _t(pdf, gpu))
    img = ops.fbp(plan, sino, 2)[0].cpu().numpy()
    i60 = int(np.argmin(np.abs(e - 0.06)))
    mu = (ph.astype(np.float64) / 1000 * att[:, i60, None, None]).sum(0)
    r = np.corrcoef(img.ravel(), mu.ravel())[0, 1]
    assert r > 0.95


@pytest.mark.parametrize("alpha,beta", [(0.01, 2.0), (0.01, 3.7)])
def test_beam_hardening_kernel_and_fused_filter(gpu, golden, oracle, alpha, beta):
    ops = _ops()
    p = golden.sigma0_120kv()
    ref, _, _ = oracle.beam_hardening(p, alpha, beta)  # bit-identical to the reference tool
    got = ops.beam_hardening(_t(p, gpu), alpha, beta).cpu().numpy()
    assert rel_max(got, ref) < 1e-6
    assert np.count_nonzero(got != ref) < 0.01 * got.size
    plan = ops.plan_for(_geom())
    fused = ops.filter_sinogram_bhc(plan, _t(p, gpu), alpha, beta)[0].cpu().numpy()
    two_step = oracle.filter_ref(Geom(), ref, 2)
    assert rel_max(fused, two_step) < 1e-5
    mx, mn = ops.minmax(_t(p, gpu))
    assert mx == p.max() and mn == p.min()


def test_bhc_driver(gpu, golden, oracle, tmp_path):
    import subprocess
    from conftest import ROOT
    import os
    p = golden.sigma0_120kv()
    p.tofile(tmp_path / "reprojection_float.raw")
    r = subprocess.run([os.path.join(ROOT, "virtualct_amd", "bin", "mvct_bhc"), "0.01", "2.0"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "reprojection_float_cor.raw", np.float32).reshape(800, 609)
    ref, mn0, mn1 = oracle.beam_hardening(p, 0.01, 2.0)
    assert rel_max(out, ref) < 1e-6
    assert r.stdout.strip() == f"{mn0:f} {mn1:f}"


def test_wide_field_backprojection_takes_library_atan(gpu, oracle):
    """A 76.8 cm field (N=128, 0.6 cm) puts masked pixels at |tan(gamma)| up to
    0.53, beyond the polynomial: the kernel variant with the library atan
    fallback, against the oracle."""
    ops = _ops()
    N, pix, V = 128, 0.6, 160
    g = Geom(N=N, pix=pix, V=V)
    r = (N - 1) / 2.0 * pix
    assert r / np.sqrt(85.0 ** 2 - r ** 2) > 0.5
    rng = np.random.default_rng(31)
    fl = rng.normal(size=(V, 609)).astype(np.float32) * 1e3
    plan = ops.plan_for(_geom(N, pix, V))
    img = ops.backproject(plan, _t(fl, gpu))[0].cpu().numpy()
    ref = oracle.backproject_ref(g, fl)
    f64 = oracle.backproject_f64(g, fl.astype(np.float64))
    assert rel_max(img, ref) < 1e-5
    assert rel_max(img, f64) < 1e-5
    mask = oracle.mask(N).astype(bool)
    assert np.all(img[~mask] == 0)


def test_backprojection_more_views_than_one_trig_block(gpu, oracle):
    """2100 views: a view loop longer than 2048 views (exact-order kernel)."""
    ops = _ops()
    N, pix, V = 64, 0.7, 2100
    g = Geom(N=N, pix=pix, V=V)
    fl = np.random.default_rng(32).normal(size=(V, 609)).astype(np.float32)
    plan = _exact_plan(_geom(N, pix, V))
    img = ops.backproject(plan, _t(fl, gpu))[0].cpu().numpy()
    ref = oracle.backproject_ref(g, fl)
    assert rel_max(img, ref) < 1e-5
    assert np.count_nonzero(img != ref) < 0.05 * img.size


@pytest.mark.parametrize("N,pix,V,S", [(512, 0.1, 800, 16), (512, 0.1, 800, 5), (128, 0.6, 160, 6),
                                        (64, 0.7, 90, 3), (96, 0.5, 801, 2)])
def test_multi_sinogram_float_weights_match_exact_terms(gpu, oracle, N, pix, V, S):
    """By default several sinograms per call apply the interpolation and
    distance weights in float: within 1e-6 of the exact FP64 terms (same view
    order) and 1e-5 of the reference, and each image does not depend on how
    the call is split (one call of S, or S/2 + S/2)."""
    import torch
    ops = _ops()
    rng = np.random.default_rng(70 + S)
    fl = _t(rng.normal(size=(S, V, 609)).astype(np.float32) * 1e3, gpu)
    plan = ops.Plan(_geom(N, pix, V))
    fast = ops.backproject(plan, fl)
    exact_plan = _exact_plan(_geom(N, pix, V))
    exact = ops.backproject(exact_plan, fl)
    assert rel_max(fast.cpu().numpy(), exact.cpu().numpy()) < 1e-6
    ref = oracle.backproject_ref(Geom(N=N, pix=pix, V=V), fl[S - 1].cpu().numpy())
    assert rel_max(fast[S - 1].cpu().numpy(), ref) < 1e-5
    h = S // 2
    if h >= 2:
        halves = torch.cat([ops.backproject(plan, fl[:h].contiguous()), ops.backproject(plan, fl[h:].contiguous())])
        assert torch.equal(halves, fast)
    plan.close()
    exact_plan.close()
