"""The workloads of configs 3-5 at their own sizes, checked against the oracle.

* config 3: the 8-element calibration phantom, 120 kV and 6 MV from one trace,
  each sinogram followed by FBP at 512^2;
* config 4: the batched forward + FBP of phantoms of uneven size (256^2,
  384^2, 512^2, 768^2 at a 50 cm field, so 1/pixel is not an integer) x 80 kV,
  120 kV and 6 MV, run as one captured HIP graph, every sinogram and image
  against the FP64 forward and the reference-order FBP;
* config 5: the batched dual-energy decomposition at 16 x 512^2, its first
  gradient against the FP64 adjoint and against central differences of the
  FP64 forward model.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2, rel_max
from oracle.oracle import Geom

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _lnrel(a, b, floor=1e-6):
    """max relative error in the -ln domain over rays with b > floor"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = b > floor
    la, lb = -np.log(np.maximum(a[m], 1e-300)), -np.log(b[m])
    return float(np.abs(la - lb).max() / max(np.abs(lb).max(), 1e-12))


def _tables(names, M, tmp):
    """Per-spectrum (att [M,NE_s], pdf [NE_s]) and the zero-padded stack [S,M,NE], [S,NE]."""
    from virtualct_amd import spectra
    tabs = [spectra.attenuation(spectra.synthetic_spectrum_file(n, str(tmp)), M)[:2] for n in names]
    NE = max(a.shape[1] for a, _ in tabs)
    att = np.zeros((len(tabs), M, NE), np.float32)
    pdf = np.zeros((len(tabs), NE), np.float32)
    for s, (a, p) in enumerate(tabs):
        att[s, :, : a.shape[1]] = a
        pdf[s, : p.shape[0]] = p
    return tabs, att, pdf


def _geom(N=512, pix=0.1, V=800):
    from virtualct_amd import default_geometry
    return default_geometry(recon_size=N, pixel_cm=pix, n_views=V)


def test_config3_calibration_phantom_dual_energy_with_fbp(gpu, oracle, tmp_path):
    """8 elements (H C N O P Ca Mg Si), 120 kV + 6 MV in one call; each
    normalised sinogram against the FP64 forward, each FBP against the
    reference-order FBP of the same sinogram and the FP64 FBP."""
    from virtualct_amd import ops, phantoms
    ph = phantoms.calibration_phantom(2000)
    assert ph.shape == (8, 512, 512)
    tabs, att, pdf = _tables(("120kV", "6MV"), 8, tmp_path)
    plan = ops.plan_for(_geom())
    sino = ops.mbfpa_forward(plan, _t(ph, gpu), _t(att, gpu), _t(pdf, gpu))
    img = ops.fbp(plan, sino, 2).cpu().numpy()
    sino = sino.cpu().numpy()
    dens = phantoms.densities_f32(ph)
    g = Geom()
    for s, (a, p) in enumerate(tabs):
        f64 = oracle.forward_f64(g, dens, a, p)
        ref = f64 / f64.max()
        assert rel_max(sino[s], ref) < TOL, s
        assert _lnrel(sino[s], ref) < TOL, s
        assert rel_max(img[s], oracle.fbp_ref(g, sino[s], 2)) < TOL, s
        assert rel_max(img[s], oracle.fbp_f64(g, sino[s], 2)) < TOL, s


def test_config4_batched_forward_fbp_against_oracle(gpu, oracle, tmp_path):
    """One phantom per size at 50 cm / N through BatchForwardFBP (blocked
    system matrices, one persistent gather, stacked MFMA filter, multi-sinogram
    backprojection) replayed from its HIP graph: every sinogram within 1e-5 of
    the FP64 forward (max-rel and -ln), every image within 1e-5 of the
    reference-order FBP of that sinogram, and of the plain per-phantom calls."""
    import torch
    from virtualct_amd import batch, ops, phantoms
    names = ("80kV", "120kV", "6MV")
    tabs, att, pdf = _tables(names, 6, tmp_path)
    sizes = (256, 384, 512, 768)
    groups, dens = [], []
    for k, N in enumerate(sizes):
        pix = 50.0 / N
        ph = phantoms.body_phantom(6000 + 100 * k, N=N, pix_cm=pix)
        dens.append(phantoms.densities_f32(ph))
        groups.append((_geom(N, pix), _t(ph[None], gpu)))
    run = batch.BatchForwardFBP(groups, _t(att, gpu), _t(pdf, gpu), sigma=0.0)
    run.capture()
    sinos, imgs = run.replay()
    torch.cuda.synchronize()
    ops.check_errors()
    for k, N in enumerate(sizes):
        pix = 50.0 / N
        g = Geom(N=N, pix=pix)
        plan = ops.plan_for(_geom(N, pix))
        s_gpu = sinos[k].cpu().numpy()   # [S,V,W] of the group's one phantom
        i_gpu = imgs[k].cpu().numpy()
        plain_s = ops.mbfpa_forward(plan, groups[k][1][0], _t(att, gpu), _t(pdf, gpu))
        plain_i = ops.fbp(plan, plain_s, 2).cpu().numpy()
        for s, (a, p) in enumerate(tabs):
            f64 = oracle.forward_f64(g, dens[k], a, p)
            ref = f64 / f64.max()
            assert rel_max(s_gpu[s], ref) < TOL, (N, names[s])
            assert _lnrel(s_gpu[s], ref) < TOL, (N, names[s])
            assert rel_max(i_gpu[s], oracle.fbp_ref(g, s_gpu[s], 2)) < TOL, (N, names[s])
            assert rel_max(i_gpu[s], plain_i[s]) < TOL, (N, names[s])
            assert rel_l2(s_gpu[s], plain_s[s].cpu().numpy()) < 1e-6, (N, names[s])


def test_config4_batched_gathers_against_fp64_path_integrals(gpu, oracle):
    """Two phantoms at 384^2 x 50/384 cm through the blocked batched gather
    (tile-resident stage 1): path integrals within 2e-6 of the FP64 oracle's
    (lengths are cached as float)."""
    import torch
    from virtualct_amd import batch, phantoms
    from virtualct_amd import ops
    N, pix = 384, 50.0 / 384
    plan = ops.plan_for(_geom(N, pix))
    sm = batch.SystemMatrix(plan, transpose=False, blocked=True)
    d = np.stack([phantoms.densities_f32(phantoms.body_phantom(7100 + b, N=N, pix_cm=pix)) for b in range(2)])
    x = batch.interleave(_t(d, gpu))
    L = torch.empty((plan.V * plan.W, 2, 8), device=gpu)
    batch.batch_paths([(sm, x, L, 2)], 6)
    Ln = batch.from_planes(L).cpu().numpy()
    g = Geom(N=N, pix=pix)
    one = np.ones(1, np.float32)
    for b in range(2):
        _, Lo = oracle.forward_f64(g, d[b], np.zeros((6, 1), np.float32), one, want_L=True)
        assert rel_max(Ln[:, b, :6].reshape(plan.V, plan.W, 6), Lo) < 2e-6, b
    sm.close()


@pytest.fixture(scope="module")
def config5(gpu, tmp_path_factory):
    """16 phantoms at 512^2, noisy 120 kV + 6 MV targets, the water start of
    the decomposition and the blocked solver with its first gradient."""
    import torch
    from virtualct_amd import batch, decompose, ops, phantoms
    tmp = tmp_path_factory.mktemp("c5")
    tabs, att, pdf = _tables(("120kV", "6MV"), 6, tmp)
    from virtualct_amd import spectra
    e = spectra.attenuation(spectra.synthetic_spectrum_file("120kV", str(tmp)), 6)[2]
    plan = ops.plan_for(_geom())
    B = 16
    d_att, d_pdf = _t(att, gpu), _t(pdf, gpu)
    truth = [phantoms.body_phantom(5000 + b) for b in range(B)]
    targets = torch.stack([ops.mbfpa_forward(plan, _t(truth[b], gpu), d_att, d_pdf, sigma=5.37e-5, seed=0x5EED,
                                             stream_id=2 * b) for b in range(B)])
    mu = ops.fbp(plan, targets[:, 0].contiguous())
    e60 = int(np.argmin(np.abs(e - 0.06)))
    mu_w = float(0.1119 * tabs[0][0][0, e60] + 0.8881 * tabs[0][0][3, e60])
    init = decompose.water_init(mu, 6, mu_w)
    sm = batch.SystemMatrix(plan, transpose=True, blocked=True)
    solver = batch.BatchedSolver(sm, targets, d_att, d_pdf, init, lanes=16, blocks=1024)
    solver._forward_loss_grad()
    torch.cuda.synchronize()
    grad = batch.deinterleave(solver.g, 6, 512).cpu().numpy()
    loss = solver.loss.cpu().numpy()
    out = dict(tabs=tabs, plan=plan, solver=solver, grad=grad, loss=loss, init=init.cpu().numpy(),
               targets=targets.cpu().numpy().astype(np.float64))
    yield out
    sm.close()


def _loss_terms(oracle, g, x, tabs, target):
    """Per-spectrum FP64 intensities I_s and residuals r_s = ln(I_s / Imax_s) - ln t_s
    of the decomposition loss sum_s sum_rays r_s^2 (Imax_s = sum_E pdf_s)."""
    out = []
    for s, (a, p) in enumerate(tabs):
        I = oracle.forward_f64(g, x, a, p)
        imax = float(np.sum(p.astype(np.float64)))
        out.append((I, np.log(I / imax) - np.log(target[s])))
    return out


@pytest.mark.parametrize("b", [0, 9])
def test_config5_gradient_against_fp64_adjoint(gpu, oracle, config5, b):
    """The batched solver's gradient of phantom b (sum over both spectra of the
    log-domain loss) within 1e-4 of the FP64 adjoint driven by the FP64
    forward's d loss / d I = 2 r / I."""
    c = config5
    g = Geom()
    x = np.ascontiguousarray(c["init"][b], np.float32)
    ref = np.zeros_like(c["grad"][b], np.float64)
    loss = 0.0
    for s, (I, r) in enumerate(_loss_terms(oracle, g, x, c["tabs"], c["targets"][b])):
        ref += oracle.adjoint_f64(g, x, c["tabs"][s][0], c["tabs"][s][1], 2.0 * r / I)
        loss += float(np.sum(r * r))
    got = c["grad"][b]
    assert np.isfinite(got).all()
    assert rel_l2(got, ref) < 1e-4
    assert rel_max(got, ref) < 1e-4
    assert c["loss"][b] == pytest.approx(loss, rel=1e-4)


def test_config5_gradient_central_differences(gpu, oracle, config5):
    """h = 1e-3 g/cm^3 central differences of the FP64 loss at voxels of
    phantom 0 whose H or O density (the water start) exceeds h."""
    c = config5
    g = Geom()
    x0 = np.ascontiguousarray(c["init"][0], np.float32)
    got = c["grad"][0]
    h = 1e-3
    checked = 0
    for m, y, xx in [(3, 256, 256), (0, 256, 256), (3, 200, 330), (0, 300, 180), (3, 180, 240)]:
        if x0[m, y, xx] <= 2 * h:
            continue
        xp = x0.copy(); xp[m, y, xx] += np.float32(h)
        xm = x0.copy(); xm[m, y, xx] -= np.float32(h)
        dl = 0.0
        for (_, rp), (_, rm) in zip(_loss_terms(oracle, g, xp, c["tabs"], c["targets"][0]),
                                    _loss_terms(oracle, g, xm, c["tabs"], c["targets"][0])):
            dl += float(np.sum((rp - rm) * (rp + rm)))  # sum rp^2 - rm^2 without cancellation
        fd = dl / (float(xp[m, y, xx]) - float(xm[m, y, xx]))
        assert got[m, y, xx] == pytest.approx(fd, rel=2e-3, abs=1e-6 * np.abs(got).max()), (m, y, xx)
        checked += 1
    assert checked >= 3


def test_config5_iterations_lower_the_loss(gpu, config5):
    """Ten projected Barzilai-Borwein iterations of the 16-phantom solver: the
    total loss falls below 0.6 of the start, every phantom's loss falls below
    its start at some iteration (the method is not monotone), and the
    densities stay non-negative and finite."""
    c = config5
    s = c["solver"]
    hist = [s.step().clone() for _ in range(11)]
    l0 = hist[0]
    best = torch_min(hist[1:])
    assert hist[-1].sum() < 0.6 * l0.sum(), (l0, hist[-1])
    assert (best < l0).all(), (l0, best)
    assert (s.x >= 0).all() and s.x.isfinite().all()


def torch_min(ts):
    import torch
    return torch.stack(ts).amin(0)
