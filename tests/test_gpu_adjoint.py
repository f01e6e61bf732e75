"""GPU adjoint of the MBFPA forward model (ray-driven transposed projection
with hardware float atomics) behind torch.autograd: dot-product identity,
FP64 CPU adjoint at 1e-4 relative, and central finite differences
(h = 1e-3 g/cm^3) of the FP64 forward model."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2, rel_max
from oracle.oracle import Geom

pytestmark = pytest.mark.gpu


def _ops():
    from virtualct_amd import ops
    return ops


def _geom(N=512, pix=0.1, V=800):
    from virtualct_amd import default_geometry
    return default_geometry(recon_size=N, pixel_cm=pix, n_views=V)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_dot_product_identity(gpu):
    """<A x, y> = <x, A^T y> for the per-element path-integral operator."""
    import torch
    ops = _ops()
    plan = ops.plan_for(_geom())
    g = torch.Generator(device=gpu).manual_seed(3)
    M = 6
    x = torch.rand((M, 512, 512), generator=g, device=gpu)
    pk, dt = ops.pack_densities(plan, x)
    Ax = ops.trace_paths(plan, pk, dt, M)  # [V,W,8]
    y = torch.zeros((800, 609, 8), device=gpu)
    y[..., :M] = torch.rand((800, 609, M), generator=g, device=gpu)
    ATy = ops.trace_adjoint(plan, y, M)  # [N,N,8]
    lhs = (Ax.double() * y.double()).sum().item()
    rhs = (x.permute(1, 2, 0).double() * ATy[..., :M].double()).sum().item()
    assert lhs == pytest.approx(rhs, rel=1e-5)


def _problem(golden, gpu, N=128, pix=0.35, V=160, seed=7):
    from virtualct_amd import phantoms, spectra
    import tempfile, os
    ph = phantoms.body_phantom(seed, N=N, pix_cm=pix)
    with tempfile.TemporaryDirectory() as d:
        tabs = [spectra.attenuation(spectra.synthetic_spectrum_file(n, d), 6) for n in ("120kV", "6MV")]
    return ph, tabs


def test_gradient_matches_fp64_adjoint(gpu, golden, oracle):
    import torch
    ops = _ops()
    N, pix, V = 128, 0.35, 160
    ph, tabs = _problem(golden, gpu, N, pix, V)
    dens = ph.astype(np.float32) / np.float32(1000)
    plan = ops.plan_for(_geom(N, pix, V))
    rng = np.random.default_rng(11)
    for att, pdf, _ in tabs:
        gI = rng.normal(size=(V, 609))
        x = _t(dens, gpu).requires_grad_(True)
        I = ops.mbfpa_intensity(x, plan, _t(att, gpu), _t(pdf, gpu))
        (I[0] * _t(gI.astype(np.float32), gpu)).sum().backward()
        got = x.grad.cpu().numpy()
        ref = oracle.adjoint_f64(Geom(N=N, pix=pix, V=V), dens, att, pdf, gI)
        assert rel_l2(got, ref) < 1e-4
        assert rel_max(got, ref) < 1e-4


def test_dual_spectrum_gradient_is_sum(gpu, golden, oracle):
    """One adjoint trace for both spectra equals the sum of the two FP64 adjoints."""
    import torch
    ops = _ops()
    N, pix, V = 128, 0.35, 160
    ph, tabs = _problem(golden, gpu, N, pix, V)
    dens = ph.astype(np.float32) / np.float32(1000)
    NE = max(a.shape[1] for a, _, _ in tabs)
    att = np.zeros((2, 6, NE), np.float32)
    pdf = np.zeros((2, NE), np.float32)
    for s, (a, p, _) in enumerate(tabs):
        att[s, :, : a.shape[1]] = a
        pdf[s, : p.shape[0]] = p
    plan = ops.plan_for(_geom(N, pix, V))
    rng = np.random.default_rng(12)
    gI = rng.normal(size=(2, V, 609))
    x = _t(dens, gpu).requires_grad_(True)
    I = ops.mbfpa_intensity(x, plan, _t(att, gpu), _t(pdf, gpu))
    (I * _t(gI.astype(np.float32), gpu)).sum().backward()
    got = x.grad.cpu().numpy()
    ref = sum(oracle.adjoint_f64(Geom(N=N, pix=pix, V=V), dens, a, p, gI[s]) for s, (a, p, _) in enumerate(tabs))
    assert rel_l2(got, ref) < 1e-4


def test_gradient_central_differences(gpu, golden, oracle):
    """h = 1e-3 g/cm^3 central differences of the FP64 forward model."""
    import torch
    ops = _ops()
    N, pix, V = 64, 0.7, 80
    ph, tabs = _problem(golden, gpu, N, pix, V)
    dens = ph.astype(np.float32) / np.float32(1000)
    att, pdf, _ = tabs[0]
    plan = ops.plan_for(_geom(N, pix, V))
    rng = np.random.default_rng(13)
    gI = rng.normal(size=(V, 609))
    x = _t(dens, gpu).requires_grad_(True)
    I = ops.mbfpa_intensity(x, plan, _t(att, gpu), _t(pdf, gpu))
    (I[0] * _t(gI.astype(np.float32), gpu)).sum().backward()
    got = x.grad.cpu().numpy()
    g = Geom(N=N, pix=pix, V=V)
    h = 1e-3
    checked = 0
    for (m, y, xx) in [(3, 32, 32), (0, 20, 40), (5, 30, 25), (1, 40, 15), (3, 10, 32)]:
        if dens[m, y, xx] < h:
            continue
        dp = dens.copy(); dp[m, y, xx] += h
        dm = dens.copy(); dm[m, y, xx] -= h
        fd = ((gI * oracle.forward_f64(g, dp, att, pdf)).sum() - (gI * oracle.forward_f64(g, dm, att, pdf)).sum())
        fd /= float(np.float64(dp[m, y, xx]) - np.float64(dm[m, y, xx]))
        assert got[m, y, xx] == pytest.approx(fd, rel=2e-3, abs=1e-6 * np.abs(got).max())
        checked += 1
    assert checked >= 3


def test_autograd_gradcheck_small(gpu):
    """torch.autograd.gradcheck-style probe in float32 on a tiny geometry:
    directional derivative vs the adjoint."""
    import torch
    ops = _ops()
    from virtualct_amd import phantoms
    N, pix, V = 64, 0.7, 80
    plan = ops.plan_for(_geom(N, pix, V))
    dens = torch.from_numpy(phantoms.densities_f32(phantoms.body_phantom(3, N=N, pix_cm=pix))).to(gpu)
    att = torch.tensor([[0.2, 0.1], [0.18, 0.1], [0.18, 0.1], [0.18, 0.1], [0.5, 0.2], [1.5, 0.5]], device=gpu)
    pdf = torch.tensor([0.6, 0.4], device=gpu)
    x = dens.clone().requires_grad_(True)
    I = ops.mbfpa_intensity(x, plan, att, pdf)
    w = torch.randn_like(I)
    (I * w).sum().backward()
    # a direction proportional to rho keeps every density positive (|eps d| <=
    # 4% of rho); the model clamps negative densities, which is not differentiable
    d = torch.randn_like(dens).clamp(-4, 4) * dens
    eps = 1e-2
    with torch.no_grad():
        fp = (ops.mbfpa_intensity(dens + eps * d, plan, att, pdf).double() * w.double()).sum()
        fm = (ops.mbfpa_intensity(dens - eps * d, plan, att, pdf).double() * w.double()).sum()
    fd = ((fp - fm) / (2 * eps)).item()
    ad = (x.grad.double() * d.double()).sum().item()
    assert ad == pytest.approx(fd, rel=2e-3)


@pytest.mark.parametrize("mode", ["atomic", "sysmat"])
def test_both_adjoint_modes_match_fp64(gpu, golden, oracle, mode):
    """The atomic re-tracing adjoint and the gather over the cached system
    matrix both match the FP64 adjoint."""
    ops = _ops()
    N, pix, V = 128, 0.35, 160
    ph, tabs = _problem(golden, gpu, N, pix, V)
    dens = ph.astype(np.float32) / np.float32(1000)
    att, pdf, _ = tabs[0]
    plan = ops.Plan(_geom(N, pix, V))  # own plan: the mode is per plan
    plan.adjoint_mode = mode
    rng = np.random.default_rng(21)
    gI = rng.normal(size=(V, 609))
    x = _t(dens, gpu).requires_grad_(True)
    I = ops.mbfpa_intensity(x, plan, _t(att, gpu), _t(pdf, gpu))
    (I[0] * _t(gI.astype(np.float32), gpu)).sum().backward()
    ref = oracle.adjoint_f64(Geom(N=N, pix=pix, V=V), dens, att, pdf, gI)
    got = x.grad.cpu().numpy()
    assert rel_l2(got, ref) < 1e-4
    assert rel_max(got, ref) < 1e-4


def test_sysmat_adjoint_mode_matches_fp64_and_is_deterministic(gpu, golden, oracle):
    """adjoint_mode "sysmat": the gradient gathers over the cached voxel-major
    system matrix; it matches the FP64 adjoint, keeps the sign test of the
    tracer (negative densities get zero) and repeats bit for bit."""
    import torch
    ops = _ops()
    N, pix, V = 128, 0.35, 160
    ph, tabs = _problem(golden, gpu, N, pix, V)
    dens = ph.astype(np.float32) / np.float32(1000)
    dens[2, 60:64, 60:64] = -0.5  # clamped voxels: zero gradient
    att, pdf, _ = tabs[1]
    plan = ops.Plan(_geom(N, pix, V))
    plan.adjoint_mode = "sysmat"
    rng = np.random.default_rng(22)
    gI = rng.normal(size=(V, 609))
    grads = []
    for _ in range(2):
        x = _t(dens, gpu).requires_grad_(True)
        I = ops.mbfpa_intensity(x, plan, _t(att, gpu), _t(pdf, gpu))
        (I[0] * _t(gI.astype(np.float32), gpu)).sum().backward()
        grads.append(x.grad.cpu().numpy())
    assert np.array_equal(grads[0], grads[1])
    ref = oracle.adjoint_f64(Geom(N=N, pix=pix, V=V), dens, att, pdf, gI)
    assert rel_l2(grads[0], ref) < 1e-4
    assert rel_max(grads[0], ref) < 1e-4
    assert np.all(grads[0][2, 60:64, 60:64] == 0)


@pytest.mark.parametrize("N,pix,V,exact", [
    (96, 50.0 / 96, 801, False),   # odd view count (the last view is not traced), N not a multiple of 32
    (100, 0.3, 50, False),         # few views, N not a multiple of 32
    (64, 0.78125, 200, False),     # image corners beyond N/2 cm: the r >= rmax test, exact march
    (128, 0.35, 160, True),        # the exact-division march forced on every ray
])
def test_tile_adjoint_geometries_match_fp64(gpu, golden, oracle, N, pix, V, exact):
    """The atomic adjoint (tile-resident: each workgroup marches only the
    steps of the rays that can land in its 32x32 tile) against the FP64
    adjoint on geometries where the tiling meets its edge cases."""
    from virtualct_amd import _lib
    ops = _ops()
    ph, tabs = _problem(golden, gpu, N, pix, V)
    dens = ph.astype(np.float32) / np.float32(1000)
    att, pdf, _ = tabs[1]
    plan = ops.Plan(_geom(N, pix, V))
    plan.adjoint_mode = "atomic"
    if exact:
        plan.set_option(_lib.OPT_EXACT_TRACE, 1)
    rng = np.random.default_rng(31 + N)
    gI = rng.normal(size=(V, 609))
    print(f"tile adjoint N={N} pix={pix:.4f} V={V} exact={exact}", flush=True)
    x = _t(dens, gpu).requires_grad_(True)
    I = ops.mbfpa_intensity(x, plan, _t(att, gpu), _t(pdf, gpu))
    (I[0] * _t(gI.astype(np.float32), gpu)).sum().backward()
    got = x.grad.cpu().numpy()
    ref = oracle.adjoint_f64(Geom(N=N, pix=pix, V=V), dens, att, pdf, gI)
    assert rel_l2(got, ref) < 1e-4
    assert rel_max(got, ref) < 1e-4


def test_tile_adjoint_equals_sysmat_gather_at_512(gpu):
    """At 512^2 x 800 x 609 the atomic (tile) adjoint and the gather over the
    system matrix (pinned to the FP64 adjoint at 128^2) sum the same segments
    in different orders: equal to float rounding, with the sign test."""
    import torch
    ops = _ops()
    plan = ops.plan_for(_geom())
    g = torch.Generator(device=gpu).manual_seed(5)
    M = 6
    w = torch.zeros((800, 609, 8), device=gpu)
    w[..., :M] = torch.rand((800, 609, M), generator=g, device=gpu) - 0.25
    dens = torch.rand((512, 512, 8), generator=g, device=gpu)
    dens[100:140, 200:260, 3] = -1.0
    tile = ops.trace_adjoint(plan, w, M, dens_packed_f32=dens)
    gather = ops.sysmat_adjoint(plan, w, M, dens_packed_f32=dens)
    torch.cuda.synchronize()
    a = tile[..., :M].cpu().numpy().astype(np.float64)
    b = gather[..., :M].cpu().numpy().astype(np.float64)
    assert rel_l2(a, b) < 1e-6
    assert np.abs(a - b).max() < 1e-5 * np.abs(b).max()
    assert np.all(a[100:140, 200:260, 3] == 0)
    assert np.all(tile[..., M:].cpu().numpy() == 0)


def test_tile_adjoint_nonfinite_weight_reaches_its_voxels(gpu):
    """A NaN weight makes the gradient NaN exactly on the voxels its ray
    crosses, as float sums do, and leaves every other voxel as it was (the
    workgroups that see it sum in float, the others in fixed point)."""
    import torch
    ops = _ops()
    plan = ops.plan_for(_geom(128, 0.35, 160))
    M = 3
    g = torch.Generator(device=gpu).manual_seed(8)
    w = torch.zeros((160, 609, 8), device=gpu)
    w[..., :M] = torch.rand((160, 609, M), generator=g, device=gpu)
    view, ch = 37, 300
    one = torch.zeros_like(w)
    one[view, ch, 0] = 1.0
    crossed = ops.trace_adjoint(plan, one, M)[..., 0].cpu().numpy() > 0
    assert crossed.sum() > 100
    clean = ops.trace_adjoint(plan, w, M).cpu().numpy()
    w[view, ch, 0] = float("nan")
    dirty = ops.trace_adjoint(plan, w, M).cpu().numpy()
    assert np.array_equal(np.isnan(dirty[..., 0]), crossed)
    assert not np.isnan(dirty[..., 1:]).any()
    keep = ~crossed
    assert np.abs(dirty[..., 0][keep] - clean[..., 0][keep]).max() <= 1e-5 * np.abs(clean[..., 0]).max()
    assert np.abs(dirty[..., 1:] - clean[..., 1:]).max() <= 1e-5 * np.abs(clean[..., 1:]).max()


def test_tile_adjoint_scales_with_tiny_and_huge_weights(gpu):
    """The fixed-point tile sums take their scale from the workgroup's weights,
    and weights too small for that scale sum in float: the gradient of
    weights times 1e-30 or 1e30 is the gradient times 1e-30 or 1e30."""
    import torch
    ops = _ops()
    plan = ops.plan_for(_geom(128, 0.35, 160))
    g = torch.Generator(device=gpu).manual_seed(13)
    w = torch.zeros((160, 609, 8), device=gpu)
    w[..., :4] = torch.rand((160, 609, 4), generator=g, device=gpu) + 0.1
    base = ops.trace_adjoint(plan, w, 4).cpu().numpy().astype(np.float64)
    for f in (1e-30, 1e-20, 1e20, 1e30):
        got = ops.trace_adjoint(plan, w * f, 4).cpu().numpy().astype(np.float64) / f
        assert rel_max(got, base) < 1e-6, f
