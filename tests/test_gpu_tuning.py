"""Launch shapes that the autotuner picks (plan options): results do not
depend on them (bit for bit for the tracer workgroup and the backprojector's
sinograms per lane; to float rounding for the adjoint's views per
workgroup), bad values are rejected, and tune_plan / BatchForwardFBP set
values from their candidate lists."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_max

pytestmark = pytest.mark.gpu


def _geom(N, pix, V):
    from virtualct_amd import default_geometry
    return default_geometry(recon_size=N, pixel_cm=pix, n_views=V)


@pytest.mark.parametrize("N,pix,V", [(512, 0.1, 800), (96, 0.5, 162), (128, 0.35, 161)])
def test_trace_block_does_not_change_results(gpu, N, pix, V):
    import torch
    from virtualct_amd import _lib, ops, phantoms
    plan = ops.Plan(_geom(N, pix, V))
    ph = torch.from_numpy(phantoms.body_phantom(23, N=N, pix_cm=pix)).to(gpu)
    pk, dt = ops.pack_densities(plan, ph)
    out = []
    for b in (256, 64, 128):
        plan.set_option(_lib.OPT_TRACE_BLOCK, b)
        L, cnt = ops.trace_paths(plan, pk, dt, 6, counts=True)
        out.append((L.cpu().numpy(), cnt.cpu().numpy()))
    for L, cnt in out[1:]:
        assert np.array_equal(L, out[0][0]) and np.array_equal(cnt, out[0][1])
    plan.close()


def test_bp_sinograms_per_lane_do_not_change_images(gpu):
    import torch
    from virtualct_amd import _lib, ops
    plan = ops.Plan(_geom(256, 0.2, 800))
    g = torch.Generator(device=gpu).manual_seed(4)
    fl = torch.randn((12, 800, 609), generator=g, device=gpu)
    imgs = []
    for k in (0, 2, 5, 16):
        plan.set_option(_lib.OPT_BP_SINOS_PER_LANE, k)
        imgs.append(ops.backproject(plan, fl).cpu().numpy())
    for im in imgs[1:]:
        assert np.array_equal(im, imgs[0])
    plan.close()


def test_adjoint_views_change_only_rounding(gpu):
    import torch
    from virtualct_amd import _lib, ops
    plan = ops.Plan(_geom(128, 0.35, 160))
    g = torch.Generator(device=gpu).manual_seed(6)
    w = torch.zeros((160, 609, 8), device=gpu)
    w[..., :6] = torch.randn((160, 609, 6), generator=g, device=gpu)
    grads = []
    for v in (0, 1, 7, 25, 160, 256):
        plan.set_option(_lib.OPT_ADJOINT_VIEWS, v)
        grads.append(ops.trace_adjoint(plan, w, 6).cpu().numpy())
    for gr in grads[1:]:
        assert rel_max(gr, grads[0]) < 1e-6
    plan.close()


def test_launch_shape_options_reject_bad_values(gpu):
    from virtualct_amd import VctError, _lib, ops
    plan = ops.Plan(_geom(64, 0.7, 80))
    for opt, bad in ((_lib.OPT_TRACE_BLOCK, 96), (_lib.OPT_TRACE_BLOCK, 512), (_lib.OPT_BP_SINOS_PER_LANE, 1),
                     (_lib.OPT_BP_SINOS_PER_LANE, 17), (_lib.OPT_ADJOINT_VIEWS, -1), (_lib.OPT_ADJOINT_VIEWS, 257)):
        with pytest.raises(VctError) as e:
            plan.set_option(opt, bad)
        assert e.value.status == _lib.ERR_ARG
    plan.close()


def test_tune_plan_picks_from_candidates(gpu, tmp_path, monkeypatch):
    from virtualct_amd import autotune, ops
    monkeypatch.setattr(autotune, "_tuner", autotune.Tuner(str(tmp_path / "tune.json")))
    plan = ops.Plan(_geom(128, 0.35, 160))
    cfg = plan.tune(6, n_sino=6)
    assert cfg["trace_block"] in autotune.TRACE_BLOCKS
    assert cfg["adjoint_views"] in autotune.ADJOINT_VIEWS
    assert cfg["bp_sinos_per_lane"] in autotune.BP_SINOS_PER_LANE
    assert (tmp_path / "tune.json").exists()
    again = autotune.Tuner(str(tmp_path / "tune.json"))
    assert len(again.table) == 3
    plan.close()


def test_batch_forward_fbp_consults_the_tuner(gpu, tmp_path, monkeypatch):
    import torch
    from virtualct_amd import autotune, batch, phantoms
    monkeypatch.setattr(autotune, "_tuner", autotune.Tuner(str(tmp_path / "tune.json")))
    groups = [(_geom(N, 50.0 / N, 160), torch.stack([torch.from_numpy(phantoms.body_phantom(70 + b, N=N, pix_cm=50.0 / N))
                                                     for b in range(2)]).to(gpu)) for N in (64, 96)]
    att = torch.rand((2, 6, 5), device=gpu) * 0.2
    pdf = torch.full((2, 5), 0.2, device=gpu)
    tuned = batch.BatchForwardFBP(groups, att, pdf)
    fixed = batch.BatchForwardFBP(groups, att, pdf, lanes=batch.lanes_for(2), blocks=0)
    keys = [json_key for json_key in autotune.tuner().table]
    assert any("forward gather" in k for k in keys)
    cand = tuned._candidates()
    assert {"lanes": tuned.lanes, "blocks": tuned.blocks} in cand
    s1, i1 = tuned.run()
    s2, i2 = fixed.run()
    for a, b in zip(s1 + i1, s2 + i2):
        assert rel_max(a.cpu().numpy(), b.cpu().numpy()) < 1e-6
