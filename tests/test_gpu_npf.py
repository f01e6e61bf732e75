"""GPU npf (chord-weighted pixel backprojection, the MBFPA program's
npf_float.raw) against the reference's own outputs and the oracle."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from conftest import rel_max
from oracle.oracle import Geom

pytestmark = pytest.mark.gpu
TOL = 1e-5

NPF_SMALL = [("n64", 64, 0.7, 80), ("n128", 128, 0.35, 160)]
NPF_WINDOWS = [(0.4, 0.8), (1.3, 0.8), (2.0, 1.2)]


def _ops():
    from virtualct_amd import ops
    return ops


def _geom(N=512, pix=0.1, V=800):
    from virtualct_amd import default_geometry
    return default_geometry(recon_size=N, pixel_cm=pix, n_views=V)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_npf_small_matches_reference(gpu, golden, oracle):
    ops = _ops()
    t = golden.npz("ref_npf.npz")
    for k, (name, N, pix, V) in enumerate(NPF_SMALL):
        plan = ops.plan_for(_geom(N, pix, V))
        proj = np.random.default_rng(100 + k).uniform(0.05, 1.0, (V, 609)).astype(np.float32)
        mask = oracle.mask(N).astype(bool)
        for z, couch in NPF_WINDOWS:
            got = ops.lengthproj(plan, _t(proj, gpu), z, couch).cpu().numpy()
            ref = t[f"{name}_z{z}_c{couch}"]
            assert rel_max(got, ref) < TOL, (name, z, couch)
            assert np.all(got[~mask] == 0)
            f64 = oracle.lengthproj_f64(Geom(N=N, pix=pix, V=V), proj, z, couch)
            assert rel_max(got, f64) < TOL


def test_npf_full_geometry(gpu, golden, oracle):
    """512 x 512 / 800 x 609 on the noise-free 120 kV sinogram: within 1e-5
    of the reference (sample rows) and of the oracle (whole image)."""
    ops = _ops()
    sino = golden.sigma0_120kv()
    t = golden.npz("ref_npf.npz")
    plan = ops.plan_for(_geom())
    d = _t(sino, gpu)
    for z, couch in NPF_WINDOWS:
        got = ops.lengthproj(plan, d, z, couch).cpu().numpy()
        assert rel_max(got[[0, 128, 256, 384]], t[f"npf_sigma0_z{z}_c{couch}_rows"]) < TOL
        ref = oracle.lengthproj_ref(Geom(), sino, z, couch)
        assert hashlib.sha256(ref.tobytes()).hexdigest() == golden.hashes()[f"npf_sigma0_z{z}_c{couch}_sha256"]
        assert rel_max(got, ref) < TOL
        # most pixels agree to the bit; acos of the device library may differ in the last place
        assert np.count_nonzero(got != ref) < 0.05 * got.size


def test_npf_rejects_window_before_view_zero(gpu):
    ops = _ops()
    import torch
    plan = ops.plan_for(_geom(64, 0.7, 80))
    proj = torch.ones((80, 609), device=gpu)
    with pytest.raises(RuntimeError):
        ops.lengthproj(plan, proj, 0.0, 0.8)
    with pytest.raises(RuntimeError):
        ops.lengthproj(plan, proj, 0.3, 0.0)
    with pytest.raises(ValueError):
        ops.lengthproj(plan, proj[:40], 0.4, 0.8)
