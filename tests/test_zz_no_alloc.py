"""vct_mbfpa_forward_noalloc (the MBFPA forward whose work buffer is a static
device array) against vct_mbfpa_forward, both called through the C ABI:
bit-identical sinograms on two calls in a row, the second smaller, and a
request beyond the array's size rejected with VCT_ERR_ARG."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _forward_pair(gpu, N, pix, V, S, sigma, seed_ph):
    """(no-allocation sinograms, workspace-entry sinograms) of one phantom."""
    import torch
    from virtualct_amd import _lib, default_geometry, ops, phantoms
    plan = ops.plan_for(default_geometry(recon_size=N, pixel_cm=pix, n_views=V))
    ph = torch.from_numpy(phantoms.body_phantom(seed_ph, N=N, pix_cm=pix)).to(gpu)
    pk, dt = ops.pack_densities(plan, ph)
    g = torch.Generator().manual_seed(seed_ph)
    NE = 9
    att = (torch.rand((S, 6, NE), generator=g) * 0.3 + 0.05).to(gpu)
    pdf = torch.rand((S, NE), generator=g).to(gpu)
    pdf /= pdf.sum(1, keepdim=True)
    L = _lib.lib()
    stream = C.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    out = []
    for noalloc in (True, False):
        sino = torch.full((S, V, 609), float("nan"), device=gpu)
        args = (plan.handle, C.c_void_p(pk.data_ptr()), dt, 6, C.c_void_p(att.data_ptr()),
                C.c_void_p(pdf.data_ptr()), NE, S, sigma, 0x5EED, 3, C.c_void_p(sino.data_ptr()))
        if noalloc:
            st = L.vct_mbfpa_forward_noalloc(*args, stream)
        else:
            nb = int(L.vct_forward_workspace_bytes(plan.handle, S))
            work = torch.empty(nb // 4 + 1, device=gpu)
            st = L.vct_mbfpa_forward(*args, C.c_void_p(work.data_ptr()), stream)
        assert st == 0, L.vct_last_error().decode()
        torch.cuda.synchronize()
        out.append(sino.cpu().numpy())
    return out


def test_noalloc_forward_equals_workspace_forward(gpu):
    from virtualct_amd import _lib
    assert int(_lib.lib().vct_noalloc_scratch_bytes()) == 64 << 20
    print("call 1: 512^2, 800 views, 3 spectra, noise on", flush=True)
    a, b = _forward_pair(gpu, 512, 0.1, 800, 3, 5.37e-5, 101)
    assert np.isfinite(a).all() and np.array_equal(a, b)
    print("call 2: 128^2, 160 views, 1 spectrum, noise off", flush=True)
    a, b = _forward_pair(gpu, 128, 0.35, 160, 1, 0.0, 102)
    assert np.isfinite(a).all() and np.array_equal(a, b)


def test_noalloc_forward_rejects_oversized_request(gpu):
    import torch
    from virtualct_amd import _lib, default_geometry, ops
    L = _lib.lib()
    plan = ops.plan_for(default_geometry())
    S = 30  # 15.6 MB of path integrals + 30 x 1.95 MB of raw sums > 64 MiB
    assert int(L.vct_forward_workspace_bytes(plan.handle, S)) > int(L.vct_noalloc_scratch_bytes())
    pk = torch.zeros((512, 512, 8), dtype=torch.int16, device=gpu)
    att = torch.zeros((S, 6, 2), device=gpu)
    pdf = torch.ones((S, 2), device=gpu)
    sino = torch.empty((S, 800, 609), device=gpu)
    st = L.vct_mbfpa_forward_noalloc(plan.handle, C.c_void_p(pk.data_ptr()), 0, 6, C.c_void_p(att.data_ptr()),
                                     C.c_void_p(pdf.data_ptr()), 2, S, 0.0, 1, 0, C.c_void_p(sino.data_ptr()),
                                     C.c_void_p(torch.cuda.current_stream(gpu).cuda_stream))
    assert st == 1  # VCT_ERR_ARG
    assert b"no-allocation" in L.vct_last_error()
