"""Batched execution: many phantoms (and geometries) x spectra per launch.

* ``SystemMatrix`` -- the cached segment lists of one geometry
  (``vct_sysmat_build``), ray-major and optionally voxel-major and blocked
  (regrouped by image tile / view block for L2-resident two-stage gathers,
  with a partial-sum workspace per phantom count).
* ``interleave`` / ``deinterleave`` -- [B,M,N,N] <-> the phantom-interleaved
  rows of the batched kernels; ``to_planes`` / ``from_planes`` convert
  between those rows and plain [rows,B,8] (materials last).
* ``BatchedSolver`` -- projected-gradient dual-energy decomposition of B
  phantoms: one gather launch for the forward path integrals of all
  phantoms, one fused loss/weights launch for all spectra, one atomic-free
  gather launch for the adjoint, then the Barzilai-Borwein update; a step can
  be captured once in a HIP graph and replayed.
* ``BatchForwardFBP`` -- forward projection (several spectra, Gaussian
  noise) and FBP of phantoms of several sizes: one persistent launch traces
  every group, then the spectral sum, detector, one stacked MFMA filter GEMM
  and the backprojections; the whole run can be captured in a HIP graph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib, ops
from ._lib import BatchGroup, call
from .autotune import tuner


def _stream(dev: torch.device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _p(t: torch.Tensor | None) -> C.c_void_p:
    return C.c_void_p(t.data_ptr() if t is not None else 0)


class SystemMatrix:
    def __init__(self, plan: ops.Plan, transpose: bool = True, blocked: bool = False, tiles: bool = True):
        self.plan = plan
        self.handle = C.c_void_p()
        dev = torch.device("cuda", plan.device)
        flags = (_lib.SYSMAT_TRANSPOSE if transpose else 0) | (_lib.SYSMAT_BLOCKED if blocked else 0)
        if blocked and not tiles:
            flags |= _lib.SYSMAT_NO_TILE
        with torch.cuda.device(plan.device):
            call("vct_sysmat_build", plan.handle, flags, C.byref(self.handle), _stream(dev))
        L = _lib.lib()
        self.segments = int(L.vct_sysmat_segments(self.handle))
        self.bytes = int(L.vct_sysmat_bytes(self.handle))
        self.transpose = transpose
        self.blocked = blocked
        self._work: dict[tuple[int, int], torch.Tensor] = {}

    def row(self, ray: int):
        """(voxels uint32, lengths float32) of one ray's cached segments, in march order."""
        import numpy as np
        n = C.c_int(0)
        _lib.lib().vct_sysmat_row(self.handle, ray, None, None, 0, C.byref(n), _stream(torch.device("cuda", self.plan.device)))
        vox = np.zeros(max(n.value, 1), np.uint32)
        ln = np.zeros(max(n.value, 1), np.float32)
        call("vct_sysmat_row", self.handle, ray, C.c_void_p(vox.ctypes.data), C.c_void_p(ln.ctypes.data), vox.size,
             C.byref(n), _stream(torch.device("cuda", self.plan.device)))
        return vox[: n.value], ln[: n.value]

    def workspace(self, B: int, slot: int = 0) -> torch.Tensor | None:
        """Partial-sum workspace of the blocked gathers for B phantoms (one
        per slot: gathers that may overlap use different slots); None when
        the matrix has no blocked form."""
        if not self.blocked:
            return None
        key = (B, slot)
        if key not in self._work:
            n = int(_lib.lib().vct_sysmat_workspace(self.handle, B))
            self._work[key] = torch.empty(max(n // 4, 1), dtype=torch.float32,
                                          device=torch.device("cuda", self.plan.device))
        return self._work[key]

    def close(self):
        self._work.clear()
        if self.handle:
            _lib.lib().vct_sysmat_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_planes(nat: torch.Tensor) -> torch.Tensor:
    """[rows,B,8] (materials last) -> the batched kernels' rows, returned in a
    tensor of the same shape: each row of 8*B floats holds materials 0-3 of
    phantom b at 4b, materials 4-5 at 4B+2b and 6-7 at 6B+2b (for B = 1 the
    row is unchanged)."""
    R, B, _ = nat.shape
    out = torch.empty_like(nat)
    flat = out.view(R, 8 * B)
    flat[:, : 4 * B] = nat[:, :, 0:4].reshape(R, 4 * B)
    flat[:, 4 * B: 6 * B] = nat[:, :, 4:6].reshape(R, 2 * B)
    flat[:, 6 * B:] = nat[:, :, 6:8].reshape(R, 2 * B)
    return out


def from_planes(rows: torch.Tensor) -> torch.Tensor:
    """Inverse of ``to_planes``: the batched kernels' rows -> [rows,B,8]."""
    R, B, _ = rows.shape
    flat = rows.reshape(R, 8 * B)
    return torch.cat([flat[:, : 4 * B].reshape(R, B, 4), flat[:, 4 * B: 6 * B].reshape(R, B, 2),
                      flat[:, 6 * B:].reshape(R, B, 2)], dim=2)


def interleave(dens: torch.Tensor) -> torch.Tensor:
    """[B,M,N,N] float -> N*N rows of the batched layout (materials >= M zero)."""
    B, M, N, _ = dens.shape
    nat = torch.zeros((N * N, B, 8), dtype=torch.float32, device=dens.device)
    nat[:, :, :M] = dens.permute(2, 3, 0, 1).reshape(N * N, B, M)
    return to_planes(nat)


def deinterleave(x: torch.Tensor, M: int, N: int) -> torch.Tensor:
    """N*N rows of the batched layout -> [B,M,N,N]."""
    B = x.shape[1]
    return from_planes(x)[:, :, :M].reshape(N, N, B, M).permute(2, 3, 0, 1).contiguous()


def lanes_for(B: int) -> int:
    """Lanes per row of the gather: 1 (one phantom, rows split over 16 lanes)
    or the power of two in 8..64 that holds B phantoms."""
    if B == 1:
        return 1
    lanes = 8
    while lanes < B:
        lanes *= 2
    if lanes > 64:
        raise ValueError("at most 64 phantoms per group")
    return lanes


def _groups(items):
    arr = (BatchGroup * len(items))()
    seen: dict[tuple[int, int], int] = {}
    for k, (sm, d_in, d_out, B) in enumerate(items):
        arr[k].sysmat = sm.handle.value
        arr[k].d_in = d_in.data_ptr()
        arr[k].d_out = d_out.data_ptr()
        arr[k].n_phantoms = B
        arr[k].reserved = 0
        # groups of one launch that share a matrix get separate workspaces
        slot = seen.get((id(sm), B), 0)
        seen[(id(sm), B)] = slot + 1
        work = sm.workspace(B, slot)
        arr[k].d_work = work.data_ptr() if work is not None else None
    return arr


def batch_paths(items, M: int, lanes: int | None = None, blocks: int = 0):
    """items: [(SystemMatrix, densities (NN rows), out L (rays rows), B)] -> one
    launch; rows in the batched layout (``to_planes``)."""
    lanes = lanes or lanes_for(max(it[3] for it in items))
    dev = items[0][1].device
    call("vct_batch_paths", _groups(items), len(items), M, lanes, blocks, _stream(dev))


def batch_adjoint(items, M: int, lanes: int | None = None, blocks: int = 0):
    """items: [(SystemMatrix, w (rays rows), out grad (NN rows), B)] -> one
    launch; rows in the batched layout (``to_planes``)."""
    lanes = lanes or lanes_for(max(it[3] for it in items))
    dev = items[0][1].device
    call("vct_batch_adjoint", _groups(items), len(items), M, lanes, blocks, _stream(dev))


def batch_spectral(plan: ops.Plan, L: torch.Tensor, M: int, att: torch.Tensor, pdf: torch.Tensor):
    """L (rays rows, batched layout) -> raw [B,S,V,W], stats [B,S,2] (float bits in int32)."""
    B = L.shape[1]
    S, _, NE = att.shape
    raw = torch.empty((B, S, plan.V, plan.W), dtype=torch.float32, device=L.device)
    stats = torch.empty((B, S, 2), dtype=torch.int32, device=L.device)
    call("vct_batch_spectral", plan.handle, _p(L), B, M, _p(att), _p(pdf), NE, S, _p(raw), _p(stats),
         _stream(L.device))
    return raw, stats


def _ptr_array(values, ctype=C.c_void_p):
    arr = (ctype * len(values))(*values)
    return arr, C.cast(arr, C.c_void_p)


def batch_spectral_groups(plan: ops.Plan, Ls, Bs, M: int, att: torch.Tensor, pdf: torch.Tensor):
    """Several groups' L (rays rows, batched layout, B_g phantoms each) of the
    plan's detector geometry in one launch -> raw [sum B, S, V, W] and stats
    [sum B, S, 2], the phantoms in group order."""
    if not 1 <= len(Ls) <= 16 or len(Ls) != len(Bs):
        raise ValueError("1..16 groups, one phantom count each")
    for L, B in zip(Ls, Bs):
        if L.dtype != torch.float32 or L.shape != (plan.V * plan.W, B, 8) or not L.is_contiguous():
            raise ValueError(f"path integrals must be contiguous float32 [{plan.V * plan.W},{B},8]")
    S, _, NE = att.shape
    nb = sum(Bs)
    dev = Ls[0].device
    raw = torch.empty((nb, S, plan.V, plan.W), dtype=torch.float32, device=dev)
    stats = torch.empty((nb, S, 2), dtype=torch.int32, device=dev)
    keep_l, ls = _ptr_array([L.data_ptr() for L in Ls])
    keep_b, bs = _ptr_array(list(Bs), C.c_int)
    call("vct_batch_spectral_groups", plan.handle, ls, bs, len(Ls), M, _p(att), _p(pdf), NE, S, _p(raw), _p(stats),
         _stream(dev))
    return raw, stats


def backproject_groups(items, sinos_per_lane: int = 0):
    """items: [(plan, filtered sinograms [S,V,W] float32, images [S,N,N] float32)]
    backprojected in one launch (each image as ``ops.backproject`` gives it)."""
    if not 1 <= len(items) <= 16:
        raise ValueError("1..16 groups")
    for plan, fl, img in items:
        S = fl.shape[0]
        if fl.dtype != torch.float32 or fl.shape != (S, plan.V, plan.W) or not fl.is_contiguous():
            raise ValueError(f"filtered sinograms must be contiguous float32 [S,{plan.V},{plan.W}]")
        if img.dtype != torch.float32 or img.shape != (S, plan.N, plan.N) or not img.is_contiguous():
            raise ValueError(f"images must be contiguous float32 [{S},{plan.N},{plan.N}]")
    keep_p, ps = _ptr_array([it[0].handle.value for it in items])
    keep_f, fs = _ptr_array([it[1].data_ptr() for it in items])
    keep_n, ns = _ptr_array([it[1].shape[0] for it in items], C.c_int)
    keep_i, imgs = _ptr_array([it[2].data_ptr() for it in items])
    call("vct_backproject_groups", ps, fs, ns, imgs, len(items), int(sinos_per_lane), _stream(items[0][1].device))


IL_WIDTHS = (4, 8, 12, 16)


def interleave_sinograms(plan: ops.Plan, fltd: torch.Tensor, k: int, out: torch.Tensor | None = None):
    """Filtered sinograms [S,V,W] -> slices of k, [ceil(S/k),V,W,k] (zeros
    past S), the layout ``backproject_groups_il`` reads."""
    if k not in IL_WIDTHS:
        raise ValueError(f"k must be one of {IL_WIDTHS}")
    S = fltd.shape[0]
    if fltd.dtype != torch.float32 or fltd.shape != (S, plan.V, plan.W) or not fltd.is_contiguous():
        raise ValueError(f"filtered sinograms must be contiguous float32 [S,{plan.V},{plan.W}]")
    shape = ((S + k - 1) // k, plan.V, plan.W, k)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=fltd.device)
    elif out.dtype != torch.float32 or tuple(out.shape) != shape or not out.is_contiguous():
        raise ValueError(f"out must be contiguous float32 {shape}")
    call("vct_interleave_sinograms", plan.handle, _p(fltd), S, k, _p(out), _stream(fltd.device))
    return out


def backproject_groups_il(items, k: int):
    """items: [(plan, interleaved sinograms [ceil(S/k),V,W,k], S, images
    [S,N,N] float32)] backprojected in one launch, k sinograms per lane
    (each image as ``backproject_groups`` gives it)."""
    if not 1 <= len(items) <= 16:
        raise ValueError("1..16 groups")
    if k not in IL_WIDTHS:
        raise ValueError(f"k must be one of {IL_WIDTHS}")
    for plan, il, S, img in items:
        if il.dtype != torch.float32 or tuple(il.shape) != ((S + k - 1) // k, plan.V, plan.W, k) \
                or not il.is_contiguous():
            raise ValueError(f"interleaved sinograms must be contiguous float32 [{(S + k - 1) // k},{plan.V},"
                             f"{plan.W},{k}]")
        if img.dtype != torch.float32 or img.shape != (S, plan.N, plan.N) or not img.is_contiguous():
            raise ValueError(f"images must be contiguous float32 [{S},{plan.N},{plan.N}]")
    keep_p, ps = _ptr_array([it[0].handle.value for it in items])
    keep_f, fs = _ptr_array([it[1].data_ptr() for it in items])
    keep_n, ns = _ptr_array([it[2] for it in items], C.c_int)
    keep_i, imgs = _ptr_array([it[3].data_ptr() for it in items])
    call("vct_backproject_groups_il", ps, fs, ns, imgs, len(items), int(k), _stream(items[0][1].device))


def batch_loss(plan: ops.Plan, L, M: int, att, pdf, target, w=None, loss=None):
    """-> w (rays rows, batched layout: d loss / d L), loss [B]."""
    B = L.shape[1]
    S, _, NE = att.shape
    w = torch.empty_like(L) if w is None else w
    loss = torch.empty(B, dtype=torch.float32, device=L.device) if loss is None else loss
    call("vct_batch_loss", plan.handle, _p(L), B, M, _p(att), _p(pdf), NE, S, _p(target), _p(w), _p(loss),
         _stream(L.device))
    return w, loss


# ---------------------------------------------------------------- solver

class BatchedSolver:
    """Projected Barzilai-Borwein descent on B phantoms of one geometry."""

    def __init__(self, sysmat: SystemMatrix, targets: torch.Tensor, att: torch.Tensor, pdf: torch.Tensor,
                 init: torch.Tensor, alpha0: float = 1e-4, lanes: int | None = None, blocks: int = 0):
        if not sysmat.transpose:
            raise ValueError("the solver needs the voxel-major system matrix (transpose=True)")
        self.sm = sysmat
        self.plan = sysmat.plan
        self.B, self.M = init.shape[0], init.shape[1]
        self.N = init.shape[-1]
        self.att = att.contiguous()
        self.pdf = pdf.contiguous()
        self.target = targets.contiguous()  # [B,S,V,W]
        dev = init.device
        self.x = interleave(init.float())
        rays = self.plan.V * self.plan.W
        self.L = torch.empty((rays, self.B, 8), device=dev)
        self.w = torch.empty_like(self.L)
        self.g = torch.empty_like(self.x)
        self.loss = torch.empty(self.B, device=dev)
        self.prev_x = torch.zeros_like(self.x)
        self.prev_g = torch.zeros_like(self.x)
        self.alpha = torch.full((self.B,), alpha0, device=dev)
        nb = int(_lib.lib().vct_batch_bb_workspace(self.B))
        if nb == 0:
            raise ValueError("the solver takes 1..64 phantoms")
        self.bb_work = torch.empty(nb // 8, dtype=torch.float64, device=dev)
        self.have_prev = False
        self.sm.workspace(self.B)  # allocated here, never under graph capture
        key = ("gather fwd+adj", self.N, self.B, self.M, "blocked" if self.sm.blocked else "direct")
        cfg = tuner().best(key, lambda: self._candidates(), self._time_config) if lanes is None else \
            {"lanes": lanes, "blocks": blocks}
        self.lanes, self.blocks = cfg["lanes"], cfg["blocks"]
        self.graph = None

    # -- autotuning of the gather launch shape
    def _candidates(self):
        base = lanes_for(self.B)
        cus = torch.cuda.get_device_properties(self.x.device).multi_processor_count
        out = []
        for lanes in sorted({base, 8 if base == 1 else min(64, base * 2)}):
            for per_cu in (2, 3, 4, 8):
                out.append({"lanes": lanes, "blocks": cus * per_cu})
        return out

    def _time_config(self, cfg):
        # a step runs both gathers with this shape (the tile-resident forward
        # stage 1 ignores it), so both are timed
        fwd = [(self.sm, self.x, self.L, self.B)]
        adj = [(self.sm, self.w, self.g, self.B)]
        self.w.zero_()
        def pair():
            batch_paths(fwd, self.M, cfg["lanes"], cfg["blocks"])
            batch_adjoint(adj, self.M, cfg["lanes"], cfg["blocks"])
        pair()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            pair()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 3

    # -- one iteration
    def _forward_loss_grad(self):
        batch_paths([(self.sm, self.x, self.L, self.B)], self.M, self.lanes, self.blocks)
        batch_loss(self.plan, self.L, self.M, self.att, self.pdf, self.target, self.w, self.loss)
        batch_adjoint([(self.sm, self.w, self.g, self.B)], self.M, self.lanes, self.blocks)

    def _update(self, first: bool = False):
        """Projected Barzilai-Borwein step (``vct_batch_bb_step``): the
        per-phantom step length from s = x - x_prev, y = g - g_prev (kept on
        the first step), then x_prev = x, g_prev = g, x = max(x - alpha g, 0)."""
        call("vct_batch_bb_step", _p(self.x), _p(self.prev_x), _p(self.prev_g), _p(self.g), _p(self.alpha),
             self.x.shape[0], self.B, int(first), _p(self.bb_work), _stream(self.x.device))

    def step(self):
        """One iteration; returns the per-phantom loss before the update."""
        if self.graph is not None:
            self.graph.replay()
            return self.loss
        self._forward_loss_grad()
        self._update(first=not self.have_prev)
        self.have_prev = True
        return self.loss

    def capture(self):
        """Capture the steady-state iteration in a HIP graph (after one step)."""
        if not self.have_prev:
            self.step()
        torch.cuda.synchronize()
        state = [t.clone() for t in (self.x, self.prev_x, self.prev_g, self.alpha)]
        side = torch.cuda.Stream(self.x.device)
        side.wait_stream(torch.cuda.current_stream(self.x.device))
        with torch.cuda.stream(side):  # warm the allocator outside capture
            self._forward_loss_grad()
            self._update()
        torch.cuda.current_stream(self.x.device).wait_stream(side)
        torch.cuda.synchronize()
        for dst, src in zip((self.x, self.prev_x, self.prev_g, self.alpha), state):
            dst.copy_(src)  # the warm-up must not advance the iteration
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._forward_loss_grad()
            self._update()
        self.graph = g
        return g

    def densities(self) -> torch.Tensor:
        return deinterleave(self.x, self.M, self.N)


# ---------------------------------------------------- batched forward + FBP

@dataclass
class Group:
    plan: ops.Plan
    sysmat: SystemMatrix
    dens: torch.Tensor    # NN rows, batched layout
    L: torch.Tensor       # rays rows, batched layout
    B: int


class BatchForwardFBP:
    """MBFPA (S spectra, noise) + FBP for phantoms of several sizes."""

    def __init__(self, groups, att: torch.Tensor, pdf: torch.Tensor, sigma: float = 0.0, seed: int = 0x5EED,
                 filt: int = 2, lanes: int | None = None, blocks: int = 0, blocked: bool = True):
        """groups: [(Geometry, densities [B,M,N,N] int16 or float32 tensor)];
        blocked: gather through the blocked (tile-regrouped) system matrix."""
        self.att = att.contiguous()
        self.pdf = pdf.contiguous()
        if att.dim() != 3 or pdf.dim() != 2 or pdf.shape != (att.shape[0], att.shape[2]):
            raise ValueError(f"attenuation [S,M,NE] / pdf [S,NE] expected, got {tuple(att.shape)} / {tuple(pdf.shape)}")
        self.S = att.shape[0]
        self.M = att.shape[1]
        self.sigma, self.seed, self.filt = sigma, seed, filt
        self.groups: list[Group] = []
        if not groups:
            raise ValueError("no groups")
        # one stacked filter GEMM serves every group, and the filter depends on
        # the detector and source geometry: only the image grid may differ
        fixed = [f for f, _ in groups[0][0]._fields_ if f not in ("recon_size", "pixel_cm", "reserved0")]
        for geom, dens in groups:
            diff = [f for f in fixed if getattr(geom, f) != getattr(groups[0][0], f)]
            if diff:
                raise ValueError(f"groups may differ only in recon_size and pixel_cm; {geom} differs in {diff}")
            if dens.dim() != 4 or dens.shape[1] != self.M or dens.shape[2:] != (geom.recon_size, geom.recon_size):
                raise ValueError(f"densities must be [B,{self.M},{geom.recon_size},{geom.recon_size}], "
                                 f"got {tuple(dens.shape)}")
        for geom, dens in groups:
            plan = ops.plan_for(geom)
            sm = SystemMatrix(plan, transpose=False, blocked=blocked)
            d = dens.float() / 1000.0 if dens.dtype == torch.int16 else dens.float()
            x = interleave(d)
            sm.workspace(x.shape[1])
            L = torch.empty((plan.V * plan.W, x.shape[1], 8), device=x.device)
            self.groups.append(Group(plan, sm, x, L, x.shape[1]))
        V = {g.plan.V for g in self.groups}
        W = {g.plan.W for g in self.groups}
        if len(V) != 1 or len(W) != 1:
            raise ValueError("all groups must share the detector geometry (views x channels)")
        # the gather's launch shape and each group's backprojection split come
        # from the autotuner unless given (tuned here, never under capture)
        if lanes is None:
            key = ("forward gather", tuple((g.plan.N, g.B) for g in self.groups), self.M,
                   "blocked" if blocked else "direct")
            cfg = tuner().best(key, self._candidates, self._time_config)
            self.lanes, self.blocks = cfg["lanes"], cfg["blocks"]
        else:
            self.lanes, self.blocks = lanes, blocks
        # the grouped backprojection's sinograms per lane (= the interleaved
        # slice width), timed on this run's shapes (random filtered sinograms)
        key = ("grouped backprojection, interleaved", tuple((g.plan.N, g.B * self.S) for g in self.groups),
               self.groups[0].plan.V, self.groups[0].plan.W)
        self.bp_k = tuner().best(key, lambda: [{"sinos_per_lane": k} for k in IL_WIDTHS],
                                 self._time_bp)["sinos_per_lane"]
        self.il = [torch.empty(((g.B * self.S + self.bp_k - 1) // self.bp_k, g.plan.V, g.plan.W, self.bp_k),
                               dtype=torch.float32, device=g.dens.device) for g in self.groups]
        # the filter's packed operand lives in a buffer of this run, not in the
        # context scratch, so a replay of the captured graph shares nothing
        # with later eager calls
        self.n_sino = sum(g.B for g in self.groups) * self.S
        self.filter_work = ops.filter_workspace(self.groups[0].plan, self.n_sino)
        self.graph = None
        self.out = None

    def _candidates(self):
        base = lanes_for(max(g.B for g in self.groups))
        cus = torch.cuda.get_device_properties(self.groups[0].dens.device).multi_processor_count
        out = [{"lanes": base, "blocks": 0}]
        for lanes in sorted({base, 8 if base == 1 else min(64, base * 2)}):
            for per_cu in (2, 4, 8):
                out.append({"lanes": lanes, "blocks": cus * per_cu})
        return out

    def _time_config(self, cfg):
        from .autotune import events_ms
        items = [(g.sysmat, g.dens, g.L, g.B) for g in self.groups]
        return events_ms(lambda: batch_paths(items, self.M, cfg["lanes"], cfg["blocks"]))

    def _time_bp(self, cfg):
        from .autotune import events_ms
        gen = torch.Generator(device=self.groups[0].dens.device).manual_seed(2)
        k = cfg["sinos_per_lane"]
        items = []
        for g in self.groups:
            n = g.B * self.S
            fl = torch.randn((n, g.plan.V, g.plan.W), generator=gen, device=g.dens.device)
            items.append((g.plan, interleave_sinograms(g.plan, fl, k), n,
                          torch.empty((n, g.plan.N, g.plan.N), device=g.dens.device)))
        return events_ms(lambda: backproject_groups_il(items, k))

    def run(self):
        """One pass: the forward gather of every group (one launch), the
        spectral sums of every group (one launch), the detector of all
        sinograms (one launch), one stacked filter GEMM, each group's
        filtered sinograms interleaved in slices of bp_k (one launch per
        group) and the backprojections of every group (one launch)."""
        gs = self.groups
        batch_paths([(g.sysmat, g.dens, g.L, g.B) for g in gs], self.M, self.lanes, self.blocks)
        p0 = gs[0].plan
        raw, stats = batch_spectral_groups(p0, [g.L for g in gs], [g.B for g in gs], self.M, self.att, self.pdf)
        n = raw.shape[0] * raw.shape[1]
        stack = ops.detector(p0, raw.view(n, p0.V, p0.W), stats.view(n, 2), self.sigma, self.seed, 0)
        fltd = ops.filter_sinogram(p0, stack, self.filt, work=self.filter_work)
        sinos, imgs, items = [], [], []
        off = 0
        for g, il in zip(gs, self.il):
            k = g.B * self.S
            img = torch.empty((k, g.plan.N, g.plan.N), dtype=torch.float32, device=stack.device)
            sinos.append(stack[off:off + k])
            imgs.append(img)
            interleave_sinograms(g.plan, fltd[off:off + k], self.bp_k, il)
            items.append((g.plan, il, k, img))
            off += k
        backproject_groups_il(items, self.bp_k)
        self.out = (sinos, imgs)
        return self.out

    def capture(self):
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            out = self.run()
        self.graph = g
        self.out = out
        return g

    def replay(self):
        self.graph.replay()
        return self.out
