"""Torch-facing operations on top of the libvct C ABI.

All tensors are device tensors on a gfx950 GPU; PyTorch provides device
memory and the stream (each call is enqueued on the current torch stream).
Nothing here computes on the CPU: without libvct.so or without a GPU the
functions raise.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import DENS_F32, DENS_U16, FILTER_SHEPP_LOGAN, Geometry, call

_ctx_lock = threading.Lock()
_contexts: dict[int, C.c_void_p] = {}
_plans: dict[tuple, "Plan"] = {}


def _require_gpu(t: torch.Tensor | None = None):
    if not torch.cuda.is_available():
        raise RuntimeError("virtualct_amd needs a ROCm GPU (gfx950); none is visible")
    if t is not None and not t.is_cuda:
        raise ValueError("tensor must be on the GPU")


def context(device: int = 0) -> C.c_void_p:
    """The libvct context of a device (created once per process)."""
    _require_gpu()
    with _ctx_lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = C.c_void_p()
            with torch.cuda.device(device):
                call("vct_init", device, C.byref(ctx))
            _contexts[device] = ctx
        return ctx


class Plan:
    """Device tables for one fan-beam geometry (``vct_plan``)."""

    def __init__(self, geom: Geometry, device: int = 0):
        self.geom = geom
        self.device = device
        self.ctx = context(device)
        self.handle = C.c_void_p()
        with torch.cuda.device(device):
            call("vct_plan_create", self.ctx, C.byref(geom), C.byref(self.handle))
        g = geom
        self.N, self.W, self.V = g.recon_size, g.n_channels, g.n_views
        # "sysmat" (default): the autograd adjoint gathers over a cached
        # voxel-major system matrix, built on first use (8 bytes per segment
        # in each direction, 4.5 GB at 512^2 / 800 x 609; ~70 ms), atomic-free
        # and deterministic, 1.5 ms per call at 512^2; "atomic" re-traces the
        # rays tile by tile (2.7 ms per call, no memory beyond the gradient).
        self.adjoint_mode = "sysmat"
        self._sysmat = None

    def set_option(self, option: int, value: int):
        call("vct_plan_set_option", self.handle, option, value)

    def tune(self, M: int = 6, n_sino: int = 1, what=("trace", "adjoint", "backproject")) -> dict:
        """Time this geometry's launch shapes and keep the fastest
        (``autotune.tune_plan``); returns the chosen plan options."""
        from .autotune import tune_plan
        return tune_plan(self, M, n_sino, what)

    def check_errors(self):
        """``check_errors`` for this plan's device."""
        check_errors(self.device)

    def system_matrix(self):
        """The cached system matrix of this geometry (ray- and voxel-major)."""
        if self._sysmat is None:
            from .batch import SystemMatrix
            self._sysmat = SystemMatrix(self, transpose=True)
        return self._sysmat

    def close(self):
        if self._sysmat is not None:
            self._sysmat.close()
            self._sysmat = None
        if self.handle:
            _lib.lib().vct_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_errors(device: int = 0, stream=None):
    """Wait for the stream (default: the current torch stream) and raise
    ``VctError`` (status internal) if a kernel since the last check gave up
    (a ray beyond its marching step bound, a system-matrix row off its
    count); the flag is cleared either way."""
    s = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream(device).cuda_stream)
    call("vct_check_errors", context(device), s)


def plan_for(geom: Geometry, device: int = 0) -> Plan:
    key = (device, geom.key())
    p = _plans.get(key)
    if p is None:
        p = Plan(geom, device)
        _plans[key] = p
    return p


def _ptr(t: torch.Tensor | None) -> C.c_void_p:
    return C.c_void_p(t.data_ptr() if t is not None else 0)


def _stream(device: int) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dev(plan: Plan):
    return torch.device("cuda", plan.device)


# ----------------------------------------------------------------- forward

def pack_densities(plan: Plan, dens: torch.Tensor):
    """[M,N,N] int16 (mg/cm^3, the reference file format) or float32 (g/cm^3)
    -> (packed [N,N,8], dens_type)."""
    _require_gpu(dens)
    M, N, N2 = dens.shape
    if N != plan.N or N2 != plan.N:
        raise ValueError(f"phantom is {N}x{N2}, geometry expects {plan.N}")
    if not 1 <= M <= 8:
        raise ValueError("1..8 materials supported")
    dens = dens.contiguous()
    if dens.dtype == torch.int16:
        out = torch.empty((N, N, 8), dtype=torch.int16, device=dens.device)
        call("vct_pack_densities_i16", plan.handle, _ptr(dens), M, _ptr(out), _stream(plan.device))
        return out, DENS_U16
    if dens.dtype == torch.float32:
        out = torch.empty((N, N, 8), dtype=torch.float32, device=dens.device)
        call("vct_pack_densities_f32", plan.handle, _ptr(dens), M, _ptr(out), _stream(plan.device))
        return out, DENS_F32
    raise TypeError("densities must be int16 or float32")


def _check_packed(plan: Plan, packed: torch.Tensor, dens_type: int):
    _require_gpu(packed)
    want = torch.int16 if dens_type == DENS_U16 else torch.float32
    if dens_type not in (DENS_U16, DENS_F32) or packed.dtype != want or packed.shape != (plan.N, plan.N, 8) \
            or not packed.is_contiguous():
        raise ValueError(f"packed densities must be contiguous {want} [{plan.N},{plan.N},8] "
                         f"(got {packed.dtype} {tuple(packed.shape)}, type {dens_type})")


def trace_paths(plan: Plan, packed: torch.Tensor, dens_type: int, M: int, counts: bool = False,
                check: bool = True):
    """Per-ray elemental path integrals L [V,W,8] (and segment counters).

    check: wait for the tracer and raise if a ray gave up (``check_errors``);
    pipelines that enqueue many launches pass False and check once."""
    _check_packed(plan, packed, dens_type)
    L = torch.empty((plan.V, plan.W, 8), dtype=torch.float32, device=_dev(plan))
    cnt = torch.empty((plan.V, plan.W), dtype=torch.int32, device=_dev(plan)) if counts else None
    call("vct_trace_paths", plan.handle, _ptr(packed), dens_type, M, _ptr(L), _ptr(cnt), _stream(plan.device))
    if check:
        check_errors(plan.device)
    return (L, cnt) if counts else L


def spectral_sum(plan: Plan, L: torch.Tensor, M: int, att: torch.Tensor, pdf: torch.Tensor,
                 want_dIdL: bool = False):
    """raw [S,V,W], stats [S,2] (int32 holding float bits: max, min), dIdL [S,V,W,8]."""
    att = att.to(device=_dev(plan), dtype=torch.float32).contiguous()
    pdf = pdf.to(device=_dev(plan), dtype=torch.float32).contiguous()
    if att.dim() == 2:
        att = att.unsqueeze(0)
    if pdf.dim() == 1:
        pdf = pdf.unsqueeze(0)
    S, Ma, NE = att.shape
    if Ma != M or pdf.shape != (S, NE):
        raise ValueError(f"attenuation {tuple(att.shape)} / pdf {tuple(pdf.shape)} do not match M={M}")
    _require_gpu(L)
    if L.dtype != torch.float32 or L.shape != (plan.V, plan.W, 8) or not L.is_contiguous():
        raise ValueError(f"path integrals must be contiguous float32 [{plan.V},{plan.W},8], got "
                         f"{L.dtype} {tuple(L.shape)}")
    raw = torch.empty((S, plan.V, plan.W), dtype=torch.float32, device=_dev(plan))
    stats = torch.empty((S, 2), dtype=torch.int32, device=_dev(plan))
    dIdL = torch.empty((S, plan.V, plan.W, 8), dtype=torch.float32, device=_dev(plan)) if want_dIdL else None
    call("vct_spectral_sum", plan.handle, _ptr(L), M, _ptr(att), _ptr(pdf), NE, S, _ptr(raw), _ptr(stats),
         _ptr(dIdL), _stream(plan.device))
    return raw, stats, dIdL


def stats_values(stats: torch.Tensor) -> torch.Tensor:
    """[S,2] int32 bit patterns -> [S,2] float32 (max, min)."""
    return stats.view(torch.float32)


def detector(plan: Plan, raw: torch.Tensor, stats: torch.Tensor, sigma: float = 0.0,
             seed: int = 0x5EED, stream_id: int = 0):
    _require_gpu(raw)
    S = raw.shape[0]
    if raw.dtype != torch.float32 or raw.shape != (S, plan.V, plan.W) or not raw.is_contiguous():
        raise ValueError(f"raw sinograms must be contiguous float32 [S,{plan.V},{plan.W}]")
    if stats.shape != (S, 2) or stats.element_size() != 4 or not stats.is_contiguous():
        raise ValueError("stats must be contiguous 32-bit [S,2]")
    out = torch.empty_like(raw)
    call("vct_detector", plan.handle, _ptr(raw), _ptr(stats), S, float(sigma), seed & (2**64 - 1),
         stream_id & 0xFFFFFFFF, _ptr(out), _stream(plan.device))
    return out


def mbfpa_forward(plan: Plan, dens: torch.Tensor, att: torch.Tensor, pdf: torch.Tensor,
                  sigma: float = 0.0, seed: int = 0x5EED, stream_id: int = 0,
                  packed=None, check: bool = True):
    """Normalised (and optionally noisy) sinograms [S,V,W] of one phantom.

    dens: [M,N,N] int16 or float32 (ignored if ``packed=(tensor, type)``);
    att: [S,M,NE] or [M,NE]; pdf: [S,NE] or [NE].  check: wait and raise if
    a ray gave up (``check_errors``)."""
    if packed is None:
        packed = pack_densities(plan, dens)
    pk, dt = packed
    M = dens.shape[0] if dens is not None else att.shape[-2]
    att = att.to(device=_dev(plan), dtype=torch.float32).contiguous()
    pdf = pdf.to(device=_dev(plan), dtype=torch.float32).contiguous()
    if att.dim() == 2:
        att = att.unsqueeze(0)
    if pdf.dim() == 1:
        pdf = pdf.unsqueeze(0)
    S, Ma, NE = att.shape
    if Ma != M or pdf.shape != (S, NE):
        raise ValueError(f"attenuation {tuple(att.shape)} / pdf {tuple(pdf.shape)} do not match M={M}")
    _check_packed(plan, pk, dt)
    sino = torch.empty((S, plan.V, plan.W), dtype=torch.float32, device=_dev(plan))
    call("vct_mbfpa_forward", plan.handle, _ptr(pk), dt, M, _ptr(att), _ptr(pdf), NE, S, float(sigma),
         seed & (2**64 - 1), stream_id & 0xFFFFFFFF, _ptr(sino), C.c_void_p(0), _stream(plan.device))
    if check:
        check_errors(plan.device)
    return sino


# --------------------------------------------------------------------- FBP

def _sinograms(plan: Plan, x: torch.Tensor, what: str) -> torch.Tensor:
    """[V,W] or [S,V,W] float32 on the GPU -> contiguous [S,V,W]."""
    _require_gpu(x)
    if x.dtype != torch.float32:
        raise ValueError(f"{what} must be float32, got {x.dtype}")
    x = x.contiguous()
    if x.dim() == 2:
        x = x.unsqueeze(0)
    if x.dim() != 3 or x.shape[1:] != (plan.V, plan.W) or x.shape[0] < 1:
        raise ValueError(f"{what} shape {tuple(x.shape)} does not match geometry {plan.V}x{plan.W}")
    return x


def filter_workspace(plan: Plan, n_sino: int) -> torch.Tensor:
    """A workspace for ``filter_sinogram(..., work=)`` of n_sino sinograms."""
    n = int(_lib.lib().vct_filter_workspace_bytes(plan.handle, n_sino))
    return torch.empty(max(n // 4, 1), dtype=torch.float32, device=_dev(plan))


def filter_sinogram(plan: Plan, proj: torch.Tensor, filt: int = FILTER_SHEPP_LOGAN, in_mode: bool = False,
                    work: torch.Tensor | None = None):
    """Filtered sinograms [S,V,W].  work: a caller-owned workspace
    (``filter_workspace``) instead of the context scratch, for callers that
    record the filter in a HIP graph."""
    proj = _sinograms(plan, proj, "sinogram")
    S = proj.shape[0]
    out = torch.empty_like(proj, dtype=torch.float32)
    if work is None:
        call("vct_filter_sinogram", plan.handle, _ptr(proj), S, filt, int(in_mode), _ptr(out), _stream(plan.device))
    else:
        need = int(_lib.lib().vct_filter_workspace_bytes(plan.handle, S))
        if not work.is_cuda or not work.is_contiguous() or work.numel() * work.element_size() < need:
            raise ValueError(f"filter workspace must be a contiguous device buffer of >= {need} bytes")
        call("vct_filter_sinogram_ws", plan.handle, _ptr(proj), S, filt, int(in_mode), _ptr(out), _ptr(work),
             _stream(plan.device))
    return out


def backproject(plan: Plan, fltd: torch.Tensor):
    fltd = _sinograms(plan, fltd, "filtered sinogram")
    S = fltd.shape[0]
    img = torch.empty((S, plan.N, plan.N), dtype=torch.float32, device=fltd.device)
    call("vct_backproject", plan.handle, _ptr(fltd), S, _ptr(img), _stream(plan.device))
    return img


ISO_TO_DETECTOR_CM = 57.6


def lengthproj(plan: Plan, proj: torch.Tensor, z_cm: float, couch_speed_cm: float,
               iso_det_cm: float = ISO_TO_DETECTOR_CM) -> torch.Tensor:
    """npf [N,N]: chord-weighted pixel backprojection of the sinogram window
    proj [V, w] for slice z (the MBFPA program's npf_float.raw).  The window's
    rows hold the first w <= W channels of the plan's detector (measured data
    keep fewer channels; the detector's pitch still places the hits)."""
    _require_gpu(proj)
    proj = proj.contiguous()
    if proj.dtype != torch.float32 or proj.dim() != 2 or proj.shape[0] != plan.V // 2 * 2 \
            or not 2 <= proj.shape[1] <= plan.W:
        raise ValueError(f"proj must be float32 [{plan.V // 2 * 2}, w <= {plan.W}], got {tuple(proj.shape)}")
    npf = torch.empty((plan.N, plan.N), dtype=torch.float32, device=proj.device)
    call("vct_lengthproj_window", plan.handle, _ptr(proj), proj.shape[1], float(z_cm), float(couch_speed_cm),
         float(iso_det_cm), _ptr(npf), _stream(plan.device))
    return npf


def measured_shape(in_mode: bool, rotations: int, height: int, width: int, sino_fact: int = 1,
                   skip_rows: int = 0) -> _lib.Measured:
    """``vct_measured``: PR (in_mode False) or IN data of ``rotations`` x
    ``height`` views (after sino_fact) x ``width`` channels."""
    return _lib.Measured(int(bool(in_mode)), rotations, height, width, sino_fact, skip_rows)


def load_measured(m: _lib.Measured, detector_path: str, airscan_path: str, device: int = 0):
    """Air-scan normalised measured data from the detector and air-scan files
    (the MBFPA program's loadProjectionData): (volume [rot*H, W] float32,
    mask [rot*H, W] uint8, largest IN value) as numpy arrays."""
    n = m.rotations * m.height * m.width
    vol = (C.c_float * n)()
    mask = (C.c_uint8 * n)()
    mx = C.c_double(0.0)
    with torch.cuda.device(device):
        call("vct_load_measured_host", context(device), C.byref(m), detector_path.encode(), airscan_path.encode(),
             C.cast(vol, C.c_void_p), C.cast(mask, C.c_void_p), C.byref(mx))
    shape = (m.rotations * m.height, m.width)
    return (np.frombuffer(vol, np.float32).reshape(shape).copy(), np.frombuffer(mask, np.uint8).reshape(shape).copy(),
            mx.value)


def helical_window(volume: torch.Tensor, rotations: int, height: int) -> torch.Tensor:
    """The npf input window [height/2*2, W] blended from the measured volume
    [rotations*height, W] (M/main.cpp:352-371)."""
    _require_gpu(volume)
    volume = volume.contiguous()
    if volume.dtype != torch.float32 or volume.dim() != 2 or volume.shape[0] != rotations * height:
        raise ValueError(f"volume must be float32 [{rotations * height}, W]")
    W = volume.shape[1]
    win = torch.empty((height // 2 * 2, W), dtype=torch.float32, device=volume.device)
    call("vct_helical_window", _ptr(volume), rotations, height, W, _ptr(win), _stream(volume.device.index or 0))
    return win


def fbp(plan: Plan, proj: torch.Tensor, filt: int = FILTER_SHEPP_LOGAN, in_mode: bool = False,
        bhc: tuple[float, float] | None = None):
    """FBP; bhc=(alpha, beta) fuses the beam-hardening correction into the filter."""
    if bhc is None:
        return backproject(plan, filter_sinogram(plan, proj, filt, in_mode))
    return backproject(plan, filter_sinogram_bhc(plan, proj, bhc[0], bhc[1], filt, in_mode))


def filter_sinogram_bhc(plan: Plan, proj: torch.Tensor, alpha: float, beta: float,
                        filt: int = FILTER_SHEPP_LOGAN, in_mode: bool = False):
    proj = _sinograms(plan, proj, "sinogram")
    S = proj.shape[0]
    out = torch.empty_like(proj, dtype=torch.float32)
    call("vct_filter_sinogram_bhc", plan.handle, _ptr(proj), S, filt, int(in_mode), float(alpha), float(beta),
         _ptr(out), _stream(plan.device))
    return out


def beam_hardening(sino: torch.Tensor, alpha: float, beta: float) -> torch.Tensor:
    """y = -ln p, y' = |y| + alpha |y|^beta, exp(-y') elementwise."""
    _require_gpu(sino)
    sino = sino.contiguous()
    out = torch.empty_like(sino, dtype=torch.float32)
    call("vct_beam_hardening", _ptr(sino), sino.numel(), float(alpha), float(beta), _ptr(out),
         C.c_void_p(torch.cuda.current_stream(sino.device).cuda_stream))
    return out


def minmax(x: torch.Tensor):
    """(max, min) of a float tensor, via the device reduction."""
    _require_gpu(x)
    x = x.contiguous()
    out = (C.c_float * 2)()
    call("vct_minmax", _ptr(x), x.numel(), C.cast(out, C.c_void_p), C.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    return float(out[0]), float(out[1])


# ----------------------------------------------------------------- adjoint

def trace_adjoint(plan: Plan, weights: torch.Tensor, M: int, g: torch.Tensor | None = None,
                  dens_packed_f32: torch.Tensor | None = None, grad: torch.Tensor | None = None,
                  dens_type: int = DENS_F32, check: bool = True):
    """grad [N,N,8] += sum_rays weights[ray,m] * g[ray] * length(ray, voxel).

    weights: float32 [V,W,8]; g: float32 [V,W] or None; dens_packed_f32:
    float32 [N,N,8] (voxels with a negative density get no gradient) or None;
    grad: float32 [N,N,8] to accumulate into (zeros if None)."""
    _require_gpu(weights)
    if weights.dtype != torch.float32 or weights.shape != (plan.V, plan.W, 8):
        raise ValueError(f"weights must be float32 [{plan.V},{plan.W},8], got {weights.dtype} {tuple(weights.shape)}")
    if g is not None and (g.dtype != torch.float32 or g.shape != (plan.V, plan.W)):
        raise ValueError(f"g must be float32 [{plan.V},{plan.W}]")
    if dens_packed_f32 is not None:
        _check_packed(plan, dens_packed_f32, DENS_F32)
    if grad is None:
        grad = torch.zeros((plan.N, plan.N, 8), dtype=torch.float32, device=_dev(plan))
    elif grad.dtype != torch.float32 or grad.shape != (plan.N, plan.N, 8) or not grad.is_contiguous():
        raise ValueError(f"grad must be contiguous float32 [{plan.N},{plan.N},8]")
    call("vct_trace_adjoint", plan.handle, _ptr(dens_packed_f32), dens_type, M, _ptr(weights.contiguous()),
         _ptr(g.contiguous() if g is not None else None), _ptr(grad), _stream(plan.device))
    if check:
        check_errors(plan.device)
    return grad


def sysmat_adjoint(plan: Plan, weights: torch.Tensor, M: int,
                   dens_packed_f32: torch.Tensor | None = None) -> torch.Tensor:
    """The adjoint of ``trace_adjoint`` (g = None) as an atomic-free gather
    over the plan's cached voxel-major system matrix: grad [N,N,8].  Voxels
    whose packed density is negative get zero, as the tracer's sign test."""
    from .batch import batch_adjoint
    sm = plan.system_matrix()
    grad = torch.empty((plan.N, plan.N, 8), dtype=torch.float32, device=_dev(plan))
    w = weights.contiguous().view(plan.V * plan.W, 1, 8)
    batch_adjoint([(sm, w, grad.view(plan.N * plan.N, 1, 8), 1)], M, 1)
    if dens_packed_f32 is not None:
        grad.mul_((dens_packed_f32.view(plan.N, plan.N, 8) >= 0).to(grad.dtype))
    return grad


class MBFPAIntensity(torch.autograd.Function):
    """I[s,v,w] = sum_E pdf_s(E) exp(-sum_m mu_s,m(E) L_m(v,w)) of float densities
    [M,N,N] (g/cm^3).  Backward runs the hand-written adjoint once for all
    spectra (the per-ray weights of the spectra are combined first)."""

    @staticmethod
    def forward(ctx, dens, plan, att, pdf):
        M = dens.shape[0]
        pk, dt = pack_densities(plan, dens.detach().float().contiguous())
        L = trace_paths(plan, pk, dt, M)
        raw, stats, dIdL = spectral_sum(plan, L, M, att, pdf, want_dIdL=True)
        ctx.plan = plan
        ctx.M = M
        ctx.save_for_backward(pk, dIdL)
        return raw

    @staticmethod
    def backward(ctx, grad_raw):
        pk, dIdL = ctx.saved_tensors
        plan, M = ctx.plan, ctx.M
        w = (dIdL * grad_raw.contiguous().unsqueeze(-1)).sum(0)  # [V,W,8]
        if plan.adjoint_mode == "sysmat":
            grad = sysmat_adjoint(plan, w, M, pk)
        else:
            grad = trace_adjoint(plan, w, M, None, pk)
        return grad[..., :M].permute(2, 0, 1).contiguous(), None, None, None


def mbfpa_intensity(dens: torch.Tensor, plan: Plan, att: torch.Tensor, pdf: torch.Tensor):
    return MBFPAIntensity.apply(dens, plan, att, pdf)


@dataclass
class Spectrum:
    """Attenuation table [M,NE] and normalised pdf [NE] of one spectrum."""
    att: "torch.Tensor"
    pdf: "torch.Tensor"
    energy_mev: "torch.Tensor"
    name: str = ""
