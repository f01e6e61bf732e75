"""ctypes wrapper around oracle/liboracle.so and oracle/_ref/*.so.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py`` as the checker; the product package
(``virtualct_amd``) never imports this module.

``Oracle`` is the C restatement (vct_oracle.c).  ``RefMBFPA`` / ``RefFBP``
bind the reference's own functions compiled from source by ``make -C oracle
ref`` (present only where that build was run).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

# Element table: H C N O P Ca (M/prior_weight_production.cpp:249-255) + Mg Si.
ELEMENT_Z = (1, 6, 7, 8, 15, 20, 12, 14)
ELEMENT_A = (1, 12, 14, 16, 31, 40, 24, 28)


class OrcGeom(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("pix", C.c_double), ("W", C.c_int), ("V", C.c_int),
        ("start_deg", C.c_double), ("sid", C.c_double), ("det_radius", C.c_double),
        ("det_src_iso", C.c_double), ("iso_channel", C.c_double),
        ("max_alpha", C.c_double), ("max_beta", C.c_double), ("xfactor", C.c_double),
    ]


@dataclass
class Geom:
    N: int = 512
    pix: float = 0.1
    W: int = 609
    V: int = 800
    start_deg: float = 0.0
    sid: float = 85.0
    det_radius: float = 142.6
    det_src_iso: float = 85.0
    iso_channel: float = 304.5
    max_alpha: float = 0.326086956
    max_beta: float = 0.326086956
    xfactor: float = 1.0

    def c(self) -> OrcGeom:
        return OrcGeom(self.N, self.pix, self.W, self.V, self.start_deg, self.sid,
                       self.det_radius, self.det_src_iso, self.iso_channel,
                       self.max_alpha, self.max_beta, self.xfactor)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return C.CDLL(path)


class Oracle:
    def __init__(self, path: str | None = None):
        self.lib = L = _load(path or os.path.join(HERE, "liboracle.so"))
        G = C.POINTER(OrcGeom)
        L.orc_view_tables.argtypes = [G, _f64p, _f64p]
        L.orc_channel_table.argtypes = [G, _f64p, _f64p]
        L.orc_trace_ray.argtypes = [G, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int,
                                    _i32p, _f64p, C.c_int, C.POINTER(C.c_int)]
        L.orc_trace_ray.restype = C.c_int
        L.orc_trace_ray_iks.argtypes = [G, _f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int,
                                        _i32p, _f64p, _i32p, C.c_int, C.POINTER(C.c_int)]
        L.orc_trace_ray_iks.restype = C.c_int
        L.orc_segment_counts.argtypes = [G, _i32p]
        L.orc_segment_counts.restype = C.c_int64
        L.orc_attenuation.argtypes = [C.c_char_p, C.c_int, _f64p, _f64p, C.c_int,
                                      _f32p, _f32p, _f64p, C.c_int]
        L.orc_attenuation.restype = C.c_int
        L.orc_forward_ref.argtypes = [G, _f32p, C.c_int, _f32p, _f32p, C.c_int, _f32p]
        L.orc_forward_f64.argtypes = [G, _f32p, C.c_int, _f32p, _f32p, C.c_int, _f64p, C.c_void_p]
        L.orc_normalize.argtypes = [_f32p, C.c_int]
        L.orc_normalize.restype = C.c_float
        L.orc_gaussian_noise_glibc.argtypes = [_f32p, C.c_int, C.c_double, C.c_int]
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_beam_hardening.argtypes = [_f32p, C.c_int, C.c_double, C.c_double,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_const_map.argtypes = [G, C.c_int, _f64p]
        L.orc_filter_ref.argtypes = [G, _f64p, _f32p, _f32p, C.c_int]
        L.orc_filter_f64.argtypes = [G, _f64p, _f32p, _f64p]
        L.orc_mask.argtypes = [C.c_int, _u8p]
        L.orc_backproject_ref.argtypes = [G, _f32p, _f32p]
        L.orc_backproject_f64.argtypes = [G, _f64p, _f64p]
        L.orc_adjoint_f64.argtypes = [G, _f32p, C.c_int, _f32p, _f32p, C.c_int, _f64p, _f64p]
        for fn, out in (("orc_lengthproj_window_ref", _f32p), ("orc_lengthproj_window_f64", _f64p)):
            getattr(L, fn).argtypes = [G, _f32p, C.c_int, C.c_double, C.c_double, C.c_double, out]
            getattr(L, fn).restype = C.c_int
        L.orc_load_measured.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p, _u8p,
                                        C.POINTER(C.c_double)]
        L.orc_load_measured.restype = C.c_int
        L.orc_helical_window.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, _f32p]

    # -- tables
    def view_tables(self, g: Geom):
        c = np.zeros(g.V, np.float64); s = np.zeros(g.V, np.float64)
        self.lib.orc_view_tables(C.byref(g.c()), c, s)
        return c, s

    def channel_table(self, g: Geom):
        a = np.zeros(g.W, np.float64); b = np.zeros(g.W, np.float64)
        self.lib.orc_channel_table(C.byref(g.c()), a, b)
        return a, b

    def trace_ray(self, g: Geom, view: int, ch: int, tables=None):
        c, s = tables[0] if tables else self.view_tables(g)
        px0, py0 = tables[1] if tables else self.channel_table(g)
        ms = 4 * g.N + 64
        vox = np.zeros(ms, np.int32); rr = np.zeros(ms, np.float64); idx = np.zeros(ms, np.int32)
        iks = C.c_int(0)
        n = self.lib.orc_trace_ray_iks(C.byref(g.c()), c, s, px0, py0, view, ch, vox, rr, idx, ms, C.byref(iks))
        if n < 0:
            raise RuntimeError("segment buffer overflow")
        return vox[:n].copy(), rr[:n].copy(), iks.value, idx[:n].copy()

    def segment_counts(self, g: Geom):
        out = np.zeros((g.V, g.W), np.int32)
        tot = self.lib.orc_segment_counts(C.byref(g.c()), out)
        return out, int(tot)

    def attenuation(self, path: str, M: int = 6, nthin: int = 1):
        Z = np.array(ELEMENT_Z[:M], np.float64); A = np.array(ELEMENT_A[:M], np.float64)
        att = np.zeros(M * 1024, np.float32); pdf = np.zeros(1024, np.float32)
        en = np.zeros(1024, np.float64)
        ne = self.lib.orc_attenuation(path.encode(), M, Z, A, nthin, att, pdf, en, 1024)
        if ne < 0:
            raise RuntimeError(f"cannot read spectrum {path}")
        return att[: M * ne].reshape(M, ne).copy(), pdf[:ne].copy(), en[:ne].copy()

    # -- forward
    def forward_ref(self, g: Geom, dens: np.ndarray, att: np.ndarray, pdf: np.ndarray):
        dens = np.ascontiguousarray(dens, np.float32)
        M = dens.shape[0]
        out = np.zeros((g.V, g.W), np.float32)
        self.lib.orc_forward_ref(C.byref(g.c()), dens.reshape(-1), M,
                                 np.ascontiguousarray(att, np.float32).reshape(-1),
                                 np.ascontiguousarray(pdf, np.float32), pdf.shape[0], out)
        return out

    def forward_f64(self, g: Geom, dens, att, pdf, want_L: bool = False):
        dens = np.ascontiguousarray(dens, np.float32)
        M = dens.shape[0]
        out = np.zeros((g.V, g.W), np.float64)
        Lp = np.zeros((g.V, g.W, M), np.float64) if want_L else None
        self.lib.orc_forward_f64(C.byref(g.c()), dens.reshape(-1), M,
                                 np.ascontiguousarray(att, np.float32).reshape(-1),
                                 np.ascontiguousarray(pdf, np.float32), pdf.shape[0], out,
                                 Lp.ctypes.data if want_L else None)
        return (out, Lp) if want_L else out

    def normalize(self, sino: np.ndarray):
        s = np.ascontiguousarray(sino, np.float32).copy()
        mx = self.lib.orc_normalize(s.reshape(-1), s.size)
        return s, mx

    def noise_glibc(self, sino: np.ndarray, sigma: float, order: int = 0):
        s = np.ascontiguousarray(sino, np.float32).copy()
        self.lib.orc_gaussian_noise_glibc(s.reshape(-1), s.size, sigma, order)
        return s

    def philox(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
        return out

    def beam_hardening(self, sino, alpha: float, beta: float):
        s = np.ascontiguousarray(sino, np.float32).copy()
        a, b = C.c_double(), C.c_double()
        self.lib.orc_beam_hardening(s.reshape(-1), s.size, alpha, beta, C.byref(a), C.byref(b))
        return s, a.value, b.value

    # -- FBP
    def const_map(self, g: Geom, filt: int = 2):
        Cm = np.zeros((g.W, g.W), np.float64)
        self.lib.orc_const_map(C.byref(g.c()), filt, Cm)
        return Cm

    def filter_ref(self, g: Geom, proj, filt: int = 2, in_mode: bool = False, Cm=None):
        Cm = self.const_map(g, filt) if Cm is None else Cm
        out = np.zeros((g.V, g.W), np.float32)
        self.lib.orc_filter_ref(C.byref(g.c()), Cm, np.ascontiguousarray(proj, np.float32), out, int(in_mode))
        return out

    def filter_f64(self, g: Geom, proj, filt: int = 2, Cm=None):
        Cm = self.const_map(g, filt) if Cm is None else Cm
        out = np.zeros((g.V, g.W), np.float64)
        self.lib.orc_filter_f64(C.byref(g.c()), Cm, np.ascontiguousarray(proj, np.float32), out)
        return out

    def mask(self, N: int):
        m = np.zeros((N, N), np.uint8)
        self.lib.orc_mask(N, m)
        return m

    def backproject_ref(self, g: Geom, fltd):
        img = np.zeros((g.N, g.N), np.float32)
        self.lib.orc_backproject_ref(C.byref(g.c()), np.ascontiguousarray(fltd, np.float32), img)
        return img

    def backproject_f64(self, g: Geom, fltd):
        img = np.zeros((g.N, g.N), np.float64)
        self.lib.orc_backproject_f64(C.byref(g.c()), np.ascontiguousarray(fltd, np.float64), img)
        return img

    def fbp_ref(self, g: Geom, proj, filt: int = 2, in_mode: bool = False):
        return self.backproject_ref(g, self.filter_ref(g, proj, filt, in_mode))

    def fbp_f64(self, g: Geom, proj, filt: int = 2):
        return self.backproject_f64(g, self.filter_f64(g, proj, filt))

    def adjoint_f64(self, g: Geom, dens, att, pdf, gI):
        dens = np.ascontiguousarray(dens, np.float32)
        M = dens.shape[0]
        grad = np.zeros((M, g.N, g.N), np.float64)
        self.lib.orc_adjoint_f64(C.byref(g.c()), dens.reshape(-1), M,
                                 np.ascontiguousarray(att, np.float32).reshape(-1),
                                 np.ascontiguousarray(pdf, np.float32), pdf.shape[0],
                                 np.ascontiguousarray(gI, np.float64), grad)
        return grad


    def lengthproj_ref(self, g: Geom, proj, z_cm: float, couch_cm: float, iso_det: float = 57.6):
        """npf [N,N] float (reference rounding) of a window proj [V/2*2, w], whose
        rows hold the first w <= g.W channels of the detector; None when the
        view window is invalid."""
        proj = np.ascontiguousarray(proj, np.float32)
        out = np.zeros((g.N, g.N), np.float32)
        rc = self.lib.orc_lengthproj_window_ref(C.byref(g.c()), proj, proj.shape[-1], z_cm, couch_cm, iso_det, out)
        return out if rc == 0 else None

    def lengthproj_f64(self, g: Geom, proj, z_cm: float, couch_cm: float, iso_det: float = 57.6):
        proj = np.ascontiguousarray(proj, np.float32)
        out = np.zeros((g.N, g.N), np.float64)
        rc = self.lib.orc_lengthproj_window_f64(C.byref(g.c()), proj, proj.shape[-1], z_cm, couch_cm, iso_det, out)
        return out if rc == 0 else None


    # -- measured data
    def load_measured(self, in_mode: int, rot: int, H: int, W: int, nskip: int, air, raw):
        """(volume [rot*H, W] float32, mask uint8, max) or None for a shape
        the reference reads out of bounds.  air [6,800,609]; raw PR
        [rot,800,609], IN [rot,H,609]."""
        out = np.zeros((rot * H, W), np.float32)
        mask = np.zeros((rot * H, W), np.uint8)
        mx = C.c_double(0.0)
        rc = self.lib.orc_load_measured(in_mode, rot, H, W, nskip, np.ascontiguousarray(air, np.float32),
                                        np.ascontiguousarray(raw, np.float32), out, mask, C.byref(mx))
        return (out, mask, mx.value) if rc == 0 else None

    def helical_window(self, volume, rot: int, H: int, W: int):
        win = np.zeros((H // 2 * 2, W), np.float32)
        self.lib.orc_helical_window(np.ascontiguousarray(volume, np.float32), rot, H, W, win)
        return win


class RefMBFPA:
    """The reference's own MBFPA functions (oracle/_ref/libref_mbfpa.so)."""

    def __init__(self):
        self.lib = L = _load(os.path.join(REF_DIR, "libref_mbfpa.so"))
        L.ref_attenuation.argtypes = [C.c_char_p, C.c_int, _f32p, _f32p]
        L.ref_attenuation.restype = C.c_int
        L.ref_reprojection.argtypes = [C.c_int, _f32p, _f32p, C.c_double, C.c_int, C.c_int,
                                       _f64p, _f64p, C.c_double, _f32p, _f32p, C.c_int,
                                       C.c_int, C.c_int, C.c_void_p]
        L.ref_reprojection.restype = C.c_int
        L.ref_reprojection_full.argtypes = [C.c_int, _f32p, _f32p, C.c_double, C.c_int, C.c_int,
                                            _f64p, _f64p, C.c_double, _f32p, _f32p, C.c_int,
                                            C.c_int, C.c_int, C.c_int]
        L.ref_reprojection_full.restype = C.c_int
        L.ref_lengthproj.argtypes = [C.c_int, _f32p, _f32p, C.c_double, C.c_int, C.c_int, _f64p, _f64p,
                                     C.c_double, C.c_double]
        L.ref_lengthproj.restype = C.c_int
        L.ref_load_measured.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_char_p,
                                        C.c_char_p, _f32p, _u8p, C.POINTER(C.c_int)]
        L.ref_load_measured.restype = C.c_int

    def load_measured(self, inpr: str, W: int, H_txt: int, rot: int, sino_fact: int, rate_orig: float,
                      proj_file: str, air_file: str):
        """The reference's loadProjectionData on real files: (result, volume
        [rot*H', W], mask, H').  It writes projections.raw into the working
        directory and stops the process when an IN value exceeds 1."""
        H = H_txt // sino_fact
        vol = np.zeros((rot * H, W), np.float32)
        mask = np.zeros((rot * H, W), np.uint8)
        rows = C.c_int(0)
        rc = self.lib.ref_load_measured(inpr.encode(), W, H_txt, rot, sino_fact, rate_orig, proj_file.encode(),
                                        air_file.encode(), vol, mask, C.byref(rows))
        return rc, vol, mask, rows.value

    def lengthproj(self, g: Geom, proj, cosT, sinT, z_cm: float, couch_cm: float):
        """The reference's lengthproj_TOMO: npf [N,N] (zeros outside the mask)."""
        npf = np.zeros((g.N, g.N), np.float32)
        p = np.ascontiguousarray(proj, np.float32).copy()
        rc = self.lib.ref_lengthproj(g.N, npf, p, g.pix, g.W, g.V, np.ascontiguousarray(cosT, np.float64),
                                     np.ascontiguousarray(sinT, np.float64), z_cm, couch_cm)
        if rc != 0:
            raise MemoryError("reference cache allocation failed")
        return npf

    def attenuation(self, path: str, M: int = 6):
        att = np.zeros(M * 1024, np.float32); pdf = np.zeros(1024, np.float32)
        ne = self.lib.ref_attenuation(path.encode(), M, att, pdf)
        return att[: M * ne].reshape(M, ne).copy(), pdf[:ne].copy()

    def forward(self, g: Geom, dens, att, pdf, cosT, sinT, ke_range=None, seg_counts=False):
        dens = np.ascontiguousarray(dens, np.float32).copy()
        M, NE = att.shape
        out = np.zeros((g.V, g.W), np.float32)
        a = np.ascontiguousarray(att, np.float32).reshape(-1)
        p = np.ascontiguousarray(pdf, np.float32)
        if seg_counts:
            cnt = np.zeros((g.V, g.W), np.int32)
            rc = self.lib.ref_reprojection(g.N, dens.reshape(-1), out, g.pix, g.W, g.V, cosT, sinT,
                                           g.xfactor, a, p, M, NE, 0, cnt.ctypes.data)
            if rc != 0:
                raise MemoryError("reference cache allocation failed")
            return out, cnt
        k0, k1 = ke_range if ke_range else (0, NE)
        rc = self.lib.ref_reprojection_full(g.N, dens.reshape(-1), out, g.pix, g.W, g.V, cosT, sinT,
                                            g.xfactor, a, p, M, NE, k0, k1)
        if rc != 0:
            raise MemoryError("reference cache allocation failed")
        return out


class RefFBP:
    """The reference's own FBP functions (oracle/_ref/libref_fbp.so); W fixed at 609."""

    def __init__(self):
        self.lib = L = _load(os.path.join(REF_DIR, "libref_fbp.so"))
        L.ref_fbp_channels.restype = C.c_int
        L.ref_const_map.argtypes = [C.c_int, _f64p]
        L.ref_filter.argtypes = [C.c_int, _f32p, _f32p, C.c_int, C.c_int]
        L.ref_backproject.argtypes = [C.c_int, C.c_double, C.c_int, _f32p, _f64p, _f64p,
                                      C.c_double, _f32p]
        self.W = L.ref_fbp_channels()

    def const_map(self, filt: int = 2):
        Cm = np.zeros((self.W, self.W), np.float64)
        self.lib.ref_const_map(filt, Cm)
        return Cm

    def filter(self, proj, filt: int = 2, in_mode: bool = False):
        proj = np.ascontiguousarray(proj, np.float32)
        out = np.zeros_like(proj)
        self.lib.ref_filter(proj.shape[0], proj, out, filt, int(in_mode))
        return out

    def backproject(self, g: Geom, fltd, cosT, sinT):
        img = np.zeros((g.N, g.N), np.float32)
        self.lib.ref_backproject(g.N, g.pix, g.V, np.ascontiguousarray(fltd, np.float32),
                                 cosT, sinT, g.xfactor, img)
        return img


def ref_available() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, f)) for f in ("libref_mbfpa.so", "libref_fbp.so"))
