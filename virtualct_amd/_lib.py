"""ctypes binding of libvct.so (the C ABI declared in include/vct.h).

The library is loaded from ``virtualct_amd/lib/libvct.so`` (built in-tree by
``virtualct_amd._build``).  There is no fallback: if the library or a gfx950
device is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libvct.so")

VCT_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of memory",
          4: "I/O error", 5: "unsupported", 6: "internal limit"}
DENS_U16, DENS_F32 = 0, 1
FILTER_RAMLAK, FILTER_SHEPP_LOGAN = 1, 2
OPT_EXACT_TRACE = 1
OPT_MAX_STEPS = 2
OPT_EXACT_BACKPROJECT = 4
OPT_TRACE_ORBIT = 5
OPT_TRACE_BLOCK = 6
OPT_BP_SINOS_PER_LANE = 7
OPT_ADJOINT_VIEWS = 8
ERR_ARG = 1
ERR_INTERNAL = 6
SYSMAT_TRANSPOSE = 1
SYSMAT_BLOCKED = 2
SYSMAT_NO_TILE = 4


class VctError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")
        self.status = status


class Geometry(C.Structure):
    """Mirror of ``vct_geometry`` (include/vct.h)."""

    _fields_ = [
        ("recon_size", C.c_int32), ("n_channels", C.c_int32), ("n_views", C.c_int32),
        ("reserved0", C.c_int32), ("pixel_cm", C.c_double), ("start_view_deg", C.c_double),
        ("sid_cm", C.c_double), ("det_radius_cm", C.c_double), ("det_src_iso_cm", C.c_double),
        ("iso_channel", C.c_double), ("max_alpha", C.c_double), ("max_beta", C.c_double),
        ("xfactor", C.c_double),
    ]

    def key(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)

    def __repr__(self):
        return ("Geometry(N=%d, pix=%g, W=%d, V=%d, xfactor=%g)"
                % (self.recon_size, self.pixel_cm, self.n_channels, self.n_views, self.xfactor))


_lib = None
_lock = threading.Lock()

_P = C.c_void_p
_SIGS = {
    "vct_geometry_default": (None, [C.POINTER(Geometry)]),
    "vct_status_string": (C.c_char_p, [C.c_int]),
    "vct_last_error": (C.c_char_p, []),
    "vct_abi_version": (C.c_int, []),
    "vct_init": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "vct_terminate": (C.c_int, [_P]),
    "vct_check_errors": (C.c_int, [_P, _P]),
    "vct_plan_create": (C.c_int, [_P, C.POINTER(Geometry), C.POINTER(_P)]),
    "vct_plan_destroy": (C.c_int, [_P]),
    "vct_plan_set_option": (C.c_int, [_P, C.c_int, C.c_int]),
    "vct_attenuation_coefficient": (C.c_int, [C.c_char_p, C.c_int, _P, _P, C.c_int, _P, _P,
                                               C.POINTER(C.c_int), _P]),
    "vct_const_map": (C.c_int, [C.POINTER(Geometry), C.c_int, _P]),
    "vct_pack_densities_i16": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "vct_pack_densities_f32": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "vct_trace_paths": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, _P]),
    "vct_spectral_sum": (C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
    "vct_detector": (C.c_int, [_P, _P, _P, C.c_int, C.c_double, C.c_uint64, C.c_uint32, _P, _P]),
    "vct_forward_workspace_bytes": (C.c_size_t, [_P, C.c_int]),
    "vct_noalloc_scratch_bytes": (C.c_size_t, []),
    "vct_mbfpa_forward_noalloc": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, C.c_double,
                                             C.c_uint64, C.c_uint32, _P, _P]),
    "vct_mbfpa_forward": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, C.c_double,
                                     C.c_uint64, C.c_uint32, _P, _P, _P]),
    "vct_filter_sinogram": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "vct_filter_workspace_bytes": (C.c_size_t, [_P, C.c_int]),
    "vct_filter_sinogram_ws": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
    "vct_backproject": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "vct_backproject_groups": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, _P]),
    "vct_interleaved_bytes": (C.c_size_t, [_P, C.c_int, C.c_int]),
    "vct_interleave_sinograms": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P]),
    "vct_backproject_groups_il": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, _P]),
    "vct_lengthproj": (C.c_int, [_P, _P, C.c_double, C.c_double, C.c_double, _P, _P]),
    "vct_lengthproj_window": (C.c_int, [_P, _P, C.c_int, C.c_double, C.c_double, C.c_double, _P, _P]),
    "vct_beam_hardening": (C.c_int, [_P, C.c_int, C.c_double, C.c_double, _P, _P]),
    "vct_filter_sinogram_bhc": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _P, _P]),
    "vct_minmax": (C.c_int, [_P, C.c_int, _P, _P]),
    "vct_trace_adjoint": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
    "vct_reprojection": (C.c_int, [_P, C.POINTER(Geometry), _P, C.c_int, _P, _P, C.c_int, _P]),
    "vct_gaussian_noise": (C.c_int, [_P, _P, C.c_int, C.c_double, C.c_uint64, _P]),
    "vct_filter_sinogram_host": (C.c_int, [_P, C.POINTER(Geometry), _P, C.c_int, C.c_int, _P]),
    "vct_reconstruct_image_host": (C.c_int, [_P, C.POINTER(Geometry), _P, _P]),
    "vct_lengthproj_host": (C.c_int, [_P, C.POINTER(Geometry), _P, C.c_double, C.c_double, C.c_double, _P]),
    "vct_lengthproj_window_host": (C.c_int, [_P, C.POINTER(Geometry), _P, C.c_int, C.c_double, C.c_double,
                                             C.c_double, _P]),
}


class Measured(C.Structure):
    """Mirror of ``vct_measured``."""

    _fields_ = [("in_mode", C.c_int32), ("rotations", C.c_int32), ("height", C.c_int32), ("width", C.c_int32),
                ("sino_fact", C.c_int32), ("skip_rows", C.c_int32)]


_SIGS.update({
    "vct_measured_normalise": (C.c_int, [C.POINTER(Measured), _P, _P, _P, _P, _P, _P]),
    "vct_helical_window": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "vct_load_measured_host": (C.c_int, [_P, C.POINTER(Measured), C.c_char_p, C.c_char_p, _P, _P,
                                         C.POINTER(C.c_double)]),
    "vct_helical_window_host": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P]),
})


class BatchGroup(C.Structure):
    """Mirror of ``vct_batch_group``."""

    _fields_ = [("sysmat", C.c_void_p), ("d_in", C.c_void_p), ("d_out", C.c_void_p),
                ("n_phantoms", C.c_int32), ("reserved", C.c_int32), ("d_work", C.c_void_p)]


_SIGS.update({
    "vct_sysmat_build": (C.c_int, [_P, C.c_int, C.POINTER(_P), _P]),
    "vct_sysmat_destroy": (C.c_int, [_P]),
    "vct_sysmat_segments": (C.c_int64, [_P]),
    "vct_sysmat_row": (C.c_int, [_P, C.c_int, _P, _P, C.c_int, C.POINTER(C.c_int), _P]),
    "vct_sysmat_bytes": (C.c_size_t, [_P]),
    "vct_sysmat_workspace": (C.c_size_t, [_P, C.c_int]),
    "vct_batch_bb_workspace": (C.c_size_t, [C.c_int]),
    "vct_batch_bb_step": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "vct_batch_paths": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "vct_batch_adjoint": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "vct_batch_spectral": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, _P, _P, _P]),
    "vct_batch_spectral_groups": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, _P, _P,
                                            _P]),
    "vct_batch_loss": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
})
_OPTIONAL = {}


def lib() -> C.CDLL:
    """Load libvct.so once; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise FileNotFoundError(
                    f"{LIB_PATH} is not built; run `python -m virtualct_amd._build` "
                    "(there is no CPU fallback)")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            for name, (res, args) in _OPTIONAL.items():
                fn = getattr(L, name, None)
                if fn is not None:
                    fn.restype = res
                    fn.argtypes = args
            _lib = L
        return _lib


def check(status: int, where: str):
    if status != VCT_OK:
        msg = lib().vct_last_error()
        raise VctError(status, where, msg.decode() if msg else "")


def call(name: str, *args):
    fn = getattr(lib(), name)
    st = fn(*args)
    check(st, name)
    return st


def default_geometry(**overrides) -> Geometry:
    g = Geometry()
    lib().vct_geometry_default(C.byref(g))
    for k, v in overrides.items():
        if not hasattr(g, k):
            raise AttributeError(f"vct_geometry has no field {k}")
        setattr(g, k, v)
    return g


def exported_symbols():
    return sorted(_SIGS)
