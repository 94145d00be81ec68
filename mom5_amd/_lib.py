"""ctypes binding of libmom5_gpu.so (the C-ABI in include/mom5_gpu.h).

The library is built in-tree by ``python -m mom5_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in mom5_amd.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmom5_gpu.so")

MOM_GPU_OK = 0
MOM_ADVECT_QUICKER = 5
MOM_ADVECT_MDFL_SWEBY = 9
MOM_ADVECT_DST_LINEAR = 10

_vp = C.c_void_p
_dp = C.c_void_p


class Domain(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("isd", "ied", "jsd", "jed", "isc", "iec", "jsc", "jec",
                                        "nk", "cyclic_x", "cyclic_y")]


class Grid(C.Structure):
    _fields_ = [(n, _vp) for n in ("dxt", "dyt", "dat", "datr", "dxte", "dyte", "dxtn", "dytn",
                                    "dxu", "dyu", "dxur", "dyur", "f", "kmt", "kmu", "tmask",
                                    "umask", "dzt")]


class BtParams(C.Structure):
    _fields_ = [("nts", C.c_int), ("dtbt", C.c_double), ("pred_corr_gamma", C.c_double),
                ("rho0", C.c_double), ("stepping_a", C.c_int), ("initsum_with_bar", C.c_int)]


class BtVariant(C.Structure):
    _fields_ = [("cgrid", C.c_int), ("pressure_based", C.c_int), ("grav", C.c_double)] + \
               [(n, _vp) for n in ("x_tau", "x_bar_tau", "udrho_tau", "forcing_bt", "thick",
                                   "tmasken", "patm", "rho_g", "dpatm_dt", "pme", "river",
                                   "source", "smooth", "x_bar_taup1", "udrho_taup1",
                                   "conv_rho_ud_t")]


class TracerStep(C.Structure):
    _fields_ = [("ntr", C.c_int), ("T_taum1", C.POINTER(_vp)), ("T_taup1", C.POINTER(_vp)),
                ("dtype", C.POINTER(C.c_int)), ("th_tendency", C.POINTER(_vp)),
                ("stf", C.POINTER(_vp)), ("btf", C.POINTER(_vp)), ("index_salt", C.c_int),
                ("uhrho_et", _vp), ("vhrho_nt", _vp), ("wrho_bt", _vp), ("rho_dzt_tau", _vp),
                ("rho_dzt_taum1", _vp), ("rho_dztr", _vp), ("rho_dzt_taup1", _vp),
                ("dzwt", _vp), ("diff_cbt", _vp), ("rho0", C.c_double), ("dtime", C.c_double),
                ("aidif", C.c_double), ("scheme", C.POINTER(C.c_int)),
                ("min_limit", C.POINTER(C.c_double)), ("max_limit", C.POINTER(C.c_double)),
                ("index_temp", C.c_int), ("tmask_limit_ts_same", C.c_int),
                ("limit_with_upwind", C.c_int)]


class BtFields(C.Structure):
    _fields_ = [(n, _vp) for n in ("eta_t_tau", "eta_t_bar_tau", "udrho_tau", "forcing_bt",
                                    "thicku", "patm", "pme", "river", "source", "eta_smooth",
                                    "rho_g", "eta_t_bar_taup1", "udrho_taup1", "conv_rho_ud_t")]


# name -> (restype, argtypes)
_PROTOS = {
    "mom_gpu_create": (C.c_int, [C.POINTER(Domain), C.POINTER(Grid), C.POINTER(_vp)]),
    "mom_gpu_destroy": (C.c_int, [_vp]),
    "mom_gpu_sync": (C.c_int, [_vp]),
    "mom_gpu_last_error": (C.c_char_p, [_vp]),
    "mom_gpu_malloc": (C.c_int, [C.POINTER(_vp), C.c_size_t]),
    "mom_gpu_free": (C.c_int, [_vp]),
    "mom_gpu_memcpy_htod": (C.c_int, [_vp, _vp, C.c_size_t]),
    "mom_gpu_memcpy_dtoh": (C.c_int, [_vp, _vp, C.c_size_t]),
    "mom_gpu_memcpy_dtod": (C.c_int, [_vp, _vp, C.c_size_t]),
    "mom_gpu_copy": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "mom_gpu_advection_velocity": (C.c_int, [_vp] + [_vp] * 10),
    "mom_gpu_barotropic_forcing": (C.c_int, [_vp] * 5),
    "mom_gpu_update_ocean_barotropic_variant": (C.c_int, [_vp, C.POINTER(BtParams),
                                                          C.POINTER(BtVariant)]),
    "mom_gpu_eta_and_pbot_update": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double,
                                              C.c_double] + [_dp] * 10),
    "mom_gpu_set_remap_metrics": (C.c_int, [_vp] + [_dp] * 5),
    "mom_gpu_barotropic_closure": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double] + [_dp] * 10),
    "mom_gpu_ensemble_create": (C.c_int, [C.c_int, C.POINTER(_vp), C.POINTER(_vp)]),
    "mom_gpu_ensemble_destroy": (C.c_int, [_vp]),
    "mom_gpu_ensemble_step_tracers": (C.c_int, [_vp, C.POINTER(TracerStep)]),
    "mom_gpu_ensemble_update_barotropic": (C.c_int, [_vp, C.POINTER(BtParams),
                                                     C.POINTER(BtFields)]),
    "mom_gpu_graph_begin": (C.c_int, [_vp]),
    "mom_gpu_graph_end": (C.c_int, [_vp, C.POINTER(C.c_void_p)]),
    "mom_gpu_graph_launch": (C.c_int, [_vp, _vp]),
    "mom_gpu_graph_nodes": (C.c_int, [_vp]),
    "mom_gpu_graph_destroy": (C.c_int, [_vp]),
    "mom_gpu_memset": (C.c_int, [_vp, C.c_int, C.c_size_t]),
    "mom_gpu_fill": (C.c_int, [_vp, _vp, C.c_double, C.c_size_t]),
    "mom_gpu_update_domains": (C.c_int, [_vp, _vp, C.c_int]),
    "mom_gpu_set_tuning": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "mom_gpu_get_tuning": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_int)]),
    "mom_gpu_mark": (C.c_int, [_vp, C.c_int]),
    "mom_gpu_elapsed": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "mom_gpu_horz_advect_tracer": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                             C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp]),
    "mom_gpu_vert_advect_tracer": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mom_gpu_advect_sweby_all": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), _dp, _dp, _dp, _dp,
                                           C.c_double, C.POINTER(_vp), C.POINTER(_vp)]),
    "mom_gpu_compute_tmask_limit": (C.c_int, [_vp, C.c_int, C.POINTER(_vp),
                                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "mom_gpu_update_tracer_explicit": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_double, _dp]),
    "mom_gpu_invtri": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp,
                                 C.c_double]),
    "mom_gpu_vert_diffuse_implicit": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp),
                                                C.POINTER(_vp), _dp, C.POINTER(_vp), C.c_int,
                                                C.c_double, _dp, _dp, C.c_double, C.c_double]),
    "mom_gpu_update_ocean_barotropic": (C.c_int, [_vp, C.POINTER(BtParams), _dp, _dp, _dp, _dp,
                                                  _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                                  _dp]),
    "mom_gpu_step_tracers": (C.c_int, [_vp, C.POINTER(TracerStep)]),
    "mom_synth_tracer": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_ulonglong, _vp]),
    "mom_synth_loguniform": (C.c_int, [_vp, _vp, C.c_int, C.c_ulonglong, C.c_double,
                                       C.c_double]),
    "mom_synth_profile": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "mom_synth_transports_unit": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp,
                                            _vp]),
    "mom_synth_maxabs": (C.c_int, [_vp, _vp, C.c_size_t, C.POINTER(C.c_double)]),
    "mom_synth_transports_combine": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double]),
    "mom_synth_cfl": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_double, _vp,
                                C.POINTER(C.c_double)]),
    "mom_synth_transports_finish": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp]),
    "mom_gpu_tracer_error": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_double,
                                       C.POINTER(C.c_double)]),
    "mom_gpu_tracer_content_f32": (C.c_int, [_vp, _vp, _vp, C.POINTER(C.c_double)]),
    "mom_gpu_tracer_content": (C.c_int, [_vp, _dp, _dp, C.POINTER(C.c_double)]),
    "mom_gpu_tracer_content_noalloc": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
}

EXPORTED = tuple(_PROTOS)

_lib = None


class MomGpuError(RuntimeError):
    pass


def load(path: str | None = None):
    """Load (once) and return the library; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise MomGpuError(f"{p} not built: run `python -m mom5_amd.build` (no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, ctx=None, what: str = ""):
    if rc != MOM_GPU_OK:
        msg = load().mom_gpu_last_error(ctx)
        raise MomGpuError(f"{what} failed ({rc}): {msg.decode(errors='replace') if msg else ''}")
