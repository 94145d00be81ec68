"""Host-side mirror of MOM5's hot-path interface over the MI355X kernels.

Method names and argument meaning follow the reference routines
(``src/mom5`` of mom-ocean/MOM5):

==============================  ============================================
``horz_advect_tracer``          ocean_tracers/ocean_tracer_advect.F90:1898
``vert_advect_tracer``          ocean_tracers/ocean_tracer_advect.F90:2095
``advect_tracer_sweby_all``     ocean_tracers/ocean_tracer_advect.F90:4104
``compute_tmask_limit``         ocean_tracers/ocean_tracer.F90:2788
``update_tracer_explicit``      ocean_tracers/ocean_tracer.F90:2341-2350
``invtri``                      ocean_core/ocean_util.F90:153
``vert_diffuse_implicit``       ocean_param/vertical/ocean_vert_mix.F90:3426
``update_ocean_barotropic``     ocean_core/ocean_barotropic.F90:2763
``update_ocean_tracer``         ocean_tracers/ocean_tracer.F90:2243 (order)
==============================  ============================================

Every array argument is a :class:`DeviceArray` (or a raw device address);
errors raise :class:`MomGpuError` with the library's message, as the
reference's ``mpp_error(FATAL, ...)`` would stop the run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BtParams, BtVariant, Domain, Grid, MomGpuError, TracerStep, check, load

ADVECT_QUICKER = _lib.MOM_ADVECT_QUICKER
ADVECT_MDFL_SWEBY = _lib.MOM_ADVECT_MDFL_SWEBY
ADVECT_DST_LINEAR = _lib.MOM_ADVECT_DST_LINEAR
SCHEMES = {"quicker": ADVECT_QUICKER, "mdfl_sweby": ADVECT_MDFL_SWEBY,
           "dst_linear": ADVECT_DST_LINEAR}


class DeviceArray:
    """A device allocation holding one field in its Fortran memory image."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(load().mom_gpu_malloc(C.byref(p), self.nbytes), None, "mom_gpu_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes, (arr.shape, self.shape)
        check(load().mom_gpu_memcpy_htod(self.ptr, arr.ctypes.data, self.nbytes), None, "H2D")

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        check(load().mom_gpu_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), None, "D2H")
        return out

    def zero(self):
        check(load().mom_gpu_memset(self.ptr, 0, self.nbytes), None, "memset")

    def copy(self):
        """A new device array with the same contents."""
        d = DeviceArray(self.shape, self.dtype)
        check(load().mom_gpu_memcpy_dtod(d.ptr, self.ptr, self.nbytes), None, "D2D")
        return d

    def element_size(self):
        return self.dtype.itemsize

    def free(self):
        if getattr(self, "ptr", None):
            load().mom_gpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    return int(a)


def _elem_size(a):
    return int(np.dtype(a.dtype).itemsize)


def _ptab(arrs):
    tab = (C.c_void_p * max(len(arrs), 1))()
    for n, a in enumerate(arrs):
        tab[n] = _p(a)
    return tab


def tracer_step_struct(T_taum1s, T_taup1s, *, uhrho_et, vhrho_nt, wrho_bt, rho_dzt_tau, dtime,
                       stfs=None, btfs=None, dzwt=None, diff_cbt=None, index_salt=1, rho0=1035.0,
                       aidif=1.0, th_tendencies=None, rho_dzt_taum1=None, rho_dztr=None,
                       rho_dzt_taup1=None, schemes=None, min_limits=None, max_limits=None,
                       index_temp=0, ts_same=True, limit_with_upwind=False):
    """The mom_gpu_tracer_step_t of one domain and the ctypes tables it
    points to (keep both alive until the call returns).  schemes: per tracer
    "mdfl_sweby" (default) or "quicker"; Quicker tracers need min_limits and
    max_limits (their tmask_limit ranges)."""
    n = len(T_taum1s)
    dt = (C.c_int * max(n, 1))(*[_elem_size(a) for a in T_taum1s])
    keep = [dt]
    sch = lo = hi = None
    if schemes is not None:
        sch = (C.c_int * max(n, 1))(*[SCHEMES[x] for x in schemes])
        keep.append(sch)
        if min_limits is not None:
            lo = (C.c_double * max(n, 1))(*[float(x) for x in min_limits])
            hi = (C.c_double * max(n, 1))(*[float(x) for x in max_limits])
            keep += [lo, hi]

    def tab(xs):
        if xs is None:
            return None
        t = _ptab(xs)
        keep.append(t)
        return C.cast(t, C.POINTER(C.c_void_p))

    s = TracerStep(ntr=n, T_taum1=tab(T_taum1s), T_taup1=tab(T_taup1s), dtype=dt,
                   th_tendency=tab(th_tendencies), stf=tab(stfs), btf=tab(btfs),
                   index_salt=int(index_salt), uhrho_et=_p(uhrho_et), vhrho_nt=_p(vhrho_nt),
                   wrho_bt=_p(wrho_bt), rho_dzt_tau=_p(rho_dzt_tau),
                   rho_dzt_taum1=_p(rho_dzt_taum1), rho_dztr=_p(rho_dztr),
                   rho_dzt_taup1=_p(rho_dzt_taup1), dzwt=_p(dzwt), diff_cbt=_p(diff_cbt),
                   rho0=float(rho0), dtime=float(dtime), aidif=float(aidif), scheme=sch,
                   min_limit=lo, max_limit=hi, index_temp=int(index_temp),
                   tmask_limit_ts_same=int(bool(ts_same)),
                   limit_with_upwind=int(bool(limit_with_upwind)))
    return s, keep


class Graph:
    """An instantiated hipGraph of captured entry-point calls."""

    def __init__(self, gpu, handle):
        self.gpu, self.h = gpu, handle
        self.nodes = gpu.L.mom_gpu_graph_nodes(handle)

    def launch(self):
        self.gpu._chk(self.gpu.L.mom_gpu_graph_launch(self.gpu.ctx, self.h), "graph_launch")

    def close(self):
        if self.h:
            self.gpu.L.mom_gpu_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OceanGPU:
    """One MOM domain (global or regional) resident on the GPU.

    ``dom`` holds the Fortran bounds (isd..jed, isc..jec, nk, cyclic_x,
    cyclic_y); ``grid`` the host arrays of ocean_grid_type fields named as in
    MOM (dxt, dyt, dat, datr, dxte, dyte, dxtn, dytn, dxu, dyu, dxur, dyur,
    f, kmt, kmu, tmask, umask, dzt).  The grid is copied to the device.
    Entries may also be :class:`DeviceArray` (read in place), and tmask/umask
    may be left out: the masks are then k <= kmt and k <= kmu.
    """

    GRID_FIELDS = ("dxt", "dyt", "dat", "datr", "dxte", "dyte", "dxtn", "dytn", "dxu", "dyu",
                   "dxur", "dyur", "f", "kmt", "kmu", "tmask", "umask", "dzt")

    def __init__(self, dom: dict, grid: dict):
        self.L = load()
        self.dom = dict(dom)
        self.ni = dom["iec"] - dom["isc"] + 1
        self.nj = dom["jec"] - dom["jsc"] + 1
        self.nk = dom["nk"]
        self.shape2 = (self.nj + 2, self.ni + 2)
        self.shape3 = (self.nk,) + self.shape2
        staged, ptrs = {}, {}
        for name in self.GRID_FIELDS:
            a = grid.get(name)
            if a is None and name in ("tmask", "umask"):
                ptrs[name] = None
            elif isinstance(a, DeviceArray):
                ptrs[name] = a.ptr
            else:
                dt = np.int32 if name in ("kmt", "kmu") else np.float64
                staged[name] = DeviceArray.from_host(np.asarray(a, dtype=dt))
                ptrs[name] = staged[name].ptr
        g = Grid(**ptrs)
        self._dom_c = Domain(*[int(dom[n]) for n, _ in Domain._fields_])
        ctx = C.c_void_p()
        check(self.L.mom_gpu_create(C.byref(self._dom_c), C.byref(g), C.byref(ctx)), None,
              "mom_gpu_create")
        self.ctx = ctx.value
        # the context copied what it needs
        for v in staged.values():
            v.free()
        self.grid_host = grid

    # ---- helpers ---------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.L.mom_gpu_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        check(rc, self.ctx, what)

    def sync(self):
        self._chk(self.L.mom_gpu_sync(self.ctx), "sync")

    def mark(self, slot):
        self._chk(self.L.mom_gpu_mark(self.ctx, slot), "mark")

    def elapsed_ms(self, s0, s1):
        ms = C.c_float()
        self._chk(self.L.mom_gpu_elapsed(self.ctx, s0, s1, C.byref(ms)), "elapsed")
        return float(ms.value)

    def set_tuning(self, knob, value):
        self._chk(self.L.mom_gpu_set_tuning(self.ctx, knob.encode(), int(value)), "set_tuning")

    def get_tuning(self, knob):
        v = C.c_int()
        self._chk(self.L.mom_gpu_get_tuning(self.ctx, knob.encode(), C.byref(v)), "get_tuning")
        return v.value

    def fill(self, arr, value=0.0):
        """Fill a float64 device array on the context's stream."""
        n = arr.nbytes // 8
        self._chk(self.L.mom_gpu_fill(self.ctx, _p(arr), float(value), n), "fill")

    def update_domains(self, arr):
        """Single-PE halo update (mpp_update_domains) of a float64 device
        array of one or more data-domain levels."""
        nz = arr.nbytes // 8 // (self.shape2[0] * self.shape2[1])
        self._chk(self.L.mom_gpu_update_domains(self.ctx, _p(arr), int(nz)), "update_domains")

    def copy(self, dst, src):
        """Copy a float64 device array into another on the context's stream."""
        n = src.nbytes // 8
        assert dst.nbytes >= src.nbytes
        self._chk(self.L.mom_gpu_copy(self.ctx, _p(dst), _p(src), n), "copy")

    def field3(self, value=None, dtype=np.float64):
        a = DeviceArray(self.shape3, dtype)
        if value is None:
            a.zero()
        else:
            a.upload(value)
        return a

    # ---- advection velocity ------------------------------------------------
    def advection_velocity(self, u, rho_dzu, pme, river, uhrho_et, vhrho_nt, wrho_bt,
                           rho_dzt_tendency=None, mass_source=None, diverge_t=None):
        """ocean_advection_velocity (B-grid): T-cell transports and wrho_bt
        from the U-cell velocities u (two components back to back)."""
        self._chk(self.L.mom_gpu_advection_velocity(
            self.ctx, _p(u), _p(rho_dzu), _p(pme), _p(river), _p(rho_dzt_tendency),
            _p(mass_source), _p(uhrho_et), _p(vhrho_nt), _p(wrho_bt), _p(diverge_t)),
            "advection_velocity")

    def barotropic_forcing(self, accel, smf, bmf, forcing_bt):
        """ocean_barotropic_forcing (B-grid, default namelist)."""
        self._chk(self.L.mom_gpu_barotropic_forcing(self.ctx, _p(accel), _p(smf), _p(bmf),
                                                    _p(forcing_bt)), "barotropic_forcing")

    def eta_and_pbot_update(self, x_taum1, x_taup1, tendency, *, conv_rho_ud_t_tau, pme, river,
                            source, dtime, rho0=1035.0, grav=9.80, pressure_based=False,
                            dpatm_dt=None, ht=None, pbot_t_taup1=None):
        """eta_and_pbot_tendency + eta_and_pbot_update: eta_t (or anompb)
        at taup1 and its tendency; pbot_t at taup1 when pressure-based."""
        self._chk(self.L.mom_gpu_eta_and_pbot_update(
            self.ctx, int(bool(pressure_based)), float(rho0), float(grav), float(dtime),
            _p(conv_rho_ud_t_tau), _p(pme), _p(river), _p(source), _p(dpatm_dt), _p(ht),
            _p(x_taum1), _p(x_taup1), _p(pbot_t_taup1), _p(tendency)), "eta_and_pbot_update")

    def set_remap_metrics(self, dte, dtw, dus, dun, daur):
        """Grid widths of REMAP_BT_TO_BU (host arrays on the data domain)."""
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (dte, dtw, dus, dun, daur)]
        for a in keep:
            if a.shape != self.shape2:
                raise ValueError(f"remap metric of shape {a.shape}, expected {self.shape2}")
        self._chk(self.L.mom_gpu_set_remap_metrics(self.ctx, *[a.ctypes.data for a in keep]),
                  "set_remap_metrics")

    def barotropic_closure(self, eta_t_bar, patm, rho_g, eta_u, ps, grad_ps, *, anompb_bar=None,
                           grad_anompb=None, pressure_based=False, ht=None, rho0=1035.0,
                           grav=9.80, tmasken=None):
        """eta_u (pbot_u) closure, ps, grad_ps and grad_anompb of
        update_ocean_barotropic (C-grid gradients when tmasken is given)."""
        self._chk(self.L.mom_gpu_barotropic_closure(
            self.ctx, int(bool(pressure_based)), float(rho0), float(grav), _p(eta_t_bar),
            _p(anompb_bar), _p(patm), _p(rho_g), _p(ht), _p(tmasken), _p(eta_u), _p(ps),
            _p(grad_ps),
            _p(grad_anompb)), "barotropic_closure")

    # ---- tracer advection ------------------------------------------------
    def horz_advect_tracer(self, scheme, T_taum1, th_tendency, *, uhrho_et, vhrho_nt,
                           wrho_bt=None, rho_dzt_tau=None, T_tau=None, dtime=0.0,
                           tmask_limit=None, limit_with_upwind=False, wrk1=None,
                           flux_x=None, flux_y=None, flux_z=None):
        sch = SCHEMES.get(scheme, scheme)
        self._chk(self.L.mom_gpu_horz_advect_tracer(
            self.ctx, int(sch), _p(T_taum1), _p(T_tau), _p(uhrho_et), _p(vhrho_nt), _p(wrho_bt),
            _p(rho_dzt_tau), _p(tmask_limit), int(bool(limit_with_upwind)), float(dtime),
            _p(th_tendency), _p(wrk1), _p(flux_x), _p(flux_y), _p(flux_z)), "horz_advect_tracer")

    def vert_advect_tracer(self, scheme, T_taum1, th_tendency, *, wrho_bt=None, T_tau=None,
                           tmask_limit=None, wrk1=None, flux_z=None):
        sch = SCHEMES.get(scheme, scheme)
        self._chk(self.L.mom_gpu_vert_advect_tracer(
            self.ctx, int(sch), _p(T_taum1), _p(T_tau), _p(wrho_bt), _p(tmask_limit),
            _p(th_tendency), _p(wrk1), _p(flux_z)), "vert_advect_tracer")

    def advect_tracer_sweby_all(self, T_taum1s, th_tendencies, *, uhrho_et, vhrho_nt, wrho_bt,
                                rho_dzt_tau, dtime, wrk1s=None):
        n = len(T_taum1s)
        w = _ptab(wrk1s) if wrk1s is not None else None
        self._chk(self.L.mom_gpu_advect_sweby_all(
            self.ctx, n, _ptab(T_taum1s), _p(uhrho_et), _p(vhrho_nt), _p(wrho_bt),
            _p(rho_dzt_tau), float(dtime), _ptab(th_tendencies), w), "advect_sweby_all")

    def compute_tmask_limit(self, T_taus, min_limits, max_limits, tmask_limits, index_temp=0,
                            index_salt=1, ts_same=True):
        n = len(T_taus)
        mn = (C.c_double * n)(*[float(x) for x in min_limits])
        mx = (C.c_double * n)(*[float(x) for x in max_limits])
        self._chk(self.L.mom_gpu_compute_tmask_limit(
            self.ctx, n, _ptab(T_taus), mn, mx, int(index_temp), int(index_salt), int(ts_same),
            _ptab(tmask_limits)), "compute_tmask_limit")

    def update_tracer_explicit(self, T_taum1, th_tendency, rho_dzt_taum1, rho_dztr, dtime,
                               T_taup1):
        self._chk(self.L.mom_gpu_update_tracer_explicit(
            self.ctx, _p(T_taum1), _p(th_tendency), _p(rho_dzt_taum1), _p(rho_dztr),
            float(dtime), _p(T_taup1)), "update_tracer_explicit")

    # ---- vertical diffusion ------------------------------------------------
    def invtri(self, z, topbc, botbc, dcb, tdt, kmz, mask, dh, dhw, aidif=1.0):
        self._chk(self.L.mom_gpu_invtri(self.ctx, _p(z), _p(topbc), _p(botbc), _p(dcb),
                                        float(tdt), _p(kmz), _p(mask), _p(dh), _p(dhw),
                                        float(aidif)), "invtri")

    def vert_diffuse_implicit(self, T_taup1s, stfs, btfs, diff_cbt, rho_dzt_taup1, dzwt,
                              dtime_t, *, index_salt=1, aidif=1.0, rho0=1035.0, K33s=None):
        n = len(T_taup1s)
        k33 = _ptab(K33s) if K33s is not None else None
        self._chk(self.L.mom_gpu_vert_diffuse_implicit(
            self.ctx, n, _ptab(T_taup1s), _ptab(stfs), _ptab(btfs), _p(diff_cbt), k33,
            int(index_salt), float(rho0), _p(rho_dzt_taup1), _p(dzwt), float(dtime_t),
            float(aidif)), "vert_diffuse_implicit")

    # ---- barotropic ----------------------------------------------------------
    def update_ocean_barotropic(self, *, nts, dtbt, eta_t_bar_tau, udrho_tau, forcing_bt, thicku,
                                patm, pme, river, source, eta_smooth, rho_g, eta_t_bar_taup1,
                                udrho_taup1, conv_rho_ud_t=None, eta_t_tau=None,
                                pred_corr_gamma=0.2, rho0=1035.0, stepping_a=False,
                                initsum_with_bar=True):
        p = BtParams(nts=int(nts), dtbt=float(dtbt), pred_corr_gamma=float(pred_corr_gamma),
                     rho0=float(rho0), stepping_a=int(bool(stepping_a)),
                     initsum_with_bar=int(bool(initsum_with_bar)))
        self._chk(self.L.mom_gpu_update_ocean_barotropic(
            self.ctx, C.byref(p), _p(eta_t_tau), _p(eta_t_bar_tau), _p(udrho_tau),
            _p(forcing_bt), _p(thicku), _p(patm), _p(pme), _p(river), _p(source),
            _p(eta_smooth), _p(rho_g), _p(eta_t_bar_taup1), _p(udrho_taup1),
            _p(conv_rho_ud_t)), "update_ocean_barotropic")

    def update_ocean_barotropic_variant(self, *, cgrid, pressure_based, nts, dtbt, x_bar_tau,
                                        udrho_tau, forcing_bt, thick, pme, river, source, smooth,
                                        x_bar_taup1, udrho_taup1, conv_rho_ud_t=None, x_tau=None,
                                        tmasken=None, patm=None, rho_g=None, dpatm_dt=None,
                                        grav=9.80, pred_corr_gamma=0.2, rho0=1035.0,
                                        stepping_a=False, initsum_with_bar=True):
        """pred_corr_tropic_{depth,press}_{bgrid,cgrid} and the time-mean
        tail of update_ocean_barotropic (x: eta_t or anompb)."""
        p = BtParams(nts=int(nts), dtbt=float(dtbt), pred_corr_gamma=float(pred_corr_gamma),
                     rho0=float(rho0), stepping_a=int(bool(stepping_a)),
                     initsum_with_bar=int(bool(initsum_with_bar)))
        v = BtVariant(cgrid=int(bool(cgrid)), pressure_based=int(bool(pressure_based)),
                      grav=float(grav), x_tau=_p(x_tau), x_bar_tau=_p(x_bar_tau),
                      udrho_tau=_p(udrho_tau), forcing_bt=_p(forcing_bt), thick=_p(thick),
                      tmasken=_p(tmasken), patm=_p(patm), rho_g=_p(rho_g),
                      dpatm_dt=_p(dpatm_dt), pme=_p(pme), river=_p(river), source=_p(source),
                      smooth=_p(smooth), x_bar_taup1=_p(x_bar_taup1),
                      udrho_taup1=_p(udrho_taup1), conv_rho_ud_t=_p(conv_rho_ud_t))
        self._chk(self.L.mom_gpu_update_ocean_barotropic_variant(self.ctx, C.byref(p),
                                                                 C.byref(v)),
                  "update_ocean_barotropic_variant")

    # ---- batched tracer step -------------------------------------------------
    def step_tracers(self, T_taum1s, T_taup1s, *, uhrho_et, vhrho_nt, wrho_bt, rho_dzt_tau,
                     dtime, stfs=None, btfs=None, dzwt=None, diff_cbt=None, index_salt=1,
                     rho0=1035.0, aidif=1.0, th_tendencies=None, rho_dzt_taum1=None,
                     rho_dztr=None, rho_dzt_taup1=None, schemes=None, min_limits=None,
                     max_limits=None, index_temp=0, ts_same=True, limit_with_upwind=False):
        """Advection (MDFL-Sweby, or Quicker where schemes says so), explicit
        update and implicit vertical diffusion of every tracer; each
        tracer's dtype is its array's."""
        s, keep = tracer_step_struct(T_taum1s, T_taup1s, uhrho_et=uhrho_et, vhrho_nt=vhrho_nt,
                                     wrho_bt=wrho_bt, rho_dzt_tau=rho_dzt_tau, dtime=dtime,
                                     stfs=stfs, btfs=btfs, dzwt=dzwt, diff_cbt=diff_cbt,
                                     index_salt=index_salt, rho0=rho0, aidif=aidif,
                                     th_tendencies=th_tendencies, rho_dzt_taum1=rho_dzt_taum1,
                                     rho_dztr=rho_dztr, rho_dzt_taup1=rho_dzt_taup1,
                                     schemes=schemes, min_limits=min_limits,
                                     max_limits=max_limits, index_temp=index_temp,
                                     ts_same=ts_same, limit_with_upwind=limit_with_upwind)
        self._chk(self.L.mom_gpu_step_tracers(self.ctx, C.byref(s)), "step_tracers")

    # ---- hipGraph capture --------------------------------------------------------
    def graph_begin(self):
        self._chk(self.L.mom_gpu_graph_begin(self.ctx), "graph_begin")

    def graph_end(self):
        """End the capture; returns a :class:`Graph` to replay."""
        h = C.c_void_p()
        self._chk(self.L.mom_gpu_graph_end(self.ctx, C.byref(h)), "graph_end")
        return Graph(self, h.value)

    # ---- diagnostics ----------------------------------------------------------
    def tracer_error(self, a, b, rel_floor=0.0):
        """Error of tracer a (FP64 or FP32) against the FP64 result b over
        the wet compute cells: max_abs, max_ref, rms_rel (rms of the error
        over rms of b), normwise (max_abs / max_ref), cells, cells_differ
        (FP64 bit patterns), max_rel_cell (where |b| >= rel_floor)."""
        out = (C.c_double * 7)()
        self._chk(self.L.mom_gpu_tracer_error(self.ctx, _p(a), _elem_size(a), _p(b),
                                              float(rel_floor), out), "tracer_error")
        v = list(out)
        return dict(max_abs=v[0], max_ref=v[1],
                    rms_rel=(v[2] / v[3]) ** 0.5 if v[3] > 0 else 0.0,
                    normwise=v[0] / v[1] if v[1] > 0 else 0.0,
                    cells=int(v[4]), cells_differ=int(v[5]), max_rel_cell=v[6])


    def tracer_content(self, T, rho_dzt):
        """Masked content sum(T*rho_dzt*dat) (compensated FP64 sums) of an
        FP64 or FP32 tracer."""
        out = C.c_double()
        fn = (self.L.mom_gpu_tracer_content if _elem_size(T) == 8
              else self.L.mom_gpu_tracer_content_f32)
        self._chk(fn(self.ctx, _p(T), _p(rho_dzt), C.byref(out)), "tracer_content")
        return out.value


__all__ = ["DeviceArray", "OceanGPU", "MomGpuError", "ADVECT_QUICKER", "ADVECT_MDFL_SWEBY",
           "ADVECT_DST_LINEAR", "SCHEMES"]
