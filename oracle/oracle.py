"""ctypes front end of the CPU restatement (oracle/mom5_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product package
(mom5_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmom5_oracle.so")
# the same source built with -fopenmp: loops over independent levels, rows or
# points run on every host core (OMP_NUM_THREADS), results identical
LIB_OMP_PATH = os.path.join(HERE, "_build", "libmom5_oracle_omp.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Dom(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("isd", "ied", "jsd", "jed", "isc", "iec",
                                        "jsc", "jec", "nk", "cyclic_x", "cyclic_y")]


class Grid(C.Structure):
    _fields_ = [(n, _dp) for n in ("dxt", "dyt", "dat", "datr", "dxte", "dyte",
                                    "dxtn", "dytn", "dxu", "dyu", "dxur", "dyur", "f")] + \
               [("kmt", _ip), ("tmask", _dp), ("umask", _dp), ("dzt", _dp)]


class BtParams(C.Structure):
    _fields_ = [("nts", C.c_int), ("dtbt", C.c_double), ("pred_corr_gamma", C.c_double),
                ("rho0", C.c_double), ("stepping_a", C.c_int), ("initsum_with_bar", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def lib(threads=False):
    """The serial restatement, or with threads=True its all-core build."""
    path = LIB_OMP_PATH if threads else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        _libs[path] = C.CDLL(path)
    return _libs[path]


def _p(a):
    """Pointer to a float64 C-contiguous array (kept alive by the caller)."""
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _parr(arrs):
    arr = (_dp * len(arrs))()
    for n, a in enumerate(arrs):
        arr[n] = _p(a) if a is not None else None
    return arr


class Oracle:
    """Binds one synthetic case (mom5_amd.synth.Case) to the oracle."""

    def __init__(self, case, threads=False):
        self.case = case
        self.dom = Dom(**case.dom)
        g = case.grid
        self._keep = {k: np.ascontiguousarray(v) for k, v in g.items()}
        self.grid = Grid(**{k: _p(self._keep[k]) for k in
                            ("dxt", "dyt", "dat", "datr", "dxte", "dyte", "dxtn", "dytn",
                             "dxu", "dyu", "dxur", "dyur", "f", "tmask", "umask", "dzt")},
                         kmt=_pi(self._keep["kmt"]))
        self.L = lib(threads)
        nk, NJ, NI = case.nk, case.nj + 2, case.ni + 2
        self.shape3 = (nk, NJ, NI)
        self.shape2 = (NJ, NI)
        self.shapec = (nk, case.nj, case.ni)

    # --- advection -------------------------------------------------
    def advect_mdfl_sweby(self, T, dtime, lim=1.0, want_flux=False, fields=None):
        f = fields or self.case.fields
        adv = np.zeros(self.shapec)
        fl = [np.zeros(self.shape3) for _ in range(3)] if want_flux else [None] * 3
        self.L.orc_advect_tracer_mdfl_sweby(
            C.byref(self.dom), C.byref(self.grid), _p(T), _p(f["uhrho_et"]), _p(f["vhrho_nt"]),
            _p(f["wrho_bt"]), _p(f["rho_dzt"]), C.c_double(dtime), C.c_double(lim), _p(adv),
            *[(_p(a) if a is not None else None) for a in fl])
        return (adv, fl) if want_flux else adv

    def horz_advect_mdfl(self, T, th_tendency, dtime, lim=1.0, fields=None):
        f = fields or self.case.fields
        wrk1 = np.zeros(self.shape3)
        self.L.orc_horz_advect_tracer_mdfl(
            C.byref(self.dom), C.byref(self.grid), _p(T), _p(f["uhrho_et"]), _p(f["vhrho_nt"]),
            _p(f["wrho_bt"]), _p(f["rho_dzt"]), C.c_double(dtime), C.c_double(lim),
            _p(th_tendency), _p(wrk1))
        return wrk1

    def sweby_all(self, Ts, th_tendencies, dtime, fields=None):
        f = fields or self.case.fields
        wrk1 = [np.zeros(self.shape3) for _ in Ts]
        self.L.orc_advect_tracer_sweby_all(
            C.byref(self.dom), C.byref(self.grid), C.c_int(len(Ts)), _parr(Ts),
            _p(f["uhrho_et"]), _p(f["vhrho_nt"]), _p(f["wrho_bt"]), _p(f["rho_dzt"]),
            C.c_double(dtime), _parr(th_tendencies), _parr(wrk1))
        return wrk1

    def quicker_init(self):
        P2 = self.shape2[0] * self.shape2[1]
        nk = self.case.nk
        out = dict(quick_x=np.zeros(2 * P2), quick_y=np.zeros(2 * P2),
                   curv_xp=np.zeros(3 * P2), curv_xn=np.zeros(3 * P2),
                   curv_yp=np.zeros(3 * P2), curv_yn=np.zeros(3 * P2),
                   quick_z=np.zeros(2 * nk), curv_zp=np.zeros(3 * nk), curv_zn=np.zeros(3 * nk))
        self.L.orc_quicker_init(C.byref(self.dom), C.byref(self.grid),
                                *[_p(out[k]) for k in ("quick_x", "quick_y", "curv_xp", "curv_xn",
                                                       "curv_yp", "curv_yn", "quick_z", "curv_zp",
                                                       "curv_zn")])
        return out

    def horz_quicker(self, T_taum1, T_tau, tmask_limit=None, limit_with_upwind=False,
                     want_flux=False, fields=None):
        f = fields or self.case.fields
        res = np.zeros(self.shapec)
        tl = tmask_limit if tmask_limit is not None else np.zeros(self.shape3)
        fl = [np.zeros(self.shape3) for _ in range(2)] if want_flux else [None] * 2
        self.L.orc_horz_advect_tracer_quicker(
            C.byref(self.dom), C.byref(self.grid), _p(T_taum1), _p(T_tau), _p(f["uhrho_et"]),
            _p(f["vhrho_nt"]), _p(tl), C.c_int(int(limit_with_upwind)), _p(res),
            *[(_p(a) if a is not None else None) for a in fl])
        return (res, fl) if want_flux else res

    def vert_quicker(self, T_taum1, T_tau, tmask_limit=None, want_flux=False, fields=None):
        f = fields or self.case.fields
        res = np.zeros(self.shapec)
        tl = tmask_limit if tmask_limit is not None else np.zeros(self.shape3)
        fz = np.zeros(self.shape3) if want_flux else None
        self.L.orc_vert_advect_tracer_quicker(
            C.byref(self.dom), C.byref(self.grid), _p(T_taum1), _p(T_tau), _p(f["wrho_bt"]),
            _p(tl), _p(res), _p(fz) if fz is not None else None)
        return (res, fz) if want_flux else res

    def tmask_limit(self, Ts, mins, maxs, index_temp=0, index_salt=1, ts_same=True):
        outs = [np.zeros(self.shape3) for _ in Ts]
        mn = np.ascontiguousarray(mins, dtype=np.float64)
        mx = np.ascontiguousarray(maxs, dtype=np.float64)
        self.L.orc_compute_tmask_limit(C.byref(self.dom), C.byref(self.grid), C.c_int(len(Ts)),
                                       _parr(Ts), _p(mn), _p(mx), C.c_int(index_temp),
                                       C.c_int(index_salt), C.c_int(int(ts_same)), _parr(outs))
        return outs

    def update_explicit(self, T_taum1, th_tendency, dtime, T_taup1=None, fields=None):
        f = fields or self.case.fields
        out = np.zeros(self.shape3) if T_taup1 is None else T_taup1
        self.L.orc_update_tracer_explicit(C.byref(self.dom), _p(T_taum1), _p(th_tendency),
                                          _p(f["rho_dzt"]), _p(f["rho_dztr"]), C.c_double(dtime),
                                          _p(out))
        return out

    # --- vertical diffusion -----------------------------------------
    def invtri(self, z, topbc, botbc, dcb, tdt, dh, dhw, aidif=1.0, kmz=None, mask=None):
        g = self._keep
        kmz = g["kmt"] if kmz is None else np.ascontiguousarray(kmz, dtype=np.int32)
        mask = g["tmask"] if mask is None else mask
        self.L.orc_invtri(C.byref(self.dom), _p(z), _p(topbc), _p(botbc), _p(dcb), C.c_double(tdt),
                          _pi(kmz), _p(mask), _p(dh), _p(dhw), C.c_double(aidif))
        return z

    def vert_diffuse_implicit(self, Tp1s, stfs, btfs, dtime_t, index_salt=1, aidif=1.0,
                              K33=None, rho0=1035.0, fields=None):
        if self.case.dom["nk"] < 2:
            # invtri (ocean_util.F90:175-236) sets the bottom row at level
            # max(2, kmt): a column of one level has no such row
            raise ValueError("vert_diffuse_implicit: nk >= 2 required")
        f = fields or self.case.fields
        k33 = _parr(K33) if K33 is not None else None
        self.L.orc_vert_diffuse_implicit(
            C.byref(self.dom), C.byref(self.grid), C.c_int(len(Tp1s)), _parr(Tp1s), _parr(stfs),
            _parr(btfs), _p(f["diff_cbt"]), k33, C.c_int(index_salt), C.c_double(rho0),
            _p(f["rho_dzt"]), _p(f["dzwt"]), C.c_double(dtime_t), C.c_double(aidif))
        return Tp1s

    # --- barotropic ---------------------------------------------------
    def update_ocean_barotropic(self, nts, dtbt, gamma=0.2, stepping_a=False,
                                initsum_with_bar=True, rho0=1035.0, fields=None,
                                want_conv=True):
        f = fields or self.case.fields
        p = BtParams(nts=nts, dtbt=dtbt, pred_corr_gamma=gamma, rho0=rho0,
                     stepping_a=int(stepping_a), initsum_with_bar=int(initsum_with_bar))
        eta_bar = np.zeros(self.shape2)
        ud = np.zeros((2,) + self.shape2)
        conv = np.zeros(self.shape2) if want_conv else None
        self.L.orc_update_ocean_barotropic(
            C.byref(self.dom), C.byref(self.grid), C.byref(p), _p(f["eta_t"]), _p(f["eta_t_bar"]),
            _p(f["udrho"]), _p(f["forcing_bt"]), _p(f["thicku"]), _p(f["patm"]), _p(f["pme"]),
            _p(f["river"]), _p(f["source"]), _p(f["eta_smooth"]), _p(f["rho_g"]),
            _p(eta_bar), _p(ud), _p(conv) if conv is not None else None)
        return eta_bar, ud, conv

    def update_ocean_barotropic_variant(self, nts, dtbt, *, cgrid, press, fields, gamma=0.2,
                                        stepping_a=False, initsum_with_bar=True, rho0=1035.0,
                                        grav=9.80, want_conv=True):
        """One of the four split-explicit schemes + tail.  fields: x_tau,
        xbar_tau, udrho, forcing_bt, thick, tmasken (C-grid), patm, rho_g
        (depth), dpatm_dt (press), pme, river, source, smooth.  Returns
        (xbar, udrho, conv)."""
        p = BtParams(nts=nts, dtbt=dtbt, pred_corr_gamma=gamma, rho0=rho0,
                     stepping_a=int(stepping_a), initsum_with_bar=int(initsum_with_bar))
        f = {k: (np.ascontiguousarray(v, dtype=np.float64) if v is not None else None)
             for k, v in fields.items()}
        g = lambda k: _p(f[k]) if f.get(k) is not None else None
        xbar = np.zeros(self.shape2)
        ud = np.zeros((2,) + self.shape2)
        conv = np.zeros(self.shape2) if want_conv else None
        self.L.orc_update_ocean_barotropic_variant(
            C.byref(self.dom), C.byref(self.grid), C.byref(p), C.c_int(int(cgrid)),
            C.c_int(int(press)), C.c_double(grav), g("x_tau"), g("xbar_tau"), g("udrho"),
            g("forcing_bt"), g("thick"), g("tmasken"), g("patm"), g("rho_g"), g("dpatm_dt"),
            g("pme"), g("river"), g("source"), g("smooth"), _p(xbar), _p(ud),
            _p(conv) if conv is not None else None)
        return xbar, ud, conv

    def tracer_content(self, T, rho_dzt=None):
        rd = self.case.fields["rho_dzt"] if rho_dzt is None else rho_dzt
        self.L.orc_tracer_content.restype = C.c_double
        return self.L.orc_tracer_content(C.byref(self.dom), C.byref(self.grid), _p(T), _p(rd))

    # --- advection velocity ------------------------------------------
    def advection_velocity(self, u, rho_dzu, pme=None, river=None, rho_dzt_tendency=None,
                           mass_source=None):
        """ocean_advection_velocity.F90:568-672 (B-grid): u is (2, nk, NJ, NI).
        Returns uhrho_et, vhrho_nt, wrho_bt (nk+1 levels), diverge_t."""
        z2 = np.zeros(self.shape2)
        u = np.ascontiguousarray(u, dtype=np.float64)
        rho_dzu = np.ascontiguousarray(rho_dzu, dtype=np.float64)
        pme = z2 if pme is None else np.ascontiguousarray(pme)
        river = z2 if river is None else np.ascontiguousarray(river)
        ue, vn, div = (np.zeros(self.shape3) for _ in range(3))
        wb = np.zeros((self.case.nk + 1,) + self.shape2)
        self.L.orc_advection_velocity(
            C.byref(self.dom), C.byref(self.grid), _p(u), _p(rho_dzu), _p(pme), _p(river),
            _p(rho_dzt_tendency) if rho_dzt_tendency is not None else None,
            _p(mass_source) if mass_source is not None else None, _p(ue), _p(vn), _p(wb),
            _p(div))
        return ue, vn, wb, div

    # --- barotropic forcing ------------------------------------------
    def barotropic_forcing(self, accel, smf, bmf):
        """ocean_barotropic_forcing (default namelist, B-grid): accel is
        (2, nk, NJ, NI), smf and bmf (2, NJ, NI); returns forcing_bt."""
        accel = np.ascontiguousarray(accel, dtype=np.float64)
        smf = np.ascontiguousarray(smf, dtype=np.float64)
        bmf = np.ascontiguousarray(bmf, dtype=np.float64)
        out = np.zeros((2,) + self.shape2)
        self.L.orc_barotropic_forcing(C.byref(self.dom), C.byref(self.grid), _p(accel), _p(smf),
                                      _p(bmf), _p(out))
        return out

    def eta_and_pbot_update(self, x_taum1, conv_tau, pme, river, source, dtime, *,
                            pressure_based=False, dpatm_dt=None, ht=None, x_taup1=None,
                            rho0=1035.0, grav=9.80):
        """eta_and_pbot_tendency + eta_and_pbot_update: returns (x_taup1,
        tendency, pbot_taup1 or None).  x_taup1 (optional) is the array's
        previous contents (closed halos keep them)."""
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        ins = [f64(a) for a in (conv_tau, pme, river, source, x_taum1)]
        out = f64(np.zeros(self.shape2) if x_taup1 is None else np.array(x_taup1, np.float64))
        tend = np.zeros(self.shape2)
        if pressure_based:
            dpatm_dt, ht = f64(dpatm_dt), f64(ht)
            pbot = np.zeros(self.shape2)
        self.L.orc_eta_and_pbot_update(
            C.byref(self.dom), C.byref(self.grid), C.c_int(int(pressure_based)),
            C.c_double(rho0), C.c_double(grav), C.c_double(dtime), *[_p(a) for a in ins[:4]],
            _p(dpatm_dt) if pressure_based else None, _p(ht) if pressure_based else None,
            _p(ins[4]), _p(out), _p(pbot) if pressure_based else None, _p(tend))
        return out, tend, (pbot if pressure_based else None)

    def barotropic_closure(self, eta_t_bar, patm, rho_g, ucell, *, anompb_bar=None,
                           pressure_based=False, ht=None, eta_u=None, rho0=1035.0, grav=9.80,
                           tmasken=None):
        """eta_u (pbot_u) closure, ps, grad_ps and grad_anompb.  ucell: dict
        of dte, dtw, dus, dun, daur.  Returns (eta_u, ps, grad_ps,
        grad_anompb or None)."""
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        eta_t_bar, patm, rho_g = f64(eta_t_bar), f64(patm), f64(rho_g)
        uc = [f64(ucell[k]) for k in ("dte", "dtw", "dus", "dun", "daur")]
        ab = f64(anompb_bar) if anompb_bar is not None else None
        htp = f64(ht) if ht is not None else None
        tm = f64(tmasken) if tmasken is not None else None
        eu = f64(np.zeros(self.shape2) if eta_u is None else np.array(eta_u, np.float64))
        ps = np.zeros(self.shape2)
        gps = np.zeros((2,) + self.shape2)
        gab = np.zeros((2,) + self.shape2) if ab is not None else None
        self.L.orc_barotropic_closure(
            C.byref(self.dom), C.byref(self.grid), C.c_int(int(pressure_based)),
            C.c_double(rho0), C.c_double(grav), _p(eta_t_bar), _p(ab) if ab is not None else None,
            _p(patm), _p(rho_g), _p(htp) if htp is not None else None,
            _p(tm) if tm is not None else None, *[_p(a) for a in uc],
            _p(eu), _p(ps), _p(gps), _p(gab) if gab is not None else None)
        return eu, ps, gps, gab
