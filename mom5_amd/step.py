"""One baroclinic step of the hot path, in the reference's calling order.

``ocean_model.F90:update_ocean_model`` (:1441-2108) runs, for the part this
package implements:

1. ``compute_tmask_limit`` (:1567, ``ocean_tracer.F90:2788``);
2. ``update_ocean_tracer`` (:1869): per tracer ``horz_advect_tracer`` and
   ``vert_advect_tracer`` into ``th_tendency``, the explicit update to
   ``taup1`` (``ocean_tracer.F90:2295-2350``), then ``vert_diffuse_implicit``
   for all tracers (:2432);
3. ``update_ocean_barotropic`` (:1941).

Two-level time stepping (``taum1 = tau``) is assumed, as in the default
configuration (``ocean_model.F90:1489-1491``): each step reads one tracer
time level and writes the other, then the levels swap.

:class:`PlainStep` makes one C-ABI call per reference routine and tracer,
and updates the tracer halos after the step as ``ocean_model.F90:2049``
does; :class:`BatchedStep` advances every tracer with one ``step_tracers``
call (grouped, fused advection + explicit update + implicit vertical
diffusion, FP64 or FP32 per tracer).  The batched step forms the
``tmask_limit`` of its Quicker tracers inside their pass (its only readers
are the Quicker and QuickMOM3 schemes, ``ocean_tracer_advect.F90:2609-2625,
2738-2753, 3006-3111``) and reads every halo value as the halo update would
leave it, so it needs no halo updates of its own.
"""
from __future__ import annotations

import numpy as np

from .ocean import DeviceArray, OceanGPU


def _resident(x):
    """A device copy of a host array; device arrays are used in place."""
    return x if isinstance(x, DeviceArray) else DeviceArray.from_host(x)


def _copy(x):
    """A private device copy (the step writes its time levels in place)."""
    return x.copy() if isinstance(x, DeviceArray) else DeviceArray.from_host(x)


def _zeros_like(x):
    a = DeviceArray(x.shape, x.dtype)
    a.zero()
    return a


# default tracer range limits (ocean_tempsalt.F90:270-273, ocean_passive.F90:239-240)
LIMITS = {"temp": (-2.0, 32.0), "salt": (1.0, 42.0)}
PASSIVE_LIMITS = (-0.1, 1.0)


class _Captured:
    """hipGraph capture of whole steps (shared by both step drivers)."""

    def capture(self, nsteps=2):
        """Capture `nsteps` steps (even: the time levels come back to where
        they were) into a graph; a replay advances the state by nsteps."""
        if nsteps % 2:
            raise ValueError("capture an even number of steps (two-level time stepping)")
        g = self.gpu
        g.sync()
        g.graph_begin()
        try:
            self.run(nsteps)
        except Exception:
            try:
                g.graph_end()
            except Exception:
                pass
            raise
        graph = g.graph_end()
        graph.steps = nsteps
        return graph


class PlainStep(_Captured):
    """Per-tracer FP64 path: one C-ABI call per reference routine."""

    def __init__(self, case, schemes=None, gpu: OceanGPU | None = None):
        self.case = case
        self.gpu = gpu or OceanGPU(case.dom, case.grid)
        g = self.gpu
        f = case.fields
        self.dev = {k: _resident(f[k]) for k in
                    ("uhrho_et", "vhrho_nt", "wrho_bt", "rho_dzt", "rho_dztr", "dzwt",
                     "diff_cbt", "eta_t_bar", "udrho", "forcing_bt", "thicku", "patm",
                     "pme", "river", "source", "eta_smooth", "rho_g")}
        self.names = [n for n, _ in case.tracers]
        self.ntr = len(self.names)
        if schemes is None:
            schemes = ["mdfl_sweby" if n < 2 else "quicker" for n in range(self.ntr)]
        self.schemes = list(schemes)
        self.T = []
        for _, t in case.tracers:
            if not isinstance(t, DeviceArray):
                t = DeviceArray.from_host(np.asarray(t, np.float64))
            if t.element_size() != 8:
                raise ValueError("the plain path advances FP64 tracers only")
            self.T.append([t, g.field3()])
        # th_tendency is consumed by the explicit update right after the
        # advection of its tracer, so one array serves every tracer
        self.th = g.field3()
        self.tlim = [g.field3() for _ in range(self.ntr)]
        self.stf = [_resident(s) for s in f["stf"][:self.ntr]]
        self.btf = [_resident(s) for s in f["btf"][:self.ntr]]
        self.mins, self.maxs = [], []
        for n in self.names:
            lo, hi = LIMITS.get(n, PASSIVE_LIMITS)
            self.mins.append(lo)
            self.maxs.append(hi)
        self.eta_bar = [_copy(f["eta_t_bar"]), _zeros_like(self.dev["eta_t_bar"])]
        self.udrho = [_copy(f["udrho"]), _zeros_like(self.dev["udrho"])]
        self.conv = _zeros_like(self.dev["eta_t_bar"])
        self.cur = 0

    def tracer(self, n):
        """Device array of tracer n at the current time level."""
        return self.T[n][self.cur]

    def run(self, nsteps=1):
        g, d, c = self.gpu, self.dev, self.case
        for _ in range(nsteps):
            now, nxt = self.cur, 1 - self.cur
            Tn = [self.T[n][now] for n in range(self.ntr)]
            Tp = [self.T[n][nxt] for n in range(self.ntr)]
            g.compute_tmask_limit(Tn, self.mins, self.maxs, self.tlim,
                                  index_temp=self.names.index("temp") if "temp" in self.names else -1,
                                  index_salt=self.names.index("salt") if "salt" in self.names else -1)
            for n in range(self.ntr):
                th = self.th
                g.fill(th, 0.0)
                g.horz_advect_tracer(self.schemes[n], Tn[n], th, uhrho_et=d["uhrho_et"],
                                     vhrho_nt=d["vhrho_nt"], wrho_bt=d["wrho_bt"],
                                     rho_dzt_tau=d["rho_dzt"], T_tau=Tn[n], dtime=c.dt,
                                     tmask_limit=self.tlim[n])
                g.vert_advect_tracer(self.schemes[n], Tn[n], th, wrho_bt=d["wrho_bt"],
                                     T_tau=Tn[n], tmask_limit=self.tlim[n])
                g.update_tracer_explicit(Tn[n], th, d["rho_dzt"], d["rho_dztr"], c.dt, Tp[n])
            g.vert_diffuse_implicit(Tp, self.stf, self.btf, d["diff_cbt"], d["rho_dzt"],
                                    d["dzwt"], c.dt,
                                    index_salt=self.names.index("salt") if "salt" in self.names else -1)
            for t in Tp:                     # ocean_model.F90:2049
                g.update_domains(t)
            if c.dtbt > 0.0:
                g.update_ocean_barotropic(
                    nts=c.nts, dtbt=c.dtbt, eta_t_bar_tau=self.eta_bar[now],
                    udrho_tau=self.udrho[now], forcing_bt=d["forcing_bt"], thicku=d["thicku"],
                    patm=d["patm"], pme=d["pme"], river=d["river"], source=d["source"],
                    eta_smooth=d["eta_smooth"], rho_g=d["rho_g"],
                    eta_t_bar_taup1=self.eta_bar[nxt], udrho_taup1=self.udrho[nxt],
                    conv_rho_ud_t=self.conv)
            self.cur = nxt


class BatchedStep(_Captured):
    """All tracers MDFL-Sweby through ``mom_gpu_step_tracers``, then the
    barotropic subcycle.  Each tracer keeps the dtype of its initial array
    (FP64, or FP32 for passive tracers); fields may be host arrays (copied
    to the device) or device arrays (used in place: a tracer's array is its
    time level 0 and is overwritten on odd steps)."""

    FIELDS = ("uhrho_et", "vhrho_nt", "wrho_bt", "rho_dzt", "rho_dztr", "dzwt", "diff_cbt",
              "eta_t_bar", "udrho", "forcing_bt", "thicku", "patm", "pme", "river", "source",
              "eta_smooth", "rho_g")

    def __init__(self, case, gpu: OceanGPU | None = None, tracers=None, diffuse=True,
                 barotropic=True, schemes=None):
        self.case = case
        self.gpu = gpu or OceanGPU(case.dom, case.grid)
        f = case.fields
        self.dev = {k: _resident(f[k]) for k in self.FIELDS}
        trs = case.tracers if tracers is None else tracers
        self.names = [n for n, _ in trs]
        self.ntr = len(self.names)
        # per-tracer scheme ("mdfl_sweby" or "quicker") and tmask_limit ranges
        self.schemes = None if schemes is None or all(x == "mdfl_sweby" for x in schemes) \
            else list(schemes)
        self.mins = [LIMITS.get(n, PASSIVE_LIMITS)[0] for n in self.names]
        self.maxs = [LIMITS.get(n, PASSIVE_LIMITS)[1] for n in self.names]
        self.T = [[_resident(t), None] for _, t in trs]
        for pair in self.T:
            pair[1] = _zeros_like(pair[0])
        self.diffuse = diffuse
        self.stf = [_resident(x) for x in f["stf"][:self.ntr]] if diffuse else None
        self.btf = [_resident(x) for x in f["btf"][:self.ntr]] if diffuse else None
        self.index_salt = self.names.index("salt") if "salt" in self.names else -1
        self.barotropic = barotropic and case.dtbt > 0.0
        if self.barotropic:
            self.eta_bar = [_copy(f["eta_t_bar"]), _zeros_like(self.dev["eta_t_bar"])]
            self.udrho = [_copy(f["udrho"]), _zeros_like(self.dev["udrho"])]
            self.conv = _zeros_like(self.dev["eta_t_bar"])
        # other th_tendency contributions per tracer (device arrays), or None
        self.th = None
        self.cur = 0

    def tracer(self, n):
        """Device array of tracer n at the current time level."""
        return self.T[n][self.cur]

    def step_args(self):
        """Keyword arguments of this step's tracer call at the current level."""
        d, c = self.dev, self.case
        now, nxt = self.cur, 1 - self.cur
        return dict(T_taum1s=[p[now] for p in self.T], T_taup1s=[p[nxt] for p in self.T],
                    uhrho_et=d["uhrho_et"], vhrho_nt=d["vhrho_nt"], wrho_bt=d["wrho_bt"],
                    rho_dzt_tau=d["rho_dzt"], rho_dztr=d["rho_dztr"], dtime=c.dt,
                    stfs=self.stf, btfs=self.btf, dzwt=d["dzwt"] if self.diffuse else None,
                    diff_cbt=d["diff_cbt"] if self.diffuse else None,
                    index_salt=self.index_salt, th_tendencies=self.th, schemes=self.schemes,
                    min_limits=self.mins, max_limits=self.maxs,
                    index_temp=self.names.index("temp") if "temp" in self.names else -1)

    def bt_args(self):
        """Barotropic parameters and fields of this step at the current level."""
        d, c = self.dev, self.case
        now, nxt = self.cur, 1 - self.cur
        return dict(nts=c.nts, dtbt=c.dtbt, eta_t_bar_tau=self.eta_bar[now],
                    udrho_tau=self.udrho[now], forcing_bt=d["forcing_bt"], thicku=d["thicku"],
                    patm=d["patm"], pme=d["pme"], river=d["river"], source=d["source"],
                    eta_smooth=d["eta_smooth"], rho_g=d["rho_g"],
                    eta_t_bar_taup1=self.eta_bar[nxt], udrho_taup1=self.udrho[nxt],
                    conv_rho_ud_t=self.conv)

    def run(self, nsteps=1):
        g, d, c = self.gpu, self.dev, self.case
        for _ in range(nsteps):
            now, nxt = self.cur, 1 - self.cur
            a = self.step_args()
            g.step_tracers(a.pop("T_taum1s"), a.pop("T_taup1s"), **a)
            if self.barotropic:
                g.update_ocean_barotropic(
                    nts=c.nts, dtbt=c.dtbt, eta_t_bar_tau=self.eta_bar[now],
                    udrho_tau=self.udrho[now], forcing_bt=d["forcing_bt"], thicku=d["thicku"],
                    patm=d["patm"], pme=d["pme"], river=d["river"], source=d["source"],
                    eta_smooth=d["eta_smooth"], rho_g=d["rho_g"],
                    eta_t_bar_taup1=self.eta_bar[nxt], udrho_taup1=self.udrho[nxt],
                    conv_rho_ud_t=self.conv)
            self.cur = nxt
