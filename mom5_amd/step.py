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
"""
from __future__ import annotations

import numpy as np

from .ocean import DeviceArray, OceanGPU

# default tracer range limits (ocean_tempsalt.F90:270-273, ocean_passive.F90:239-240)
LIMITS = {"temp": (-2.0, 32.0), "salt": (1.0, 42.0)}
PASSIVE_LIMITS = (-0.1, 1.0)


class PlainStep:
    """Per-tracer FP64 path: one C-ABI call per reference routine."""

    def __init__(self, case, schemes=None, gpu: OceanGPU | None = None):
        self.case = case
        self.gpu = gpu or OceanGPU(case.dom, case.grid)
        g = self.gpu
        f = case.fields
        self.dev = {k: DeviceArray.from_host(f[k]) for k in
                    ("uhrho_et", "vhrho_nt", "wrho_bt", "rho_dzt", "rho_dztr", "dzwt",
                     "diff_cbt", "eta_t", "eta_t_bar", "udrho", "forcing_bt", "thicku", "patm",
                     "pme", "river", "source", "eta_smooth", "rho_g")}
        self.names = [n for n, _ in case.tracers]
        self.ntr = len(self.names)
        if schemes is None:
            schemes = ["mdfl_sweby" if n < 2 else "quicker" for n in range(self.ntr)]
        self.schemes = list(schemes)
        self.T = [[DeviceArray.from_host(np.asarray(t, np.float64)), g.field3()]
                  for _, t in case.tracers]
        self.th = [g.field3() for _ in range(self.ntr)]
        self.tlim = [g.field3() for _ in range(self.ntr)]
        self.stf = [DeviceArray.from_host(s) for s in f["stf"]]
        self.btf = [DeviceArray.from_host(s) for s in f["btf"]]
        self.mins, self.maxs = [], []
        for n in self.names:
            lo, hi = LIMITS.get(n, PASSIVE_LIMITS)
            self.mins.append(lo)
            self.maxs.append(hi)
        sh = g.shape2
        self.eta_bar = [self.dev["eta_t_bar"], DeviceArray(sh)]
        self.udrho = [self.dev["udrho"], DeviceArray((2,) + sh)]
        self.conv = DeviceArray(sh)
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
                th = self.th[n]
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
            if c.dtbt > 0.0:
                g.update_ocean_barotropic(
                    nts=c.nts, dtbt=c.dtbt, eta_t_bar_tau=self.eta_bar[now],
                    udrho_tau=self.udrho[now], forcing_bt=d["forcing_bt"], thicku=d["thicku"],
                    patm=d["patm"], pme=d["pme"], river=d["river"], source=d["source"],
                    eta_smooth=d["eta_smooth"], rho_g=d["rho_g"],
                    eta_t_bar_taup1=self.eta_bar[nxt], udrho_taup1=self.udrho[nxt],
                    conv_rho_ud_t=self.conv)
            self.cur = nxt
