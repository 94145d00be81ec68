"""Ensembles of independent domains advanced together (mom_gpu_ensemble_*).

Each member is a :class:`~mom5_amd.step.BatchedStep` on its own context (its
own grid, bathymetry, time step and barotropic subcycle count).  One
ensemble step makes one call for the tracers of every member and one for
their barotropic subcycles; inside, each phase is a single launch over all
members (a device-side work queue of (domain, tile) items for the tracer
passes), so a few small domains fill the GPU together.  Results equal, bit
for bit, each member stepped on its own.

:func:`make_regional_ensemble` draws the uneven regional domains of the
ensemble workload from a seed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BtFields, BtParams, TracerStep, check, load
from .ocean import _p, tracer_step_struct
from .step import BatchedStep, _Captured


class EnsembleStep(_Captured):
    """Advance several BatchedSteps as one ensemble.  Launches and graph
    capture use the first member's context (its stream)."""

    def __init__(self, members):
        self.members = list(members)
        self.gpu = self.members[0].gpu
        self.L = load()
        n = len(self.members)
        ctxs = (C.c_void_p * n)(*[m.gpu.ctx for m in self.members])
        h = C.c_void_p()
        check(self.L.mom_gpu_ensemble_create(n, ctxs, C.byref(h)), self.gpu.ctx,
              "mom_gpu_ensemble_create")
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            self.L.mom_gpu_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps=1):
        n = len(self.members)
        for _ in range(nsteps):
            structs = (TracerStep * n)()
            keep = []
            for k, m in enumerate(self.members):
                a = m.step_args()
                s, kp = tracer_step_struct(a.pop("T_taum1s"), a.pop("T_taup1s"), **a)
                structs[k] = s
                keep.append(kp)
            check(self.L.mom_gpu_ensemble_step_tracers(self.h, structs), self.gpu.ctx,
                  "ensemble_step_tracers")
            bt = [m for m in self.members if m.barotropic]
            if bt:
                if len(bt) != n:
                    raise ValueError("either every member or none has a barotropic part")
                params = (BtParams * n)()
                fields = (BtFields * n)()
                for k, m in enumerate(self.members):
                    a = m.bt_args()
                    params[k] = BtParams(nts=int(a["nts"]), dtbt=float(a["dtbt"]),
                                         pred_corr_gamma=0.2, rho0=1035.0, stepping_a=0,
                                         initsum_with_bar=1)
                    fields[k] = BtFields(**{f: _p(a.get(f)) for f, _ in BtFields._fields_})
                check(self.L.mom_gpu_ensemble_update_barotropic(self.h, params, fields),
                      self.gpu.ctx, "ensemble_update_barotropic")
            for m in self.members:
                m.cur = 1 - m.cur

    def cells(self):
        return sum(m.case.ni * m.case.nj * m.case.nk for m in self.members)

    def tracer_cells(self):
        return sum(m.ntr * m.case.ni * m.case.nj * m.case.nk for m in self.members)


def regional_shapes(ndom=24, seed=0x4D4F4D35 + 4, nx=(200, 800), ny=(150, 600), nz=(30, 50),
                    nts=(40, 120)):
    """Seeded (ni, nj, nk, nts, lat_s, lon_span, seed) of each member."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(ndom):
        ni = int(rng.integers(nx[0], nx[1] + 1))
        nj = int(rng.integers(ny[0], ny[1] + 1))
        nk = int(rng.integers(nz[0], nz[1] + 1))
        ns = int(rng.integers(nts[0], nts[1] + 1))
        ns += ns % 2                      # even subcycle counts: dt = dtbt*nts/2
        lat_s = float(rng.uniform(-75.0, 75.0 - 0.1 * nj))   # stay clear of the poles
        out.append(dict(ni=ni, nj=nj, nk=nk, nts=ns, lat_s=lat_s,
                        lat_n=lat_s + 0.1 * nj, lon_span=0.1 * ni,
                        seed=int(rng.integers(1, 2 ** 31))))
    return out


def make_regional_ensemble(shapes, n_passive=6, passive_dtype=np.float32, land_frac=0.25):
    """One device-built closed regional case per shape (mom5_amd.synth)."""
    from .synth import make_case_device
    cases = []
    for s in shapes:
        cases.append(make_case_device(s["ni"], s["nj"], s["nk"], cyclic_x=False,
                                      lat_s=s["lat_s"], lat_n=s["lat_n"],
                                      lon_span=s["lon_span"], land_frac=land_frac,
                                      seed=s["seed"], n_passive=n_passive, nts=s["nts"],
                                      passive_dtype=passive_dtype))
    return cases


__all__ = ["EnsembleStep", "make_regional_ensemble", "regional_shapes"]
