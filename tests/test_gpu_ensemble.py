"""Ensembles (mom_gpu_ensemble_*): every member of an ensemble step equals,
bit for bit, the same member stepped alone through mom_gpu_step_tracers and
mom_gpu_update_ocean_barotropic (FP64 and FP32 tracers, barotropic state),
eagerly and from replays of a captured graph."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [dict(ni=60, nj=40, nk=8, nts=10, lat_s=10.0, lat_n=14.0, lon_span=6.0, seed=11),
          dict(ni=131, nj=70, nk=12, nts=16, lat_s=-40.0, lat_n=-33.0, lon_span=13.1, seed=12),
          dict(ni=37, nj=23, nk=7, nts=6, lat_s=30.0, lat_n=32.3, lon_span=3.7, seed=13),
          dict(ni=90, nj=95, nk=10, nts=12, lat_s=-5.0, lat_n=4.5, lon_span=9.0, seed=14),
          # nk = 21: the ensemble's register-resident FP32 solver runs every
          # domain with 16 register slots a wavefront, the others alone use 8
          dict(ni=48, nj=30, nk=21, nts=8, lat_s=20.0, lat_n=23.0, lon_span=4.8, seed=15)]


def _members(cases, fresh):
    from mom5_amd.step import BatchedStep
    out = []
    for c in cases:
        if fresh:
            dt = [np.float64, np.float64] + [np.float32] * (len(c.tracers) - 2)
            c = dataclasses.replace(c, tracers=[(nm, c.info["tracer"](i, dt[i]))
                                                for i, (nm, _) in enumerate(c.tracers)])
        out.append(BatchedStep(c, gpu=c.info["gpu"]))
    return out


def _assert_same(a_members, b_members, tag):
    for k, (a, b) in enumerate(zip(a_members, b_members)):
        for n in range(a.ntr):
            assert np.array_equal(a.tracer(n).to_host(), b.tracer(n).to_host()), (tag, k, n)
        assert np.array_equal(a.eta_bar[a.cur].to_host(), b.eta_bar[b.cur].to_host()), (tag, k)
        assert np.array_equal(a.udrho[a.cur].to_host(), b.udrho[b.cur].to_host()), (tag, k)


def _sync(members):
    for m in members:
        m.gpu.sync()


def test_ensemble_equals_members_alone_and_graph_replays():
    # the members alone and the ensemble share each domain's context (its
    # solver and subcycle scratch), so the two never run at the same time
    from mom5_amd.ensemble import EnsembleStep, make_regional_ensemble
    print("[ensemble: 5 regional domains, 2 FP64 + 3 FP32 tracers each]", flush=True)
    cases = make_regional_ensemble(SHAPES, n_passive=3)
    alone = _members(cases, fresh=True)
    ens_m = _members(cases, fresh=True)
    # other th_tendency terms on two of the domains (FP64 and FP32 tracers)
    from mom5_amd import DeviceArray
    rng = np.random.default_rng(3)
    for k in (1, 3):
        ths = [DeviceArray.from_host(1e-3 * rng.standard_normal(alone[k].tracer(n).shape))
               for n in range(alone[k].ntr)]
        alone[k].th = ens_m[k].th = ths
    ens = EnsembleStep(ens_m)
    for m in alone:
        m.run(3)
    _sync(alone)
    ens.run(3)
    ens.gpu.sync()
    _assert_same(alone, ens_m, "eager")
    # one more step (both time-level parities seen), then graph replays
    for m in alone:
        m.run(1)
    _sync(alone)
    ens.run(1)
    graph = ens.capture(2)
    for r in range(2):
        graph.launch()
        ens.gpu.sync()
        for m in alone:
            m.run(2)
        _sync(alone)
        _assert_same(alone, ens_m, f"replay {r}")
    graph.close()
    ens.close()


def test_ensemble_members_match_oracle():
    """Every member of an ensemble stepped three times against the oracle's
    own step sequence on that member's domain (MDFL-Sweby advection,
    explicit update, implicit vertical diffusion, B-grid barotropic
    subcycles with their time averages): T and S within 1e-12 per cell (the
    kernels' limited fluxes are division-free), the barotropic state bit for
    bit, the FP32 passive tracers within 1e-5 of max |T| (the FP64
    oracle carries them in FP64).  One member also adds other th_tendency
    terms."""
    from mom5_amd import DeviceArray
    from mom5_amd.ensemble import EnsembleStep, make_regional_ensemble
    from oracle.oracle import Oracle
    from _cases import assert_close_cells, host_case, oracle_tracer_step
    nsteps = 3
    print("[ensemble vs oracle: 5 regional domains, 2 FP64 + 3 FP32 tracers, 3 steps]",
          flush=True)
    cases = make_regional_ensemble(SHAPES, n_passive=3)
    members = _members(cases, fresh=True)
    rng = np.random.default_rng(5)
    th_host = {2: [1e-3 * rng.standard_normal(members[2].tracer(n).shape)
                   for n in range(members[2].ntr)]}
    members[2].th = [DeviceArray.from_host(t) for t in th_host[2]]
    # the oracle's initial state: each member's initial arrays widened to FP64
    init = [[m.tracer(n).to_host().astype(np.float64) for n in range(m.ntr)] for m in members]
    ens = EnsembleStep(members)
    ens.run(nsteps)
    ens.gpu.sync()
    for k, (c, m) in enumerate(zip(cases, members)):
        hc = host_case(c)
        o = Oracle(hc)
        f = hc.fields
        Ts = init[k]
        eb, ud = f["eta_t_bar"].copy(), f["udrho"].copy()
        for _ in range(nsteps):
            Ts = oracle_tracer_step(o, Ts, f["stf"][:m.ntr], f["btf"][:m.ntr], hc.dt,
                                    ths=th_host.get(k))
            fl = dict(f)
            fl.update(eta_t=eb, eta_t_bar=eb, udrho=ud)
            eb, ud, _ = o.update_ocean_barotropic(hc.nts, hc.dtbt, fields=fl)
        wet = hc.grid["tmask"][:, 1:-1, 1:-1] > 0
        for n in range(m.ntr):
            got = m.tracer(n).to_host()[:, 1:-1, 1:-1]
            ref = Ts[n][:, 1:-1, 1:-1]
            if n < 2:
                assert_close_cells(got, ref, what=f"member {k} tracer {n}")
            else:
                err = np.abs(got.astype(np.float64)[wet] - ref[wet]).max()
                assert err <= 1e-5 * np.abs(ref[wet]).max(), (k, n, err)
        assert np.array_equal(m.eta_bar[m.cur].to_host(), eb), k
        assert np.array_equal(m.udrho[m.cur].to_host(), ud), k
    ens.close()
