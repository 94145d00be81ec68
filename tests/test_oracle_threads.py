"""The all-core build of the CPU restatement (-fopenmp, used for the all-core
CPU baseline) gives the serial build's results bit for bit: its parallel
loops run over independent levels, rows or points.  Run with four threads on a
cyclic case with land, so that every parallel loop has work on several
threads."""
import os

import numpy as np
import pytest

from mom5_amd.synth import make_case


@pytest.fixture(scope="module")
def pair(oracle_lib):
    os.environ.setdefault("OMP_NUM_THREADS", "4")
    case = make_case(40, 26, 9, nts=6, n_passive=1, seed=77)
    return case, oracle_lib.Oracle(case), oracle_lib.Oracle(case, threads=True)


def test_threaded_build_is_another_library(pair):
    _, serial, threaded = pair
    assert serial.L is not threaded.L
    assert serial.L._name != threaded.L._name


def test_mdfl_update_and_diffusion_identical(pair):
    case, serial, threaded = pair
    f = case.fields
    out = []
    for o in (serial, threaded):
        res = []
        for _, T in case.tracers:
            th = np.zeros_like(T)
            o.horz_advect_mdfl(T, th, case.dt)
            res.append(o.update_explicit(T, th, case.dt))
        n = len(res)
        o.vert_diffuse_implicit(res, f["stf"][:n], f["btf"][:n], case.dt, index_salt=1)
        out.append(res)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_quicker_and_tmask_limit_identical(pair):
    case, serial, threaded = pair
    Ts = [t for _, t in case.tracers]
    out = []
    for o in (serial, threaded):
        lim = o.tmask_limit(Ts, [-2.0, 30.0, -0.1], [25.0, 36.0, 0.5], 0, 1, True)
        h = o.horz_quicker(Ts[0], Ts[0], lim[0])
        v = o.vert_quicker(Ts[0], Ts[0], lim[0])
        out.append((lim, h, v))
    (l1, h1, v1), (l2, h2, v2) = out
    assert all(np.array_equal(a, b) for a, b in zip(l1, l2))
    assert np.array_equal(h1, h2) and np.array_equal(v1, v2)


def test_barotropic_identical(pair):
    case, serial, threaded = pair
    r1 = serial.update_ocean_barotropic(case.nts, case.dtbt)
    r2 = threaded.update_ocean_barotropic(case.nts, case.dtbt)
    for a, b in zip(r1, r2):
        assert np.array_equal(a, b)
