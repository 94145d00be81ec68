"""Pins the invtri restatement (ocean_util.F90:153-222) and the tracer loop
of vert_diffuse_implicit (ocean_vert_mix.F90:3426-3513):

* each wet column solves the backward-Euler discretisation of
  d(dh*z)/dt = d/dz(dcb dz/dz) with flux boundary conditions, checked
  against a dense solve of that system assembled from the PDE;
* total column content sum(dh*z) changes exactly by tdt*(topbc - botbc);
* land columns come back zero; a one-level column keeps only the top flux
  (the bottom condition lands at level max(2,kmt), a masked level);
* nmix = 2 (salt) reads the second diff_cbt plane.
"""
import numpy as np
import pytest

from _cases import uniform_case


def _column_case(nk=9, ni=7, nj=5, seed=0):
    rng = np.random.default_rng(seed)
    kmt = np.zeros((nj + 2, ni + 2), dtype=np.int32)
    kmt[1:-1, 1:-1] = rng.integers(0, nk + 1, (nj, ni))
    kmt[1, 1] = 1
    kmt[1, 2] = 0
    kmt[2, 2] = nk
    c = uniform_case(ni, nj, nk, kmt=kmt, cyclic_x=False)
    dh = 1035.0 * (1.0 + rng.random((nk, nj + 2, ni + 2)))
    dhw = 1.0 + 3.0 * rng.random((nk + 1, nj + 2, ni + 2))
    dcb = 1035.0 * 10 ** rng.uniform(-5, -1, (nk, nj + 2, ni + 2))
    z = 10.0 + rng.standard_normal((nk, nj + 2, ni + 2))
    top = rng.standard_normal((nj + 2, ni + 2))
    bot = rng.standard_normal((nj + 2, ni + 2))
    return c, kmt, dh, dhw, dcb, z, top, bot


def _dense_solution(zc, dhc, dhwc, dcbc, kb, tdt, top, bot):
    """Assemble (dh_k z_k - tdt [F_{k-1/2} - F_{k+1/2}]) = dh_k z0_k + BC."""
    n = kb
    A = np.zeros((n, n))
    rhs = dhc[:n] * zc[:n]
    for k in range(n):
        A[k, k] = dhc[k]
        if k > 0:                       # diffusive flux through the top face of k
            g = tdt * dcbc[k - 1] / dhwc[k]
            A[k, k] += g
            A[k, k - 1] -= g
        if k < n - 1:                   # through the bottom face of k
            g = tdt * dcbc[k] / dhwc[k + 1]
            A[k, k] += g
            A[k, k + 1] -= g
    rhs[0] += tdt * top
    rhs[n - 1] -= tdt * bot
    return np.linalg.solve(A, rhs)


def test_matches_dense_solve_and_conserves(oracle_lib):
    c, kmt, dh, dhw, dcb, z, top, bot = _column_case()
    o = oracle_lib.Oracle(c)
    tdt = 3600.0
    out = o.invtri(z.copy(), top, bot, dcb, tdt, dh, dhw)
    nk = c.nk
    checked = 0
    for j in range(1, c.nj + 1):
        for i in range(1, c.ni + 1):
            kb = kmt[j, i]
            if kb >= 2:
                ref = _dense_solution(z[:, j, i], dh[:, j, i], dhw[:, j, i], dcb[:, j, i], kb,
                                      tdt, top[j, i], bot[j, i])
                assert np.allclose(out[:kb, j, i], ref, rtol=1e-12, atol=1e-12)
                assert np.all(out[kb:, j, i] == 0.0)
                c0 = (dh[:kb, j, i] * z[:kb, j, i]).sum() + tdt * (top[j, i] - bot[j, i])
                c1 = (dh[:kb, j, i] * out[:kb, j, i]).sum()
                assert abs(c1 - c0) <= 1e-13 * np.abs(dh[:kb, j, i] * z[:kb, j, i]).sum()
                checked += 1
            elif kb == 0:
                assert np.all(out[:, j, i] == 0.0)
            else:       # kmt = 1: top flux only, bottom condition lost
                exp = (z[0, j, i] + top[j, i] * tdt / dh[0, j, i]) / (1.0 + 1e-30)
                assert out[0, j, i] == pytest.approx(exp, rel=1e-15)
                assert np.all(out[1:, j, i] == 0.0)
    assert checked > 5


def test_wrong_operand_order_would_fail(oracle_lib):
    """A transposed dhw index (dhw(k) in place of dhw(k-1)) breaks the dense check."""
    c, kmt, dh, dhw, dcb, z, top, bot = _column_case(seed=1)
    o = oracle_lib.Oracle(c)
    out = o.invtri(z.copy(), top, bot, dcb, 3600.0, dh, dhw)
    j, i = 2, 2
    kb = kmt[j, i]
    bad = _dense_solution(z[:, j, i], dh[:, j, i], np.roll(dhw[:, j, i], -1), dcb[:, j, i], kb,
                          3600.0, top[j, i], bot[j, i])
    assert not np.allclose(out[:kb, j, i], bad, rtol=1e-8)


def test_vert_diffuse_implicit_uses_nmix(oracle_lib):
    from mom5_amd.synth import make_case
    case = make_case(20, 14, 8, nts=10, seed=31, stf_scale=1e-3)
    o = oracle_lib.Oracle(case)
    f = case.fields
    Ts = [t.copy() for _, t in case.tracers]
    out = o.vert_diffuse_implicit([t.copy() for t in Ts], f["stf"], f["btf"], case.dt,
                                  index_salt=1)
    for n, plane in ((0, 0), (1, 1)):
        ref = o.invtri(Ts[n].copy(), f["stf"][n], f["btf"][n], 1035.0 * f["diff_cbt"][plane],
                       case.dt, f["rho_dzt"], f["dzwt"])
        assert np.array_equal(out[n][:, 1:-1, 1:-1], ref[:, 1:-1, 1:-1])
    wrong = o.invtri(Ts[1].copy(), f["stf"][1], f["btf"][1], 1035.0 * f["diff_cbt"][0],
                     case.dt, f["rho_dzt"], f["dzwt"])
    assert not np.array_equal(out[1][:, 1:-1, 1:-1], wrong[:, 1:-1, 1:-1])


def test_one_level_refused(oracle_lib):
    """invtri sets its bottom row at level max(2, kmt): the restatement
    refuses a one-level domain (as the library does) instead of writing past
    the column."""
    from mom5_amd.synth import make_case
    case = make_case(12, 6, 1, land_frac=0.2, nts=2, seed=5, n_passive=0)
    o = oracle_lib.Oracle(case)
    T = case.tracers[0][1].copy()
    with pytest.raises(ValueError):
        o.vert_diffuse_implicit([T], case.fields["stf"][:1], case.fields["btf"][:1], case.dt)
