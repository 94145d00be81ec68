"""The CPU restatement of ocean_barotropic_forcing (default namelist,
B-grid; ocean_barotropic.F90:3058-3141) pinned by closed forms:

* constant accel A in every wet U-cell and no stress: forcing_bt = kmu*A
  (the vertical integral counts exactly the wet levels; dyadic A: exact);
* stress only: forcing_bt = umask(k=1)*(smf - bmf);
* accel below kmu is ignored (a mask error adds it);
* halo: cyclic-x columns wrap on the compute rows, closed rows and the
  corners stay zero.
"""
import numpy as np

from _cases import uniform_case


def _case():
    c = uniform_case(10, 7, 5, cyclic_x=True)
    kmt = c.grid["kmt"]
    kmt[3, 4] = 2
    kmt[5, 7] = 0
    kmt[:, 0] = kmt[:, -2]
    kmt[:, -1] = kmt[:, 1]
    kmu = np.zeros_like(kmt)
    kmu[:-1, :-1] = np.minimum.reduce([kmt[:-1, :-1], kmt[:-1, 1:], kmt[1:, :-1], kmt[1:, 1:]])
    kmu[:, -1] = kmu[:, 1]
    c.grid["kmu"] = kmu
    kk = np.arange(1, 6)[:, None, None]
    c.grid["tmask"] = (kk <= kmt[None]).astype(np.float64)
    c.grid["umask"] = (kk <= kmu[None]).astype(np.float64)
    return c


def test_constant_accel_counts_wet_levels(oracle_lib):
    c = _case()
    o = oracle_lib.Oracle(c)
    A = 0.375
    accel = np.full((2, 5, 9, 12), A)
    z = np.zeros((2, 9, 12))
    f = o.barotropic_forcing(accel, z, z)
    kmu = c.grid["kmu"].astype(np.float64)
    for n in range(2):
        assert np.array_equal(f[n][1:-1, 1:-1], (kmu * A)[1:-1, 1:-1])


def test_stress_only_and_mask(oracle_lib):
    c = _case()
    o = oracle_lib.Oracle(c)
    rng = np.random.default_rng(3)
    smf = rng.standard_normal((2, 9, 12))
    bmf = rng.standard_normal((2, 9, 12))
    accel = np.zeros((2, 5, 9, 12))
    # large values below the bottom must not count
    um = c.grid["umask"]
    accel[:, um == 0.0] = 1.0e6
    f = o.barotropic_forcing(accel, smf, bmf)
    want = um[0] * (smf - bmf)
    assert np.array_equal(f[:, 1:-1, 1:-1], want[:, 1:-1, 1:-1])


def test_halo_wrap(oracle_lib):
    c = _case()
    o = oracle_lib.Oracle(c)
    rng = np.random.default_rng(4)
    accel = rng.standard_normal((2, 5, 9, 12))
    f = o.barotropic_forcing(accel, np.zeros((2, 9, 12)), np.zeros((2, 9, 12)))
    assert np.array_equal(f[:, 1:-1, 0], f[:, 1:-1, -2])
    assert np.array_equal(f[:, 1:-1, -1], f[:, 1:-1, 1])
    assert np.all(f[:, 0, :] == 0.0) and np.all(f[:, -1, :] == 0.0)
