"""The elimination order of the register-resident FP32 column solver
(k_vdiff_twin), restated in numpy and checked on the CPU against a dense
solve of the same tridiagonal system.

A column's nk levels are split in two halves of nt = ceil(nk/2) and
nb = nk - nt rows held in H register slots each, the real rows last (slots
H-nact .. H-1, inert rows P = E = 0 before them).  The top half eliminates
downward, the bottom half upward (sub- and super-diagonal swapped); the two
meet at the interface between levels nt and nt+1, where
    x_top = d_top - E_top x_bot,   x_bot = d_bot - E_bot x_top
is solved, and each half back-substitutes.  The system is invtri's
(ocean_util.F90:175-236) in increment form: a x(k-1) + b x(k) + c x(k+1) =
bfl(k) - a (T(k-1) - T(k)) - c (T(k+1) - T(k)), with b = 1 - a - c, masked
rows (k > kmt) reducing to x = 0 and T(taup1) = (T + x) * tmask.
"""
import numpy as np
import pytest


def system(nk, rng, strong, kmt=None):
    """One column: a, c (<= 0, a(1) = c(nk) = 0, zero below kmt), T, bfl."""
    kmt = int(rng.integers(0, nk + 1)) if kmt is None else kmt
    d = 1e-4 + (10.0 if strong else 0.01) * rng.random(nk + 1) ** 4
    a = np.zeros(nk + 1)
    c = np.zeros(nk + 1)
    for k in range(1, nk + 1):
        if k > 1 and k <= kmt:
            a[k] = -d[k - 1]
        if k < nk and k + 1 <= kmt:
            c[k] = -d[k]
    T = rng.standard_normal(nk + 1) * 3 + 10
    bfl = np.zeros(nk + 1)
    if kmt >= 1:
        bfl[1] = 1e-3 * rng.standard_normal()
        bfl[max(2, kmt)] += 1e-3 * rng.standard_normal() if kmt >= 2 else 0.0
    return kmt, a, c, T, bfl


def dense(nk, kmt, a, c, T, bfl):
    """T(taup1) from a dense solve of the increment system."""
    M = np.zeros((nk, nk))
    r = np.zeros(nk)
    for k in range(1, nk + 1):
        i = k - 1
        wet = k <= kmt
        if not wet:
            M[i, i] = 1.0
            continue
        M[i, i] = 1.0 - a[k] - c[k]
        if k > 1:
            M[i, i - 1] = a[k]
        if k < nk:
            M[i, i + 1] = c[k]
        tm1 = T[k - 1] if k > 1 else 0.0
        tp1 = T[k + 1] if k < nk else 0.0
        r[i] = bfl[k] - a[k] * (tm1 - T[k]) - c[k] * (tp1 - T[k])
    x = np.linalg.solve(M, r)
    return np.array([(T[k] + x[k - 1]) if k <= kmt else 0.0 for k in range(1, nk + 1)])


def twin(nk, H, kmt, a, c, T, bfl, swap_meeting=False):
    """The two-ended elimination with k_vdiff_twin's slot layout."""
    nt = (nk + 1) // 2
    halves = {}
    for top in (True, False):
        nact = nt if top else nk - nt
        ph = H - nact
        lev = [None] * (H + 1)
        for j in range(H + 1):
            r = max(j - ph, 0)           # inert slots read the first real level
            lev[j] = (r + 1) if top else (nk - r)
        P = np.zeros(H)
        E = np.zeros(H)
        Eprev = 0.0
        for j in range(ph, H):
            k = lev[j]
            p, q = (a[k], c[k]) if top else (c[k], a[k])
            mkk = 1.0 if k <= kmt else 0.0
            b = 1.0 - p - q
            bet = mkk / (b - p * Eprev + 1e-30)
            P[j] = p * bet
            Eprev = q * bet
            E[j] = Eprev
        # forward sweep of the increment rhs (d = bet * rhs - P * dprev)
        Tv = [T[lev[j]] for j in range(H + 1)]
        dv = np.zeros(H)
        dprev = 0.0
        betv = np.zeros(H)
        Ep = 0.0
        for j in range(ph, H):           # bet again, for the bfl injection
            k = lev[j]
            p, q = (a[k], c[k]) if top else (c[k], a[k])
            mkk = 1.0 if k <= kmt else 0.0
            betv[j] = mkk / (1.0 - p - q - p * Ep + 1e-30)
            Ep = q * betv[j]
        for j in range(H):
            t1 = (Tv[j - 1] - Tv[j] if j > 0 else 0.0) + dprev
            t2 = Tv[j + 1] - Tv[j]
            dj = -P[j] * t1 - E[j] * t2
            if j >= ph:
                dj += betv[j] * bfl[lev[j]]
            dv[j] = dj
            dprev = dj
        halves[top] = (P, E, dv, Tv, lev, ph)
    dT, dB = halves[True][2][H - 1], halves[False][2][H - 1]
    ET, EB = halves[True][1][H - 1], halves[False][1][H - 1]
    if swap_meeting:
        ET, EB = EB, ET
    out = np.zeros(nk + 1)
    for top, (ds, Es, do, Eo) in ((True, (dT, ET, dB, EB)), (False, (dB, EB, dT, ET))):
        P, E, dv, Tv, lev, ph = halves[top]
        x = (ds - Es * do) / (1.0 - Es * Eo)
        for j in range(H - 1, ph - 1, -1):
            if j < H - 1:
                x = dv[j] - E[j] * x
            k = lev[j]
            out[k] = (Tv[j] + x) if k <= kmt else 0.0
    return out[1:]


@pytest.mark.parametrize("nk", [2, 3, 8, 9, 16, 17, 33, 64, 75, 80])
def test_twin_elimination_equals_dense_solve(nk):
    rng = np.random.default_rng(nk)
    H = next(h for h in (8, 16, 24, 32, 40) if h >= (nk + 1) // 2)
    for trial in range(40):
        kmt, a, c, T, bfl = system(nk, rng, strong=trial % 2 == 1)
        want = dense(nk, kmt, a, c, T, bfl)
        for h in (H, H + 8):             # inert slots change nothing
            got = twin(nk, h, kmt, a, c, T, bfl)
            assert np.allclose(got, want, rtol=1e-11, atol=1e-11), (nk, trial, h, kmt)


def test_meeting_with_swapped_multipliers_fails():
    """A plausible slip (each half solving the 2x2 with the other's E)
    is caught by the dense comparison."""
    rng = np.random.default_rng(5)
    nk, H = 20, 16
    bad = 0
    for _ in range(20):
        kmt, a, c, T, bfl = system(nk, rng, strong=True, kmt=nk)
        want = dense(nk, kmt, a, c, T, bfl)
        got = twin(nk, H, kmt, a, c, T, bfl, swap_meeting=True)
        bad += not np.allclose(got, want, rtol=1e-9, atol=1e-9)
    assert bad == 20
