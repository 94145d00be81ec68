"""Hand-built domains for the closed-form and invariant checks."""
import types

import numpy as np


def uniform_case(ni, nj, nk, *, cyclic_x=True, dx=1.0, dy=1.0, dz=1.0, rho=1.0, f=0.0,
                 kmt=None, land_rows=True):
    """A Cartesian box: every metric constant, every interior column kmt=nk
    (unless given); closed rows j=0 and j=nj+1 are land."""
    NI, NJ = ni + 2, nj + 2
    one = np.ones((NJ, NI))
    if kmt is None:
        kmt = np.zeros((NJ, NI), dtype=np.int32)
        kmt[1:-1, 1:-1] = nk
        if cyclic_x:
            kmt[:, 0] = kmt[:, -2]
            kmt[:, -1] = kmt[:, 1]
    kmt = np.ascontiguousarray(kmt, dtype=np.int32)
    kmu = np.zeros_like(kmt)
    kmu[:-1, :-1] = np.minimum.reduce([kmt[:-1, :-1], kmt[:-1, 1:], kmt[1:, :-1], kmt[1:, 1:]])
    if cyclic_x:
        kmu[:, -1] = kmu[:, 1]
    kk = np.arange(1, nk + 1)[:, None, None]
    tmask = (kk <= kmt[None]).astype(np.float64)
    umask = (kk <= kmu[None]).astype(np.float64)
    grid = dict(dxt=dx * one, dyt=dy * one, dat=dx * dy * one, datr=one / (dx * dy),
                dxte=dx * one, dyte=dy * one, dxtn=dx * one, dytn=dy * one, dxu=dx * one,
                dyu=dy * one, dxur=one / dx, dyur=one / dy, f=f * one, kmt=kmt, kmu=kmu,
                tmask=tmask, umask=umask, dzt=dz * np.ones(nk))
    z3 = np.zeros((nk, NJ, NI))
    fields = dict(uhrho_et=z3.copy(), vhrho_nt=z3.copy(), wrho_bt=np.zeros((nk + 1, NJ, NI)),
                  rho_dzt=rho * np.ones((nk, NJ, NI)), rho_dztr=np.ones((nk, NJ, NI)) / rho,
                  dzwt=dz * np.ones((nk + 1, NJ, NI)), diff_cbt=np.zeros((2, nk, NJ, NI)))
    c = types.SimpleNamespace(ni=ni, nj=nj, nk=nk, cyclic_x=cyclic_x, grid=grid, fields=fields,
                              dom=dict(isd=0, ied=ni + 1, jsd=0, jed=nj + 1, isc=1, iec=ni,
                                       jsc=1, jec=nj, nk=nk, cyclic_x=int(cyclic_x), cyclic_y=0))
    return c


def bt_tmasken(c):
    """Grd%tmasken(:,:,1,1:2) (ocean_grids.F90:1390-1400): face masks up
    to iec/jec, halo-updated (cyclic x), zero beyond."""
    t = c.grid["tmask"][0]
    NJ, NI = t.shape
    m = np.zeros((2, NJ, NI))
    m[0, :-1, :-1] = np.minimum(t[:-1, :-1], t[:-1, 1:])
    m[1, :-1, :-1] = np.minimum(t[:-1, :-1], t[1:, :-1])
    if c.dom["cyclic_x"]:
        m[:, 1:-1, 0] = m[:, 1:-1, -2]
        m[:, 1:-1, -1] = m[:, 1:-1, 1]
    return m


def bt_variant_fields(c, rng, cgrid, scale=1.0):
    """Random barotropic-variant inputs on case c (masks respected)."""
    sh = c.grid["tmask"][0].shape
    tm = c.grid["tmask"][0]
    um = c.grid["umask"][0]
    mask2 = bt_tmasken(c) if cgrid else np.stack([um, um])
    eta = 0.05 * rng.standard_normal(sh) * tm
    ud = rng.standard_normal((2,) + sh) * mask2 * scale
    fb = 1e-3 * rng.standard_normal((2,) + sh) * mask2
    H = 1000.0 + 3000.0 * rng.random(sh)
    thick = (H[None] * mask2) if cgrid else H * um
    return dict(x_tau=eta, xbar_tau=eta.copy(), udrho=ud, forcing_bt=fb, thick=thick,
                tmasken=bt_tmasken(c) if cgrid else None, patm=np.zeros(sh),
                rho_g=np.full(sh, 1035.0 * 9.80), dpatm_dt=np.zeros(sh),
                pme=1e-4 * rng.standard_normal(sh) * tm, river=np.zeros(sh), source=np.zeros(sh),
                smooth=np.zeros(sh))


def oracle_tracer_step(orc, Ts, stfs, btfs, dt, *, index_salt=1, ths=None, fields=None,
                       schemes=None, mins=None, maxs=None, index_temp=0, ts_same=True,
                       limit_with_upwind=False):
    """The oracle's sequence for one batched tracer step of FP64 tracers Ts
    (halo-updated): compute_tmask_limit when a tracer is Quicker, per tracer
    th = (other tendencies or 0) + its advection (MDFL-Sweby, or Quicker
    horizontal then vertical), the explicit update, then the implicit
    vertical diffusion of all of them (ocean_tracer.F90:2295-2350, 2432,
    2788-2892).  Returns the T(taup1) arrays."""
    lims = None
    if schemes is not None and "quicker" in schemes:
        lims = orc.tmask_limit(Ts, mins, maxs, index_temp=index_temp, index_salt=index_salt,
                               ts_same=ts_same)
    outs = []
    for n, T in enumerate(Ts):
        th = np.zeros_like(T) if ths is None or ths[n] is None else np.array(ths[n], copy=True)
        if schemes is None or schemes[n] == "mdfl_sweby":
            orc.horz_advect_mdfl(T, th, dt, lim=1.0, fields=fields)
        else:
            rh = orc.horz_quicker(T, T, tmask_limit=lims[n], limit_with_upwind=limit_with_upwind,
                                  fields=fields)
            rv = orc.vert_quicker(T, T, tmask_limit=lims[n], fields=fields)
            th[:, 1:-1, 1:-1] = (th[:, 1:-1, 1:-1] + (-rh)) + (-rv)
        outs.append(orc.update_explicit(T, th, dt, fields=fields))
    orc.vert_diffuse_implicit(outs, stfs, btfs, dt, index_salt=index_salt, fields=fields)
    return outs


def host_case(case):
    """A host copy of a case whose 3-D fields and tracers live on the device
    (mom5_amd.synth.make_case_device), with the 3-D masks the oracle reads."""
    import dataclasses
    grid = dict(case.grid)
    nk = case.nk
    kk = np.arange(1, nk + 1)[:, None, None]
    grid["tmask"] = (kk <= grid["kmt"][None]).astype(np.float64)
    grid["umask"] = (kk <= grid["kmu"][None]).astype(np.float64)
    conv = lambda v: v.to_host() if hasattr(v, "to_host") else v
    fields = {k: (conv(v) if not isinstance(v, list) else [conv(x) for x in v])
              for k, v in case.fields.items()}
    tracers = [(nm, conv(t).astype(np.float64)) for nm, t in case.tracers]
    return dataclasses.replace(case, grid=grid, fields=fields, tracers=tracers)


# ---- FP64 tolerance of the division-free MDFL-Sweby arithmetic ---------------
# The HIP kernels form the limited flux without its divisions and the update
# factors with reciprocals (rounding differs from the reference's quotients by
# a few ulp); the bar is 1e-12 relative per cell on the updated tracer, and
# for intermediate quantities (tendencies, fluxes) 1e-12 of the quantity's
# natural scale.
RTOL64 = 1e-12


def assert_close_cells(got, ref, rtol=RTOL64, what=""):
    """|got - ref| <= rtol*|ref| per cell (a floor of 1e-6*max|ref| keeps cells
    that are exactly zero in the reference from demanding exact zeros)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    assert np.isfinite(got).all(), f"{what}: non-finite values"
    scale = np.maximum(np.abs(ref), 1e-6 * np.abs(ref).max())
    err = np.abs(got - ref)
    bad = err > rtol * scale
    if bad.any():
        i = tuple(np.argwhere(bad)[0])
        raise AssertionError(f"{what}: {int(bad.sum())} cells beyond {rtol:g} relative, first at "
                             f"{i}: {got[i]!r} vs {ref[i]!r} (max rel "
                             f"{(err / np.where(scale > 0, scale, 1.0)).max():.3e})")
    return float((err / np.where(scale > 0, scale, 1.0)).max())


def assert_close_scaled(got, ref, scale, rtol=RTOL64, what=""):
    """|got - ref| <= rtol*scale elementwise (scale: an array or a number)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    assert np.isfinite(got).all(), f"{what}: non-finite values"
    err = np.abs(got - ref)
    tol = rtol * np.broadcast_to(np.asarray(scale, dtype=np.float64), err.shape)
    bad = err > tol
    if bad.any():
        i = tuple(np.argwhere(bad)[0])
        raise AssertionError(f"{what}: {int(bad.sum())} values beyond {rtol:g} of scale, first "
                             f"at {i}: {got[i]!r} vs {ref[i]!r} (err {err[i]:.3e}, tol "
                             f"{tol[i]:.3e})")
