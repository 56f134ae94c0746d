"""Scalar 5th-order / WENO5 advection oracle (test infrastructure only).

Restates ``cuda_advection5.py`` of the reference:
  * ``scalar_advection_uses_weno``        -> cuda_advection5.py:129-147
  * ``cpu_terrain_mass_flux5_inplace``    -> cuda_advection5.py:270-437
  * ``cpu_scalar_advection5_inplace``     -> cuda_advection5.py:440-579
The tendency of a whole-domain launch (``lateral_sides`` given: every cell,
with the open-side inflow edits) has no CPU form in the reference; it follows
the kernel text ``cm1_adv5_scalar_tendency_v1`` (cuda_advection5.py:669) and
is pinned by invariants only (parity unpinned against reference words in the
outer three-cell band).
"""

from __future__ import annotations

import numpy as np

from .common import F, NG, flx3, lateral_code, recon5, upstrpd


def scalar_advection_uses_weno(nrk: int, *, nrkmax: int = 3, wflag: bool = True) -> bool:
    """cuda_advection5.py:129-147: WENO only on the final RK stage."""
    if nrkmax != 3 or nrk < 1 or nrk > nrkmax:
        raise ValueError("the sealed low-storage schedule requires nrkmax=3 and nrk=1..3")
    return bool(wflag and nrk == nrkmax)


def cpu_terrain_mass_flux5_inplace(arrays, metrics, divx, *, ni, nj, nk, dx, dy, dz):
    """cuda_advection5.py:270-437 (interior faces/columns only)."""
    ng = NG
    u, v, w = arrays["u3d"], arrays["v3d"], arrays["w3d"]
    rru, rrv, rrw = arrays["rru"], arrays["rrv"], arrays["rrw"]
    rho0 = metrics["rho0_3d"]
    half = F(0.5)
    ks = slice(1, nk + 1)
    ys = slice(ng, ng + nj)
    # :328-341 rru faces 1..ni-1
    dens = rho0[:, :, 0 : ni - 1] + rho0[:, :, 1:ni]
    rru[ks, ys, ng + 1 : ng + ni] = (half * dens) * (
        u[ks, ys, ng + 1 : ng + ni] * metrics["rgzu"][None, :, 1:ni]
    )
    # :342-355 rrv faces 1..nj-1
    dens = rho0[:, 0 : nj - 1, :] + rho0[:, 1:nj, :]
    xs = slice(ng, ng + ni)
    rrv[ks, ng + 1 : ng + nj, xs] = (half * dens) * (
        v[ks, ng + 1 : ng + nj, xs] * metrics["rgzv"][None, 1:nj, :]
    )
    rrw[1, ng : ng + nj, ng : ng + ni] = F(0.0)
    rrw[nk + 1, ng : ng + nj, ng : ng + ni] = F(0.0)
    dz32 = F(dz)
    zt = F(nk) * dz32
    rzt = F(1.0) / zt
    yi = slice(ng + 1, ng + nj - 1)
    xi = slice(ng + 1, ng + ni - 1)
    xi1 = slice(ng + 2, ng + ni)
    yi1 = slice(ng + 2, ng + nj)
    dzdx = metrics["dzdx"][1 : nj - 1, 1 : ni - 1]
    dzdy = metrics["dzdy"][1 : nj - 1, 1 : ni - 1]
    gz = metrics["gz"][1 : nj - 1, 1 : ni - 1]
    for k in range(2, nk + 1):
        sigmaf = F(k - 1) * dz32
        sigma = F(k - 1.5) * dz32
        r2 = (sigmaf - sigma) * (F(1.0) / dz32)
        r1 = half * (F(1.0) - r2)
        r2 = half * r2
        ux = r2 * (rru[k, yi, xi] + rru[k, yi, xi1]) + r1 * (rru[k - 1, yi, xi] + rru[k - 1, yi, xi1])
        vy = r2 * (rrv[k, yi, xi] + rrv[k, yi1, xi]) + r1 * (rrv[k - 1, yi, xi] + rrv[k - 1, yi1, xi])
        terrain = ux * dzdx + vy * dzdy
        terrain = terrain * (sigmaf - zt)
        terrain = terrain * gz
        terrain = terrain * rzt
        rrw[k, yi, xi] = metrics["rf0"][k - 1, 1 : nj - 1, 1 : ni - 1] * w[k, yi, xi] + terrain
    rdx = F(1.0) / F(dx)
    rdy = F(1.0) / F(dy)
    rdsf = F(1.0) / dz32
    value = (rru[ks, yi, xi1] - rru[ks, yi, xi]) * rdx + (rrv[ks, yi1, xi] - rrv[ks, yi, xi]) * rdy
    value = value + (rrw[2 : nk + 2, yi, xi] - rrw[1 : nk + 1, yi, xi]) * rdsf
    divx[:, 1 : nj - 1, 1 : ni - 1] = np.where(np.abs(value.astype(np.float64)) < 1.0e-20, F(0.0), value)


def _xface_fluxes(scalar, rru, ni, nj, nk, use_weno, weps):
    """x-face fluxes at faces 0..ni of rows 0..nj-1, levels 1..nk (:493-504)."""
    ng = NG
    S = scalar[1 : nk + 1, ng : ng + nj, :]
    x = np.arange(ni + 1) + ng
    mass = rru[1 : nk + 1, ng : ng + nj, ng : ng + ni + 1]
    pos = recon5(S[..., x - 3], S[..., x - 2], S[..., x - 1], S[..., x], S[..., x + 1], use_weno, weps)
    neg = recon5(S[..., x + 2], S[..., x + 1], S[..., x], S[..., x - 1], S[..., x - 2], use_weno, weps)
    return mass * np.where(mass >= F(0.0), pos, neg)


def _yface_fluxes(scalar, rrv, ni, nj, nk, use_weno, weps):
    """y-face fluxes at faces 0..nj of columns 0..ni-1 (:505-515)."""
    ng = NG
    S = scalar[1 : nk + 1, :, ng : ng + ni]
    y = np.arange(nj + 1) + ng
    mass = rrv[1 : nk + 1, ng : ng + nj + 1, ng : ng + ni]
    pos = recon5(S[:, y - 3], S[:, y - 2], S[:, y - 1], S[:, y], S[:, y + 1], use_weno, weps)
    neg = recon5(S[:, y + 2], S[:, y + 1], S[:, y], S[:, y - 1], S[:, y - 2], use_weno, weps)
    return mass * np.where(mass >= F(0.0), pos, neg)


def _zface_flux(scalar, rrw, c1, c2, k, ni, nj, nk, use_weno, weps):
    """z-face flux at face k (2..nk) with the lower-order closures (:516-559)."""
    ng = NG
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    s = lambda level: scalar[level, ys, xs]  # noqa: E731
    mass = rrw[k, ys, xs]
    cc1, cc2 = c1[k - 1], c2[k - 1]
    lin = cc1 * s(k - 1) + cc2 * s(k)
    if k == 2:
        neg = upstrpd(s(k + 1), s(k), s(k - 1), weps) if use_weno else flx3(s(k + 1), s(k), s(k - 1))
        recon = np.where(mass >= F(0.0), lin, neg)
    elif k == nk:
        up = upstrpd(s(k - 2), s(k - 1), s(k), weps) if use_weno else flx3(s(k - 2), s(k - 1), s(k))
        cond = (mass > F(0.0)) if use_weno else (mass >= F(0.0))
        recon = np.where(cond, up, lin)
    elif k == 3:
        pos = upstrpd(s(k - 2), s(k - 1), s(k), weps) if use_weno else flx3(s(k - 2), s(k - 1), s(k))
        neg = recon5(s(k + 2), s(k + 1), s(k), s(k - 1), s(k - 2), use_weno, weps)
        recon = np.where(mass >= F(0.0), pos, neg)
    elif k == nk - 1:
        pos = recon5(s(k - 3), s(k - 2), s(k - 1), s(k), s(k + 1), use_weno, weps)
        neg = upstrpd(s(k + 1), s(k), s(k - 1), weps) if use_weno else flx3(s(k + 1), s(k), s(k - 1))
        recon = np.where(mass >= F(0.0), pos, neg)
    else:
        pos = recon5(s(k - 3), s(k - 2), s(k - 1), s(k), s(k + 1), use_weno, weps)
        neg = recon5(s(k + 2), s(k + 1), s(k), s(k - 1), s(k - 2), use_weno, weps)
        recon = np.where(mass >= F(0.0), pos, neg)
    return mass * recon


def cpu_scalar_advection5_inplace(
    scalar, arrays, metrics, divx, tendency, fluxes, *, ni, nj, nk, nrk, dx, dy, dz,
    weps, wflag=True, lateral_sides=None, flat=None,
):
    """cuda_advection5.py:440-579; ``lateral_sides`` adds the whole-domain form.

    ``flat`` (mh, rdz of a flat stretched column) selects the
    terrain_flag=.false. tendency of cuda_gm1_child_stretched_scalar.py
    :262-352 (adv.F:458-478, 555-572, 646-656): -(fe-fw)*rdx,
    ((d*rdz)*mh(k)) vertically and the scale rr0(1,1,k).

    With ``lateral_sides=None`` only the tile-local interior
    ``3 <= i < ni-3, 3 <= j < nj-3`` is written (the reference CPU
    transcription, :563-579).  Otherwise every cell is written with the
    open-side inflow edits of ``cm1_adv5_scalar_tendency_v1`` (:628-634,
    :669).
    """
    ng = NG
    use_weno = scalar_advection_uses_weno(nrk, wflag=wflag)
    rru, rrv, rrw = arrays["rru"], arrays["rrv"], arrays["rrw"]
    fx, fy, fz = fluxes["x"], fluxes["y"], fluxes["z"]
    c1, c2, rr0, gz = metrics["c1"], metrics["c2"], metrics["rr0"], metrics["gz"]
    fx[1 : nk + 1, ng : ng + nj, ng : ng + ni + 1] = _xface_fluxes(scalar, rru, ni, nj, nk, use_weno, weps)
    fy[1 : nk + 1, ng : ng + nj + 1, ng : ng + ni] = _yface_fluxes(scalar, rrv, ni, nj, nk, use_weno, weps)
    fz[1, ng : ng + nj, ng : ng + ni] = F(0.0)
    fz[nk + 1, ng : ng + nj, ng : ng + ni] = F(0.0)
    for k in range(2, nk + 1):
        fz[k, ng : ng + nj, ng : ng + ni] = _zface_flux(scalar, rrw, c1, c2, k, ni, nj, nk, use_weno, weps)
    rdx = F(1.0) / F(dx)
    rdy = F(1.0) / F(dy)
    rdsf = F(1.0) / F(dz)
    ks = slice(1, nk + 1)
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    fw = fx[ks, ys, ng : ng + ni].copy()
    fe = fx[ks, ys, ng + 1 : ng + ni + 1].copy()
    fs = fy[ks, ng : ng + nj, xs].copy()
    fn = fy[ks, ng + 1 : ng + nj + 1, xs].copy()
    sv = scalar[ks, ys, xs]
    obc = lateral_code(lateral_sides)
    shape = (nk, nj, ni)
    iw = np.zeros(shape, bool)
    ie = np.zeros(shape, bool)
    js = np.zeros(shape, bool)
    jn = np.zeros(shape, bool)
    uw = rru[ks, ys, ng : ng + ni]
    ue = rru[ks, ys, ng + 1 : ng + ni + 1]
    vs = rrv[ks, ng : ng + nj, xs]
    vn = rrv[ks, ng + 1 : ng + nj + 1, xs]
    if obc:
        side = lambda s: (obc >> (4 * s)) & 15  # noqa: E731
        if side(0) == 2:
            iw[:, :, 0] = uw[:, :, 0] >= F(0.0)
        if side(1) == 2:
            ie[:, :, ni - 1] = ue[:, :, ni - 1] <= F(0.0)
        if side(2) == 2:
            js[:, 0, :] = vs[:, 0, :] >= F(0.0)
        if side(3) == 2:
            jn[:, nj - 1, :] = vn[:, nj - 1, :] <= F(0.0)
        fw = np.where(iw, fe, fw)
        fe = np.where(ie, fw, fe)
        fs = np.where(js, fn, fs)
        fn = np.where(jn, fs, fn)
    if flat is not None:
        ax = (-(fe - fw)) * rdx
        ay = (-(fn - fs)) * rdy
    else:
        ax = (fw - fe) * rdx
        ay = (fs - fn) * rdy
    if obc:
        cx = (sv * (ue - uw)) * rdx
        cy = (sv * (vn - vs)) * rdy
        ax = np.where(iw, ax - cx, ax)
        ax = np.where(ie, ax - cx, ax)
        ay = np.where(js, ay - cy, ay)
        ay = np.where(jn, ay - cy, ay)
    az = np.empty(shape, F)
    if flat is not None:
        rdz, mh = F(flat.rdz), flat.mh[1 : nk + 1][:, None, None]
        az[0] = (-fz[2, ys, xs]) * rdz
        az[nk - 1] = fz[nk, ys, xs] * rdz
        az[1 : nk - 1] = (-(fz[3 : nk + 1, ys, xs] - fz[2:nk, ys, xs])) * rdz
        az = az * mh
    else:
        az[0] = (-fz[2, ys, xs]) * rdsf
        az[nk - 1] = fz[nk, ys, xs] * rdsf
        az[1 : nk - 1] = (fz[2:nk, ys, xs] - fz[3 : nk + 1, ys, xs]) * rdsf
    total = (ax + ay) + az
    total = total + sv * divx
    if flat is not None:
        value = total * rr0[:, 0, 0][:, None, None]
    else:
        value = (total * rr0) * gz[None]
    if obc:
        tendency[...] = value
    else:
        tendency[:, 3 : nj - 3, 3 : ni - 3] = value[:, 3 : nj - 3, 3 : ni - 3]
