"""psolver=3 acoustic RK stage oracle (test infrastructure only).

Restates ``cuda_psolver3_terrain_open.py`` of the reference:
  * _prepare_coefficients            :300-390
  * _terrain_pressure_gradient       :393-479
  * _transformed_velocity_terms      :482-541
  * _initial_fpk                     :544-592
  * _pressure_advection              :595-675
  * _implicit_and_pressure_update    :678-745 (Thomas solve for w, pp3d/ppd update)
  * _average_velocities              :748-771
  * cpu_psolver3_terrain_open_stage_inplace :774-893
and the cadence ``sound3_stage_schedule_f32``/``rk_stage_dt_f32``
(cuda_sound3_column.py:391-424), ``terrain_coordinate_profiles``
(cuda_terrain_open_boundary.py:486-549) with Fornberg's published finite-
difference weight recursion (cm1_vertical_grid.py:387-454).

The reference's CPU stage is all-open.  For a lateral family with periodic or
rigid sides the stage composes what the reference's GPU launcher does
(cuda_psolver3_terrain_open.py:1330-1405): ssopenbc only on open sides, the
periodic boundary-face pressure gradient, and family-bound pressure fills;
those compositions are pinned by the invariants in the tests, not by
reference words.
"""

from __future__ import annotations

import math

import numpy as np

from . import boundary as bc
from .common import F, NG

CP = 1005.7
ALPH = 0.60
KDIV = 0.10
SMEPS = 1.0e-20
ZERO = F(0.0)
HALF = F(0.5)


def sound3_stage_schedule_f32(dt, nsound, nrk):
    """(nloop, dts) per RK stage (cuda_sound3_column.py:391-417)."""
    if nsound < 4 or nsound % 2:
        raise ValueError("nsound must be an even integer >= 4")
    dt32 = F(dt)
    if nrk == 1:
        nloop = int(np.floor(F(nsound) / F(3.0) + F(0.5)))
        dts = F(dt32 / F(F(nloop) * F(3.0)))
        if dts > F(dt32 / F(nsound)):
            nloop += 1
            dts = F(dt32 / F(F(nloop) * F(3.0)))
    elif nrk == 2:
        nloop = nsound // 2
        dts = F(dt32 / F(nsound))
    else:
        nloop = nsound
        dts = F(dt32 / F(nsound))
    return nloop, float(dts)


def rk_stage_dt_f32(dt, nrk):
    """dt/(4-nrk) in binary32 (cuda_sound3_column.py:420-424)."""
    return float(F(F(dt) / F(4 - nrk)))


def _fornberg_delta(alpha, x0):
    """Fornberg (1988) weight recursion, order <= 4 over 4 nodes, in binary64."""
    delta = [[[0.0 for _ in range(4)] for _ in range(6)] for _ in range(4)]
    delta[0][1][0] = 1.0
    b1 = 1.0
    for n in range(1, 4):
        b2 = 1.0
        for nu in range(n):
            b3 = alpha[n] - alpha[nu]
            b2 = b2 * b3
            delta[n - 1][n + 1][nu] = 0.0
            for mm in range(0, min(n, 4) + 1):
                delta[n][mm + 1][nu] = ((alpha[n] - x0) * delta[n - 1][mm + 1][nu] - mm * delta[n - 1][mm][nu]) / b3
        for mm in range(0, min(n, 4) + 1):
            delta[n][mm + 1][n] = (b1 / b2) * (mm * delta[n - 1][mm][n - 1] - (alpha[n - 1] - x0) * delta[n - 1][mm + 1][n - 1])
        b1 = b2
    return delta


def _extrapolation_weights(sigma, sigmaf, nk, top):
    if top:
        den = F(sigmaf[nk + 1] - sigmaf[nk])
        x0n = sigmaf[nk + 1]
        nodes = [sigma[nk], sigma[nk - 1], sigma[nk - 2], sigma[nk - 3]]
    else:
        den = F(sigmaf[2] - sigmaf[1])
        x0n = sigmaf[1]
        nodes = [sigma[1], sigma[2], sigma[3], sigma[4]]
    x0 = float(F(F(x0n) / den))
    alpha = [float(F(F(v) / den)) for v in nodes]
    d = _fornberg_delta(alpha, x0)
    return np.array([F(d[2][1][nu]) for nu in range(3)], dtype=F)


def terrain_coordinate_profiles(nk, dz, *, sigma=None, sigmaf=None):
    """sigma, sigmaf, rds, rdsf (+ cgs/cgt) with CM1 indices (cuda_terrain_open_boundary.py:486-549).

    With ``sigma``/``sigmaf`` given (length nk+2, CM1 indices) the column is
    stretched; otherwise it is uniform with layer thickness ``dz``.
    """
    dz32 = F(dz)
    s = np.zeros(nk + 2, F)
    sf = np.zeros(nk + 2, F)
    if sigma is None:
        for k in range(1, nk + 1):
            s[k] = (F(k) - F(0.5)) * dz32
        for k in range(1, nk + 2):
            sf[k] = F(k - 1) * dz32
    else:
        s[1 : nk + 1] = np.asarray(sigma, F)[1 : nk + 1]
        sf[1 : nk + 2] = np.asarray(sigmaf, F)[1 : nk + 2]
    rds = np.zeros(nk + 2, F)
    rdsf = np.zeros(nk + 2, F)
    for k in range(2, nk + 1):
        rds[k] = F(1.0) / (s[k] - s[k - 1])
    for k in range(1, nk + 1):
        rdsf[k] = F(1.0) / (sf[k + 1] - sf[k])
    return {"sigma": s, "sigmaf": sf, "rds": rds, "rdsf": rdsf,
            "cgs": _extrapolation_weights(s, sf, nk, False), "cgt": _extrapolation_weights(s, sf, nk, True)}


def prepare_coefficients(a, m, prof, *, ni, nj, nk, dts):
    """qk, pk1, pk2, mm per stage (cuda_psolver3_terrain_open.py:300-390)."""
    cp = F(CP)
    for name in ("qk", "pk1", "pk2", "mm", "fpk"):
        a[name].fill(ZERO)
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    gz = m["gz"]
    for k in range(1, nk + 1):
        tem = dts * prof["rdsf"][k]
        a["qk"][k, J, I] = (tem * a["ppterm"][k, J, I]) * gz
        tem2 = ((dts * gz) * prof["rdsf"][k]) * m["rr0"][k - 1]
        b2 = (-m["c2"][k]) * a["piadv"][k + 1, J, I] + (F(1.0) - m["c1"][k]) * a["piadv"][k, J, I]
        a["pk2"][k, J, I] = (tem2 * m["rf0"][k]) * b2
        b1 = m["c1"][k - 1] * a["piadv"][k - 1, J, I] + (m["c2"][k - 1] - F(1.0)) * a["piadv"][k, J, I]
        a["pk1"][k, J, I] = (tem2 * m["rf0"][k - 1]) * b1
        if k >= 2:
            tem1 = (dts * prof["rds"][k]) * cp
            th = m["c2"][k - 1] * a["thv"][k, J, I] + m["c1"][k - 1] * a["thv"][k - 1, J, I]
            a["mm"][k, J, I] = (tem1 * th) * gz
        if k == 1:
            a["pk1"][k, J, I] = ZERO
        if k == nk:
            a["pk2"][k, J, I] = ZERO


def terrain_pressure_gradient(a, m, prof, *, ni, nj, nk, dts, rdx, rdy):
    """u/v forward step on interior faces (cuda_psolver3_terrain_open.py:393-479)."""
    ng = NG
    ppd, thv = a["ppd"], a["thv"]
    vert = np.zeros_like(ppd)
    for k in range(2, nk + 1):
        vert[k] = (ppd[k] - ppd[k - 1]) * prof["rds"][k]
    halfcp = F(CP) * HALF
    eighth = F(0.125)
    for k in range(1, nk + 1):
        J = slice(1, nj + 1)
        fl, fr = slice(1, ni), slice(2, ni + 1)  # faces 2..ni: (face-1, face)
        thsum = thv[k, J, fr] + thv[k, J, fl]
        hz = (ppd[k, J, fr] - ppd[k, J, fl]) * rdx
        vs = (vert[k + 1, J, fr] + vert[k + 1, J, fl]) + (vert[k, J, fr] + vert[k, J, fl])
        tr = (eighth * vs) * (m["gxu"][k - 1, :, 1:ni] + m["gxu"][k, :, 1:ni])
        pg = (halfcp * thsum) * (hz + tr)
        ys, xs = slice(ng, ng + nj), slice(ng + 1, ng + ni)
        a["u3d"][k, ys, xs] = a["u3d"][k, ys, xs] + dts * (a["uten"][k, ys, xs] - pg)
        I = slice(1, ni + 1)
        gl, gr = slice(1, nj), slice(2, nj + 1)
        thsum = thv[k, gr, I] + thv[k, gl, I]
        hz = (ppd[k, gr, I] - ppd[k, gl, I]) * rdy
        vs = (vert[k + 1, gr, I] + vert[k + 1, gl, I]) + (vert[k, gr, I] + vert[k, gl, I])
        tr = (eighth * vs) * (m["gyv"][k - 1, 1:nj, :] + m["gyv"][k, 1:nj, :])
        pg = (halfcp * thsum) * (hz + tr)
        ys, xs = slice(ng + 1, ng + nj), slice(ng, ng + ni)
        a["v3d"][k, ys, xs] = a["v3d"][k, ys, xs] + dts * (a["vten"][k, ys, xs] - pg)


def _terrain_w(xu, yv, m, prof, *, ni, nj, nk, zt, weight=None):
    """w-level terrain term from face velocities (dum3 :508-540 / hadvp :628-648)."""
    out = np.zeros((nk + 2, nj, ni), F)
    rzt = F(1.0) / zt
    for kf in range(2, nk + 1):
        r2 = (prof["sigmaf"][kf] - prof["sigma"][kf - 1]) * prof["rds"][kf]
        r1 = F(1.0) - r2
        height = (prof["sigmaf"][kf] - zt) * rzt
        uh_ = xu[kf, :, :ni] + xu[kf, :, 1:]
        ul = xu[kf - 1, :, :ni] + xu[kf - 1, :, 1:]
        vh_ = yv[kf, :nj, :] + yv[kf, 1:, :]
        vl = yv[kf - 1, :nj, :] + yv[kf - 1, 1:, :]
        xt = (r2 * uh_ + r1 * ul) * m["dzdx"]
        yt = (r2 * vh_ + r1 * vl) * m["dzdy"]
        val = HALF * (xt + yt)
        if weight is None:
            out[kf] = (val * height) * m["gz"]
        else:
            out[kf] = ((val * height) * m["gz"]) * weight[kf - 1]
    return out


# Arithmetic type of the Thomas elimination and back substitution: the
# restatement's own (F) unless TRIDIAGONAL_BINARY64 is set, a checker knob that
# runs only that step in binary64 (w rounded to F when stored) to measure how
# much of a binary32 result's error the elimination contributes.
TRIDIAGONAL_BINARY64 = False


def _elimination_type():
    return np.float64 if TRIDIAGONAL_BINARY64 else F


def _eliminate_row(k, nk, aa, bb, cc, dd, ec, ed, wt, T):
    """Forward elimination of row k (k = 2..nk): pivot aa*ec(k-1)+bb; the top row solves w(nk)."""
    aa, bb, cc, dd = (x.astype(T) for x in (aa, bb, cc, dd))
    pivot = bb if k == 2 else aa * ec[k - 1] + bb
    inv = T(1.0) / pivot
    if k < nk:
        ec[k] = (-cc) * inv
        ed[k] = (dd if k == 2 else dd - aa * ed[k - 1]) * inv
    else:
        wt[k] = (dd - aa * ed[k - 1]) * inv


def _back_substitute(nk, ec, ed, wt, w, ys, xs):
    """w(k) = ec(k)*w(k+1) + ed(k) for k = nk-1..2, then w(2..nk) stored in the restatement's type."""
    for k in range(nk - 1, 1, -1):
        wt[k] = ec[k] * wt[k + 1] + ed[k]
    for k in range(2, nk + 1):
        w[k, ys, xs] = wt[k].astype(w.dtype)


def acoustic_small_step(a, m, prof, sides, *, ni, nj, nk, dts, rdx, rdy, dx, dy, zt, nrk, n, nloop,
                        alph, beta, kdiv, tridiagonal=None):
    """One small step of the stage loop (cuda_psolver3_terrain_open.py:838-863).

    ``tridiagonal``, when a dict, receives the (aa, bb, cc, dd) rows of the
    vertically implicit system (the reference's ``tri_*`` diagnostics).
    """
    ng = NG
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    bc.ssopen_terrain(a, m, prof, sides, ni=ni, nj=nj, nk=nk, dts=dts, dx=dx, dy=dy)
    terrain_pressure_gradient(a, m, prof, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy)
    bc.periodic_face_pgrad(a, m, prof, sides, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy, cp=CP)
    # transformed velocities xu, yv (:482-507)
    xu = np.zeros((nk + 2, nj, ni + 1), F)
    yv = np.zeros((nk + 2, nj + 1, ni), F)
    xu[1 : nk + 1] = a["u3d"][1 : nk + 1, ys, ng : ng + ni + 1] * m["rgzu"][None]
    yv[1 : nk + 1] = a["v3d"][1 : nk + 1, ng : ng + nj + 1, xs] * m["rgzv"][None]
    dum3 = _terrain_w(xu, yv, m, prof, ni=ni, nj=nj, nk=nk, zt=zt)
    # initial fpk (:544-592)
    smeps = F(SMEPS)
    w = a["w3d"]
    for k in range(1, nk + 1):
        hz = (xu[k, :, 1:] - xu[k, :, :ni]) * rdx + (yv[k, 1:, :] - yv[k, :nj, :]) * rdy
        vt = (dum3[k + 1] - dum3[k]) * prof["rdsf"][k]
        div = m["gz"] * (hz + vt)
        div = np.where(np.abs(div) < smeps, ZERO, div)
        w1 = ZERO if k == 1 else w[k, ys, xs]
        w2 = ZERO if k == nk else w[k + 1, ys, xs]
        src = dts * (a["ppten"][k, J, I] - a["ppterm"][k, J, I] * div)
        qp = (-a["qk"][k, J, I]) * (w2 - w1)
        a["fpk"][k, J, I] = src + beta * qp
    bc.bcwsfc(a, m, ni=ni, nj=nj, nk=nk)
    # pressure advection (:595-675)
    rho = m["rho0_3d"][:, 0, 0]
    for k in range(1, nk + 1):
        xu[k] = rho[k - 1] * xu[k]
        yv[k] = rho[k - 1] * yv[k]
    hadvp = _terrain_w(xu, yv, m, prof, ni=ni, nj=nj, nk=nk, zt=zt, weight=m["rrf0"])
    temx = (dts * HALF) * rdx
    temy = (dts * HALF) * rdy
    pia = a["piadv"]
    for k in range(1, nk + 1):
        u2 = temx * xu[k, :, 1:]
        u1 = temx * xu[k, :, :ni]
        v2 = temy * yv[k, 1:, :]
        v1 = temy * yv[k, :nj, :]
        ww2 = hadvp[k + 1] + beta * w[k + 1, ys, xs]
        ww1 = hadvp[k] + beta * w[k, ys, xs]
        xa = u2 * (pia[k, J, 2 : ni + 2] - pia[k, J, I]) + u1 * (pia[k, J, I] - pia[k, J, 0:ni])
        ya = v2 * (pia[k, 2 : nj + 2, I] - pia[k, J, I]) + v1 * (pia[k, J, I] - pia[k, 0:nj, I])
        hz = (-(xa + ya)) * m["rr0"][k - 1]
        hz = hz * m["gz"]
        vt = ww1 * a["pk1"][k, J, I] + ww2 * a["pk2"][k, J, I]
        a["fpk"][k, J, I] = a["fpk"][k, J, I] + (hz + vt)
    # implicit vertical solve and pressure update (:678-745)
    asq = alph * alph
    mm, pk1, pk2, qk, fpk, pp3d = a["mm"], a["pk1"], a["pk2"], a["qk"], a["fpk"], a["pp3d"]
    T = _elimination_type()
    ec = np.zeros((nk + 2, nj, ni), T)
    ed = np.zeros((nk + 2, nj, ni), T)
    wt = np.zeros((nk + 2, nj, ni), T)
    for k in range(2, nk + 1):
        aa = (-asq) * (mm[k, J, I] * (pk1[k - 1, J, I] + qk[k - 1, J, I])) if k >= 3 else np.zeros((nj, ni), F)
        cc = (-asq) * ((-mm[k, J, I]) * (pk2[k, J, I] - qk[k, J, I])) if k < nk else np.zeros((nj, ni), F)
        diag = (-(pk1[k, J, I] + qk[k, J, I])) + (pk2[k - 1, J, I] - qk[k - 1, J, I])
        bb = F(1.0) - asq * (mm[k, J, I] * diag)
        pdelta = pp3d[k, J, I] - pp3d[k - 1, J, I]
        fwk = dts * a["wten"][k, ys, xs] + beta * ((-mm[k, J, I]) * pdelta)
        udelta = (pp3d[k, J, I] + fpk[k, J, I]) - (pp3d[k - 1, J, I] + fpk[k - 1, J, I])
        dd = (w[k, ys, xs] + fwk) + alph * ((-mm[k, J, I]) * udelta)
        if tridiagonal is not None:
            tridiagonal.setdefault("rows", []).append((k, aa, bb, cc, dd))
        _eliminate_row(k, nk, aa, bb, cc, dd, ec, ed, wt, T)
    _back_substitute(nk, ec, ed, wt, w, ys, xs)
    for k in range(1, nk + 1):
        w1 = ZERO if k == 1 else w[k, ys, xs]
        w2 = ZERO if k == nk else w[k + 1, ys, xs]
        old = pp3d[k, J, I].copy()
        vt = w1 * (pk1[k, J, I] + qk[k, J, I]) + w2 * (pk2[k, J, I] - qk[k, J, I])
        up = (old + fpk[k, J, I]) + alph * vt
        up = np.where(np.abs(up) < smeps, ZERO, up)
        pp3d[k, J, I] = up
        tend = kdiv * (up - old)
        a["ppd"][k, J, I] = up + tend
        if nrk == 3 and n == nloop:
            a["ppx"][k, J, I] = tend
    if n < nloop:
        bc.bcp(a["ppd"], sides, ni=ni, nj=nj, nk=nk)
    # velocity averages (:748-771)
    tavg = F(1.0) / F(nloop)
    yy, xx = slice(ng, ng + nj + 1), slice(ng, ng + ni + 1)
    ks = slice(1, nk + 1)
    vu = a["rru"][ks, yy, xx] + a["u3d"][ks, yy, xx]
    vv = a["rrv"][ks, yy, xx] + a["v3d"][ks, yy, xx]
    addw = w[ks, yy, xx] if n < nloop else w[ks, yy, xx] * alph
    vw = a["rrw"][ks, yy, xx] + addw
    if n < nloop:
        a["rru"][ks, yy, xx], a["rrv"][ks, yy, xx], a["rrw"][ks, yy, xx] = vu, vv, vw
    else:
        a["rru"][ks, yy, xx], a["rrv"][ks, yy, xx], a["rrw"][ks, yy, xx] = vu * tavg, vv * tavg, vw * tavg


def cpu_psolver3_stage_inplace(
    a, m, *, ni, nj, nk, dt, nsound, nrk, dx, dy, dz, lateral_sides=(2, 2, 2, 2),
    alph=None, kdiv=None, prof=None, zt=None, radiation=True, stage_family_fills=True,
):
    """One acoustic RK stage (cuda_psolver3_terrain_open.py:774-893), family-general."""
    ng = NG
    sides = tuple(lateral_sides)
    prof = prof if prof is not None else terrain_coordinate_profiles(nk, dz)
    zt = F(nk * dz) if zt is None else F(zt)
    nloop, dts_v = sound3_stage_schedule_f32(dt, nsound, nrk)
    dts = F(dts_v)
    alph = F(ALPH if alph is None else alph)
    kdiv = F(KDIV if kdiv is None else kdiv)
    beta = F(1.0) - alph
    rdx = F(1.0) / F(dx)
    rdy = F(1.0) / F(dy)
    prepare_coefficients(a, m, prof, ni=ni, nj=nj, nk=nk, dts=dts)
    a["ppd"][1 : nk + 1] = a["ppx"][1 : nk + 1]
    if nrk != 1:
        a["pp3d"][1 : nk + 1] = a["ppi"][1 : nk + 1]
    ks = slice(1, nk + 1)
    yy, xx = slice(ng - 1, ng + nj + 1), slice(ng - 1, ng + ni + 1)
    a["rru"][ks, yy, xx] = ZERO
    a["rrv"][ks, yy, xx] = ZERO
    a["rrw"][ks, yy, xx] = a["wa"][ks, yy, xx] * beta
    if nrk != 1:
        a["u3d"][ks, yy, xx] = a["ua"][ks, yy, xx]
        a["v3d"][ks, yy, xx] = a["va"][ks, yy, xx]
        a["w3d"][ks, yy, xx] = a["wa"][ks, yy, xx]
    for n in range(1, nloop + 1):
        acoustic_small_step(a, m, prof, sides, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy, dx=dx, dy=dy,
                            zt=zt, nrk=nrk, n=n, nloop=nloop, alph=alph, beta=beta, kdiv=kdiv)
    if radiation:
        bc.radiation4(a, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dttmp=rk_stage_dt_f32(dt, nrk))
    if stage_family_fills:
        bc.fill(a["u3d"], bc.KIND_U, sides, ni=ni, nj=nj, nk=nk)
        bc.fill(a["v3d"], bc.KIND_V, sides, ni=ni, nj=nj, nk=nk)
        bc.fill(a["w3d"], bc.KIND_W, sides, ni=ni, nj=nj, nk=nk)
        bc.bcw_rigid(a["w3d"], ni=ni, nj=nj, nk=nk)
        bc.bcwsfc(a, m, ni=ni, nj=nj, nk=nk)
        if nrk < 3:
            bc.bcp(a["pp3d"], sides, ni=ni, nj=nj, nk=nk)
            if "scalar" in a:
                bc.fill(a["scalar"], bc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
    return nloop, float(dts)


# ----------------------------------------------------------------------------- flat stretched column
# sound.F with terrain_flag=.false. on a stretched column
# (cuda_gm1_child_psolver3_open.py:340-535 and _enqueue_flat_stage :1618-1792):
# the terrain stage's order less u/v reconstruction at w points and bcwsfc,
# with flat coefficients on mh/mf/rdz, the flat pressure gradient on every
# live face, the flat fpk (sound.F:491-509), the update with w(k), w(k+1) as
# they are, and the flat ssopenbc (bc.F:2410-2419 and its twins).
FLAT_SMEPS = F(1.0e-30)


def flat_prepare_coefficients(a, m, flat, *, ni, nj, nk, dts):
    """qk, pk1, pk2, mm from mh/mf/rdz and the base state at CM1's (1,1,k)."""
    cp = F(CP)
    for name in ("qk", "pk1", "pk2", "mm", "fpk"):
        a[name].fill(ZERO)
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    rdz = F(flat.rdz)
    pia, thv = a["piadv"], a["thv"]
    for k in range(1, nk + 1):
        dr = dts * rdz
        tem = dr * flat.mh[k]
        r2 = ((dr * flat.mh[k]) * m["rr0"][k - 1, 0, 0]) * m["rf0"][k, 0, 0]
        r1 = ((dr * flat.mh[k]) * m["rr0"][k - 1, 0, 0]) * m["rf0"][k - 1, 0, 0]
        tem1 = (dr * cp) * flat.mf[k]
        c1a, c1b, c2a, c2b = m["c1"][k - 1, 0, 0], m["c1"][k, 0, 0], m["c2"][k - 1, 0, 0], m["c2"][k, 0, 0]
        a["qk"][k, J, I] = tem * a["ppterm"][k, J, I]
        a["pk2"][k, J, I] = r2 * ((-c2b) * pia[k + 1, J, I] + (F(1.0) - c1b) * pia[k, J, I])
        a["pk1"][k, J, I] = r1 * (c1a * pia[k - 1, J, I] + (c2a - F(1.0)) * pia[k, J, I])
        a["mm"][k, J, I] = tem1 * (c2a * thv[k, J, I] + c1a * thv[k - 1, J, I])
        if k == 1:
            a["pk1"][k, J, I] = ZERO
        if k == nk:
            a["pk2"][k, J, I] = ZERO


def flat_ssopen(a, sides, *, ni, nj, nk, dts, rdx, rdy):
    """Open-side normal velocity: u(1) += dts*((((-radbcw*(u(2)-u(1)))*rdx)*uh(1)) + uten(1))."""
    ng = NG
    u, v = a["u3d"], a["v3d"]
    K, Y, X = slice(1, nk + 1), slice(ng, ng + nj), slice(ng, ng + ni)
    if sides[0] == 2:
        t = ((((-a["radbcw"][K, Y]) * (u[K, Y, ng + 1] - u[K, Y, ng])) * rdx) * a["uh"][ng]) + a["uten"][K, Y, ng]
        u[K, Y, ng] = u[K, Y, ng] + dts * t
    if sides[1] == 2:
        xe = ni + ng
        t = ((((-a["radbce"][K, Y]) * (u[K, Y, xe] - u[K, Y, xe - 1])) * rdx) * a["uh"][ni + ng - 1]) + a["uten"][K, Y, xe]
        u[K, Y, xe] = u[K, Y, xe] + dts * t
    if sides[2] == 2:
        t = ((((-a["radbcs"][K, X]) * (v[K, ng + 1, X] - v[K, ng, X])) * rdy) * a["vh"][ng]) + a["vten"][K, ng, X]
        v[K, ng, X] = v[K, ng, X] + dts * t
    if sides[3] == 2:
        yn = nj + ng
        t = ((((-a["radbcn"][K, X]) * (v[K, yn, X] - v[K, yn - 1, X])) * rdy) * a["vh"][nj + ng - 1]) + a["vten"][K, yn, X]
        v[K, yn, X] = v[K, yn, X] + dts * t


def flat_pressure_gradient(a, sides, *, ni, nj, nk, dts, rdx, rdy):
    """Every face i = 1..ni+1, j = 1..nj+1 with CM1's face weights (zero on open boundary faces)."""
    ng = NG
    cp = F(CP)
    tem1 = (rdx * cp) * HALF
    tem2 = (rdy * cp) * HALF
    ppd, thv = a["ppd"], a["thv"]
    pw, pe, ps, pn = (s == 1 for s in sides)
    ii = np.arange(1, ni + 2)[None, :]
    jj = np.arange(1, nj + 2)[:, None]
    wu = ((jj <= nj) & ((ii > 1) | pw) & ((ii < ni + 1) | pe)).astype(F)
    wv = ((ii <= ni) & ((jj > 1) | ps) & ((jj < nj + 1) | pn)).astype(F)
    Jp, Ip = slice(1, nj + 2), slice(1, ni + 2)
    for k in range(1, nk + 1):
        pu = tem1 * (ppd[k, Jp, Ip] - ppd[k, Jp, 0 : ni + 1])
        pu = pu * (thv[k, Jp, Ip] + thv[k, Jp, 0 : ni + 1])
        sl = (k, slice(ng, ng + nj + 1), slice(ng, ng + ni + 1))
        a["u3d"][sl] = a["u3d"][sl] + (dts * wu) * (a["uten"][sl] - pu)
        pv = tem2 * (ppd[k, Jp, Ip] - ppd[k, 0 : nj + 1, Ip])
        pv = pv * (thv[k, Jp, Ip] + thv[k, 0 : nj + 1, Ip])
        a["v3d"][sl] = a["v3d"][sl] + (dts * wv) * (a["vten"][sl] - pv)


def flat_small_step(a, m, flat, sides, *, ni, nj, nk, dts, rdx, rdy, nrk, n, nloop, alph, beta, kdiv,
                    tridiagonal=None):
    """One small step of the flat column; ``tridiagonal`` (a dict) receives the (aa, bb, cc, dd) rows."""
    ng = NG
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    flat_ssopen(a, sides, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy)
    flat_pressure_gradient(a, sides, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy)
    u, v, w, pia = a["u3d"], a["v3d"], a["w3d"], a["piadv"]
    temx = (dts * HALF) * rdx
    temy = (dts * HALF) * rdy
    P = lambda name, k: a[name][k, J, I]  # noqa: E731
    for k in range(1, nk + 1):
        div = (u[k, ys, ng + 1 : ng + ni + 1] - u[k, ys, xs]) * rdx + (v[k, ng + 1 : ng + nj + 1, xs] - v[k, ys, xs]) * rdy
        div = np.where(np.abs(div) < FLAT_SMEPS, ZERO, div)
        pc = pia[k, J, I]
        hx = u[k, ys, xs] * (pc - pia[k, J, 0:ni]) + u[k, ys, ng + 1 : ng + ni + 1] * (pia[k, J, 2 : ni + 2] - pc)
        hx = (-hx) * temx
        hy = v[k, ys, xs] * (pc - pia[k, 0:nj, I]) + v[k, ng + 1 : ng + nj + 1, xs] * (pia[k, 2 : nj + 2, I] - pc)
        hy = (-hy) * temy
        hp = hx + hy
        vert = w[k, ys, xs] * (P("pk1", k) + P("qk", k)) + w[k + 1, ys, xs] * (P("pk2", k) - P("qk", k))
        a["fpk"][k, J, I] = (dts * (P("ppten", k) - P("ppterm", k) * div) + hp) + beta * vert
    asq = alph * alph
    mm, pk1, pk2, qk, fpk, pp3d = a["mm"], a["pk1"], a["pk2"], a["qk"], a["fpk"], a["pp3d"]
    T = _elimination_type()
    ec = np.zeros((nk + 2, nj, ni), T)
    ed = np.zeros((nk + 2, nj, ni), T)
    wt = np.zeros((nk + 2, nj, ni), T)
    for k in range(2, nk + 1):
        zero = np.zeros((nj, ni), F)
        aa = (-asq) * (mm[k, J, I] * (pk1[k - 1, J, I] + qk[k - 1, J, I])) if k >= 3 else zero
        cc = (-asq) * ((-mm[k, J, I]) * (pk2[k, J, I] - qk[k, J, I])) if k < nk else zero
        diag = (-(pk1[k, J, I] + qk[k, J, I])) + (pk2[k - 1, J, I] - qk[k - 1, J, I])
        bb = F(1.0) - asq * (mm[k, J, I] * diag)
        pdelta = pp3d[k, J, I] - pp3d[k - 1, J, I]
        fwk = dts * a["wten"][k, ys, xs] + beta * ((-mm[k, J, I]) * pdelta)
        udelta = (pp3d[k, J, I] + fpk[k, J, I]) - (pp3d[k - 1, J, I] + fpk[k - 1, J, I])
        dd = (w[k, ys, xs] + fwk) + alph * ((-mm[k, J, I]) * udelta)
        if tridiagonal is not None:
            tridiagonal.setdefault("rows", []).append((k, aa, bb, cc, dd))
        _eliminate_row(k, nk, aa, bb, cc, dd, ec, ed, wt, T)
    _back_substitute(nk, ec, ed, wt, w, ys, xs)
    for k in range(1, nk + 1):
        old = pp3d[k, J, I].copy()
        vt = w[k, ys, xs] * (pk1[k, J, I] + qk[k, J, I]) + w[k + 1, ys, xs] * (pk2[k, J, I] - qk[k, J, I])
        up = (old + fpk[k, J, I]) + alph * vt
        up = np.where(np.abs(up) < FLAT_SMEPS, ZERO, up)
        pp3d[k, J, I] = up
        tend = kdiv * (up - old)
        a["ppd"][k, J, I] = up + tend
        if nrk == 3 and n == nloop:
            a["ppx"][k, J, I] = tend
    if n < nloop:
        bc.bcp(a["ppd"], sides, ni=ni, nj=nj, nk=nk)
    tavg = F(1.0) / F(nloop)
    yy, xx = slice(ng, ng + nj + 1), slice(ng, ng + ni + 1)
    ks = slice(1, nk + 1)
    vu = a["rru"][ks, yy, xx] + a["u3d"][ks, yy, xx]
    vv = a["rrv"][ks, yy, xx] + a["v3d"][ks, yy, xx]
    addw = w[ks, yy, xx] if n < nloop else w[ks, yy, xx] * alph
    vw = a["rrw"][ks, yy, xx] + addw
    if n < nloop:
        a["rru"][ks, yy, xx], a["rrv"][ks, yy, xx], a["rrw"][ks, yy, xx] = vu, vv, vw
    else:
        a["rru"][ks, yy, xx], a["rrv"][ks, yy, xx], a["rrw"][ks, yy, xx] = vu * tavg, vv * tavg, vw * tavg


def cpu_flat_stage_inplace(a, m, flat, *, ni, nj, nk, dt, nsound, nrk, dx, dy, lateral_sides=(2, 2, 2, 2),
                           alph=None, kdiv=None, radiation=True, stage_family_fills=True):
    """One acoustic RK stage on a flat stretched column (``flat``: mh, mf, rdz)."""
    ng = NG
    sides = tuple(lateral_sides)
    nloop, dts_v = sound3_stage_schedule_f32(dt, nsound, nrk)
    dts = F(dts_v)
    alph = F(ALPH if alph is None else alph)
    kdiv = F(KDIV if kdiv is None else kdiv)
    beta = F(1.0) - alph
    rdx = F(1.0) / F(dx)
    rdy = F(1.0) / F(dy)
    flat_prepare_coefficients(a, m, flat, ni=ni, nj=nj, nk=nk, dts=dts)
    a["ppd"][1 : nk + 1] = a["ppx"][1 : nk + 1]
    if nrk != 1:
        a["pp3d"][1 : nk + 1] = a["ppi"][1 : nk + 1]
    ks = slice(1, nk + 1)
    yy, xx = slice(ng - 1, ng + nj + 1), slice(ng - 1, ng + ni + 1)
    a["rru"][ks, yy, xx] = ZERO
    a["rrv"][ks, yy, xx] = ZERO
    a["rrw"][ks, yy, xx] = a["wa"][ks, yy, xx] * beta
    if nrk != 1:
        a["u3d"][ks, yy, xx] = a["ua"][ks, yy, xx]
        a["v3d"][ks, yy, xx] = a["va"][ks, yy, xx]
        a["w3d"][ks, yy, xx] = a["wa"][ks, yy, xx]
    for n in range(1, nloop + 1):
        flat_small_step(a, m, flat, sides, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy, nrk=nrk, n=n,
                        nloop=nloop, alph=alph, beta=beta, kdiv=kdiv)
    if radiation:
        bc.radiation4(a, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dttmp=rk_stage_dt_f32(dt, nrk))
    if stage_family_fills:
        bc.fill(a["u3d"], bc.KIND_U, sides, ni=ni, nj=nj, nk=nk)
        bc.fill(a["v3d"], bc.KIND_V, sides, ni=ni, nj=nj, nk=nk)
        bc.fill(a["w3d"], bc.KIND_W, sides, ni=ni, nj=nj, nk=nk)
        bc.bcw_rigid(a["w3d"], ni=ni, nj=nj, nk=nk)
        if nrk < 3:
            bc.bcp(a["pp3d"], sides, ni=ni, nj=nj, nk=nk)
            if "scalar" in a:
                bc.fill(a["scalar"], bc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)
    return nloop, float(dts)
