"""Stage-term, buoyancy, post-acoustic, scalar-RK and solve3 oracle (test infrastructure).

Restates in binary32 NumPy:
  * eqtset=2 stage terms       -> cuda_moist_stage_bridge.py:62-128 (cpu_eqtset2_cell_f32),
                                  vectorised; piadv and the halo fills follow the
                                  kernel text ``cm1_msb_terms_v1``/``cm1_msb_piadv_v1``
                                  and ``launch_cuda_eqtset2_stage_terms`` (:153-165, :378-425)
  * buoyancy on w/u/v faces    -> kernel text ``cm1_msb_buoyancy_v1`` (:177-183);
                                  no CPU form in the reference (parity unpinned against
                                  reference words; pinned by invariants)
  * getdiv / getdivx           -> kernel text ``cm1_msb_getdiv_v1``,
                                  ``cm1_msb_getdivx_{horizontal,vertical,divergence}_v1``
                                  (:185-194); no CPU form (parity unpinned)
  * theta packing              -> cuda_scalar_moist_rk.py:255-280
  * A-anchored compose         -> cuda_scalar_moist_rk.py:283-349
  * held momentum seed         -> cuda_momentum_rk.py:242-271
  * held compose / ppten       -> kernel text ``gm1_predict_and_compose_held_v1``
                                  (cuda_gm1_forecast.py:555-566; parity unpinned)
  * solve3                     -> cuda_gm1_forecast.py:1414-1457 and ``gm1_solve3_density_v1``
                                  (:568-577; parity unpinned)
"""

from __future__ import annotations

import numpy as np

from . import boundary as bc
from .common import F, NG

RD, RV, CP, CPV = F(287.04), F(461.5), F(1005.7), F(1870.0)
CPL, CPI, G = F(4190.0), F(2106.0), F(9.81)
ZERO = F(0.0)
HALF = F(0.5)
TH0R = F(300.0)
SMEPS_RK = F(1.0e-30)
SMEPS = F(1.0e-20)


def stage_terms(a, base, th3d, q3d, *, ni, nj, nk, sides):
    """thv, ppterm, piadv (pressure family), dum8 (padded scalar), thterm (cells).

    ``base`` holds profiles th0, qv0, pi0 (nk,) and optionally qc0, qi0;
    ``q3d`` is (6, nk, nj, ni) (qv, qc, qr, qi, qs, qg) or None for a dry run.
    """
    cv, cvv = CP - RD, CPV - RV
    rovcp, reps = RD / CP, RV / RD
    repsm1 = reps - F(1.0)
    shape = (nk, nj, ni)
    if q3d is None:
        q = np.zeros((6, nk, nj, ni), F)
    else:
        q = q3d
    qv = np.maximum(q[0], ZERO)
    liq = q[1] + q[2]
    ice = (q[3] + q[4]) + q[5]
    qli = np.maximum(liq + ice, ZERO)
    cpli = CPL * np.maximum(liq, ZERO) + CPI * np.maximum(ice, ZERO)
    th0 = base["th0"][:, None, None]
    qv0 = base["qv0"][:, None, None]
    pi0 = base["pi0"][:, None, None]
    qc0 = base.get("qc0", np.zeros(nk, F))[:, None, None]
    qi0 = base.get("qi0", np.zeros(nk, F))[:, None, None]
    rth = F(1.0) / th0
    theta = th0 + th3d
    b = th3d * rth + repsm1 * (qv - qv0)
    b = b - ((qli - qc0) - qi0)
    a["dum8"][1 : nk + 1, NG : NG + nj, NG : NG + ni] = G * b
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    a["thv"][1 : nk + 1, J, I] = (theta * (F(1.0) + reps * qv)) / ((F(1.0) + qv) + qli)
    cpm = (CP + CPV * qv) + cpli
    cvm = F(1.0) / ((cv + cvv * qv) + cpli)
    a["thterm"][...] = (theta * ((RD + RV * qv) - rovcp * cpm)) * cvm
    pi = pi0 + a["pp3d"][1 : nk + 1, J, I]
    a["ppterm"][1 : nk + 1, J, I] = ((pi * rovcp) * cpm) * cvm
    # piadv over the whole pressure family (k = 1..nk), base at the clamped column
    a["piadv"][1 : nk + 1] = base["pi0"][:, None, None] + a["pp3d"][1 : nk + 1]
    bc.bcp(a["thv"], sides, ni=ni, nj=nj, nk=nk)
    bc.fill(a["dum8"], bc.KIND_SCALAR, sides, ni=ni, nj=nj, nk=nk)


def buoyancy(a, m, *, ni, nj, nk):
    """wten/uten/vten += buoyancy terms (kernel text cm1_msb_buoyancy_v1)."""
    d8 = a["dum8"]
    ys, xs = slice(NG, NG + nj), slice(NG, NG + ni)
    c1, c2 = m["c1"], m["c2"]
    for kf in range(2, nk + 1):
        add = c1[kf - 1] * d8[kf - 1, ys, xs] + c2[kf - 1] * d8[kf, ys, xs]
        a["wten"][kf, ys, xs] = a["wten"][kf, ys, xs] + add

    def bf(kf):  # (nj, ni) buoyancy at w level kf, zero at the ends
        if kf == 1 or kf == nk + 1:
            return np.zeros((nj, ni), F)
        return c1[kf - 1] * d8[kf - 1, ys, xs] + c2[kf - 1] * d8[kf, ys, xs]

    for k in range(1, nk + 1):
        hi, lo = bf(k + 1), bf(k)
        d = d8[k, ys, xs]
        # u faces f = 1..ni-1 (cells f-1, f)
        t = (hi[:, 1:] + hi[:, :-1]) + (lo[:, 1:] + lo[:, :-1])
        t = F(0.125) * t
        t = t - F(0.25) * (d[:, 1:] + d[:, :-1])
        t = t * (m["gxu"][k - 1, :, 1:ni] + m["gxu"][k, :, 1:ni])
        a["uten"][k, ys, NG + 1 : NG + ni] = a["uten"][k, ys, NG + 1 : NG + ni] + t
        # v faces f = 1..nj-1
        t = (hi[1:, :] + hi[:-1, :]) + (lo[1:, :] + lo[:-1, :])
        t = F(0.125) * t
        t = t - F(0.25) * (d[1:, :] + d[:-1, :])
        t = t * (m["gyv"][k - 1, 1:nj, :] + m["gyv"][k, 1:nj, :])
        a["vten"][k, NG + 1 : NG + nj, xs] = a["vten"][k, NG + 1 : NG + nj, xs] + t


def _uniform_weights(kf, dz):
    dz32 = F(dz)
    rds = F(1.0) / dz32
    sf = F(kf - 1) * dz32
    sm = (F(kf - 1) - F(0.5)) * dz32
    r2 = (sf - sm) * rds
    r1 = HALF * (F(1.0) - r2)
    return sf, HALF * r2, r1


def getdiv(a, m, div, theta_div, thterm, *, ni, nj, nk, dx, dy, dz):
    """Divergence of the stage-averaged winds (kernel text cm1_msb_getdiv_v1)."""
    ys, xs = slice(NG, NG + nj), slice(NG, NG + ni)
    rdx, rdy, rdsf = F(1.0) / F(dx), F(1.0) / F(dy), F(1.0) / F(dz)
    zt = F(nk) * F(dz)
    rzt = F(1.0) / zt
    u, v, w = a["rru"], a["rrv"], a["rrw"]
    xu = u[:, ys, NG : NG + ni + 1] * m["rgzu"][None]
    yv = v[:, NG : NG + nj + 1, xs] * m["rgzv"][None]
    wt = np.zeros((nk + 2, nj, ni), F)
    for kf in range(2, nk + 1):
        sf, r2, r1 = _uniform_weights(kf, dz)
        ux = r2 * (xu[kf, :, :ni] + xu[kf, :, 1:]) + r1 * (xu[kf - 1, :, :ni] + xu[kf - 1, :, 1:])
        vy = r2 * (yv[kf, :nj] + yv[kf, 1:]) + r1 * (yv[kf - 1, :nj] + yv[kf - 1, 1:])
        t = ux * m["dzdx"] + vy * m["dzdy"]
        t = ((t * (sf - zt)) * m["gz"]) * rzt
        wt[kf] = w[kf, ys, xs] + t
    for k in range(1, nk + 1):
        d = (xu[k, :, 1:] - xu[k, :, :ni]) * rdx + (yv[k, 1:] - yv[k, :nj]) * rdy
        d = d + (wt[k + 1] - wt[k]) * rdsf
        d = np.where(np.abs(d) < SMEPS, ZERO, d)
        div[k - 1] = d
        theta_div[k - 1] = d * thterm[k - 1]


def getdivx(a, m, divx, *, ni, nj, nk, dx, dy, dz):
    """Rebuild rru/rrv/rrw as mass fluxes in place, then divx (cm1_msb_getdivx_*)."""
    rho0 = m["rho0_3d"]
    ks = slice(1, nk + 1)
    f = np.arange(ni + 1)
    l, r = np.maximum(f - 1, 0), np.minimum(f, ni - 1)
    den = HALF * (rho0[:, :, l] + rho0[:, :, r])
    ys = slice(NG, NG + nj)
    a["rru"][ks, ys, NG : NG + ni + 1] = (den * a["rru"][ks, ys, NG : NG + ni + 1]) * m["rgzu"][None]
    g = np.arange(nj + 1)
    s, n = np.maximum(g - 1, 0), np.minimum(g, nj - 1)
    den = HALF * (rho0[:, s, :] + rho0[:, n, :])
    xs = slice(NG, NG + ni)
    a["rrv"][ks, NG : NG + nj + 1, xs] = (den * a["rrv"][ks, NG : NG + nj + 1, xs]) * m["rgzv"][None]
    zt = F(nk) * F(dz)
    rzt = F(1.0) / zt
    rru, rrv = a["rru"], a["rrv"]
    a["rrw"][1, ys, xs] = ZERO
    a["rrw"][nk + 1, ys, xs] = ZERO
    for kf in range(2, nk + 1):
        sf, r2, r1 = _uniform_weights(kf, dz)
        ux = r2 * (rru[kf, ys, xs] + rru[kf, ys, NG + 1 : NG + ni + 1]) + r1 * (rru[kf - 1, ys, xs] + rru[kf - 1, ys, NG + 1 : NG + ni + 1])
        vy = r2 * (rrv[kf, ys, xs] + rrv[kf, NG + 1 : NG + nj + 1, xs]) + r1 * (rrv[kf - 1, ys, xs] + rrv[kf - 1, NG + 1 : NG + nj + 1, xs])
        t = ux * m["dzdx"] + vy * m["dzdy"]
        t = ((t * (sf - zt)) * m["gz"]) * rzt
        a["rrw"][kf, ys, xs] = m["rf0"][kf - 1] * a["rrw"][kf, ys, xs] + t
    rdx, rdy, rdsf = F(1.0) / F(dx), F(1.0) / F(dy), F(1.0) / F(dz)
    rrw = a["rrw"]
    d = (rru[ks, ys, NG + 1 : NG + ni + 1] - rru[ks, ys, xs]) * rdx + (rrv[ks, NG + 1 : NG + nj + 1, xs] - rrv[ks, ys, xs]) * rdy
    d = d + (rrw[2 : nk + 2, ys, xs] - rrw[1 : nk + 1, ys, xs]) * rdsf
    divx[...] = np.where(np.abs(d) < SMEPS, ZERO, d)


def pack_theta(th0, th3d, padded, *, ni, nj, nk):
    """(th0 - 300) + th3d into the padded scalar (cuda_scalar_moist_rk.py:255-280)."""
    padded[1 : nk + 1, NG : NG + nj, NG : NG + ni] = (th0[:, None, None] - TH0R) + th3d


def pack_scalar(q, padded, *, ni, nj, nk):
    padded[1 : nk + 1, NG : NG + nj, NG : NG + ni] = q


def compose_theta(tha, held, divterm, adv, dttmp):
    """A-anchored update with the smeps clamp (cuda_scalar_moist_rk.py:336-341)."""
    seed = held - divterm
    out = tha + F(dttmp) * (seed + adv)
    return np.where(np.abs(out) < SMEPS_RK, ZERO, out)


def compose_moisture(qa, held, adv, dttmp):
    out = qa + F(dttmp) * (held + adv)
    return np.where(np.abs(out) < SMEPS_RK, ZERO, out)


def seed_held(held, tend, *, ni, nj, nk):
    """Copy held u/v/w tendencies over i=1:ni+1, j=1:nj+1, k=1:nk (cuda_momentum_rk.py:242-271)."""
    region = (slice(1, nk + 1), slice(NG, NG + nj + 1), slice(NG, NG + ni + 1))
    for h, t in zip(held, tend):
        t[region] = h[region]


def predict_and_compose_held(base, rho, ppi, tha, qa, thheld, qheld, ppten, rates, dt, *, ni, nj, nk):
    """ppten from held theta/qv, then the Kessler held rates join (gm1_predict_and_compose_held_v1).

    The ``powf`` branch runs only when the closure left a held theta or qv
    tendency; with no closure it is never taken.
    """
    rd, rv = F(287.04), F(461.5)
    rp00 = F(1.0) / F(100000.0)
    rddcv = F(287.04) / (F(1005.7) - F(287.04))
    tem = F(0.0001) * F(0.0001)
    qsmall = F(1.0e-14)
    dt = F(dt)
    ht = thheld.copy()
    qvt = qheld[0].copy()
    cond = (np.abs(dt * ht) > tem) | (np.abs(dt * qvt) > qsmall)
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    out = rates["qpten"].copy()
    if cond.any():
        thnew = tha + dt * ht
        qvnew = qa[0] + dt * qvt
        b = (rho * (base["th0"][:, None, None] + thnew)) * (rd + rv * qvnew)
        b = b * rp00
        pinew = np.power(b, rddcv).astype(F) - base["pi0"][:, None, None]
        val = ((pinew - ppi) / dt) + rates["qpten"]
        out = np.where(cond, val, out)
    ppten[1 : nk + 1, J, I] = out
    thheld[...] = ht + rates["qtten"]
    qheld[0] = qvt + rates["qvten"]
    qheld[1] = qheld[1] + rates["qcten"]


def solve3_pressure(a, sides, *, ni, nj, nk):
    """ppx <- pp3d + ppx, ppi <- pp3d, and their halos (cuda_gm1_forecast.py:1417-1435)."""
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    ks = slice(1, nk + 1)
    a["ppx"][ks, J, I] = a["pp3d"][ks, J, I] + a["ppx"][ks, J, I]
    a["ppi"][ks, J, I] = a["pp3d"][ks, J, I]
    for name in ("ppx", "pp3d", "ppi"):
        bc.bcp(a[name], sides, ni=ni, nj=nj, nk=nk)


def solve3_density(rho, c1, c2):
    """rr = 1/rho, rf by c1/c2 interpolation with 3-point end extrapolation (gm1_solve3_density_v1)."""
    nk = rho.shape[0]
    rr = F(1.0) / rho
    rf = np.empty((nk + 1,) + rho.shape[1:], F)
    rf[0] = (F(1.875) * rho[0] + F(-1.25) * rho[1]) + F(0.375) * rho[2]
    rf[nk] = (F(1.875) * rho[nk - 1] + F(-1.25) * rho[nk - 2]) + F(0.375) * rho[nk - 3]
    for f in range(1, nk):
        rf[f] = c1[f] * rho[f - 1] + c2[f] * rho[f]
    return rr, rf


# ----------------------------------------------------------------------------- flat stretched column
# The terrain_flag=.false. texts of cuda_gm1_child_stretched_moist.py:155-190
# (misclibs.F:127-138, 241-270, 324-336; solve2.F:897-906).  ``flat`` carries
# the column's mh (nk+2), mf (nk+3) and rdz = 1/dz of the metric reference;
# the base state is read at the first physical column, CM1's (1,1,k).
FLAT_SMEPS = F(1.0e-30)


def getdiv_flat(a, div, theta_div, thterm, flat, *, ni, nj, nk, dx, dy):
    """((du*rdx) + (dv*rdy)) + ((dw*rdz)*mh(k)) of the stage averages, theta_div = div*thterm."""
    ys, xs = slice(NG, NG + nj), slice(NG, NG + ni)
    rdx, rdy, rdz = F(1.0) / F(dx), F(1.0) / F(dy), F(flat.rdz)
    u, v, w = a["rru"], a["rrv"], a["rrw"]
    for k in range(1, nk + 1):
        d = (u[k, ys, NG + 1 : NG + ni + 1] - u[k, ys, xs]) * rdx + (v[k, NG + 1 : NG + nj + 1, xs] - v[k, ys, xs]) * rdy
        d = d + ((w[k + 1, ys, xs] - w[k, ys, xs]) * rdz) * flat.mh[k]
        d = np.where(np.abs(d) < FLAT_SMEPS, ZERO, d)
        div[k - 1] = d
        theta_div[k - 1] = d * thterm[k - 1]


def getdivx_flat(a, m, divx, flat, *, ni, nj, nk, dx, dy):
    """rru, rrv <- (avg)*rho0(1,1,k), rrw <- (avg)*rf0(1,1,k) (0 at both lids), then divx."""
    ks = slice(1, nk + 1)
    ys, xs = slice(NG, NG + nj), slice(NG, NG + ni)
    rho_c = m["rho0_3d"][:, 0, 0][:, None, None]
    a["rru"][ks, ys, NG : NG + ni + 1] = a["rru"][ks, ys, NG : NG + ni + 1] * rho_c
    a["rrv"][ks, NG : NG + nj + 1, xs] = a["rrv"][ks, NG : NG + nj + 1, xs] * rho_c
    a["rrw"][1, ys, xs] = ZERO
    a["rrw"][nk + 1, ys, xs] = ZERO
    for kf in range(2, nk + 1):
        a["rrw"][kf, ys, xs] = a["rrw"][kf, ys, xs] * m["rf0"][kf - 1, 0, 0]
    rdx, rdy, rdz = F(1.0) / F(dx), F(1.0) / F(dy), F(flat.rdz)
    rru, rrv, rrw = a["rru"], a["rrv"], a["rrw"]
    d = (rru[ks, ys, NG + 1 : NG + ni + 1] - rru[ks, ys, xs]) * rdx + (rrv[ks, NG + 1 : NG + nj + 1, xs] - rrv[ks, ys, xs]) * rdy
    d = d + ((rrw[2 : nk + 2, ys, xs] - rrw[1 : nk + 1, ys, xs]) * rdz) * flat.mh[1 : nk + 1][:, None, None]
    divx[...] = np.where(np.abs(d) < FLAT_SMEPS, ZERO, d)


def buoyancy_flat(a, m, *, ni, nj, nk):
    """wten(k) += c1(k)*dum8(k-1) + c2(k)*dum8(k) for k = 2..nk; no horizontal term on a flat column."""
    d8 = a["dum8"]
    ys, xs = slice(NG, NG + nj), slice(NG, NG + ni)
    for kf in range(2, nk + 1):
        add = m["c1"][kf - 1] * d8[kf - 1, ys, xs] + m["c2"][kf - 1] * d8[kf, ys, xs]
        a["wten"][kf, ys, xs] = a["wten"][kf, ys, xs] + add
