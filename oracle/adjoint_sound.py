"""Adjoint of the acoustic RK stage on the C5 geometry, in binary64 (test infrastructure).

The stage is ``sound.cpu_psolver3_stage_inplace`` (cuda_psolver3_terrain_open.py
:774-893) on flat terrain metrics (gz = rgzu = rgzv = 1, dzdx = dzdy = gx* =
gy* = 0) with periodic laterals (no ssopen, no radiation).  For fixed
coefficients (thv, ppterm, piadv and through them qk, pk1, pk2, mm) the stage
is linear in its state and forcing, so its adjoint is the transpose of that
linear map, taken statement by statement in reverse:

  stage end   fills of u, v, w (periodic family copies), bcw_rigid, bcwsfc,
              bcp(pp3d) at nrk < 3
  small step  averages rru/rrv/rrw  <-  bcp(ppd)  <-  pressure update and
              filter  <-  Thomas solve (its transpose: a tridiagonal solve with
              the off-diagonals exchanged)  <-  pressure advection  <-
              bcwsfc  <-  fpk  <-  pressure gradient on faces 0..n
  stage start ppd = ppx, pp3d = ppi, rru = rrv = 0, rrw = beta*wa, u/v/w = A

The smeps clamps (|div| < 1e-20, |pp'| < 1e-20 -> 0) are taken as the
identity.  ``stage_adjoint`` maps output adjoints to input adjoints; it is
pinned by the dot-product identity against the binary64 restatement of the
forward stage (tests/test_adjoint_sound.py) and is the binary64 reference of
the device adjoint.
"""

from __future__ import annotations

import numpy as np

from .common import F, NG
from .sound import ALPH, CP, KDIV, sound3_stage_schedule_f32, terrain_coordinate_profiles

D = np.float64
STATE = ("u3d", "v3d", "w3d", "pp3d", "ppd", "ppx", "rru", "rrv", "rrw")


def _coefficients(coef, m, prof, *, ni, nj, nk, dts):
    """qk, pk1, pk2, mm of prepare_coefficients in binary64 (flat metrics)."""
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    thv, ppterm, pia = (coef[n].astype(D) for n in ("thv", "ppterm", "piadv"))
    qk = np.zeros((nk + 2, nj, ni), D)
    pk1 = np.zeros_like(qk)
    pk2 = np.zeros_like(qk)
    mm = np.zeros_like(qk)
    c1, c2, rr0, rf0 = (m[n].astype(D) for n in ("c1", "c2", "rr0", "rf0"))
    rdsf, rds = prof["rdsf"].astype(D), prof["rds"].astype(D)
    for k in range(1, nk + 1):
        qk[k] = (dts * rdsf[k]) * ppterm[k, J, I]
        tem2 = (dts * rdsf[k]) * rr0[k - 1]
        if k < nk:
            pk2[k] = (tem2 * rf0[k]) * ((-c2[k]) * pia[k + 1, J, I] + (1.0 - c1[k]) * pia[k, J, I])
        if k > 1:
            pk1[k] = (tem2 * rf0[k - 1]) * (c1[k - 1] * pia[k - 1, J, I] + (c2[k - 1] - 1.0) * pia[k, J, I])
            mm[k] = ((dts * rds[k]) * float(F(CP))) * (c2[k - 1] * thv[k, J, I] + c1[k - 1] * thv[k - 1, J, I])
    return qk, pk1, pk2, mm


def _bcp_adjoint(pb, *, ni, nj, nk):
    """Adjoint of bcp on periodic sides: each halo word's adjoint joins its source, then clears."""
    ks = slice(1, nk + 1)
    # forward order: west, east, south, north (the south/north copies read filled west/east columns)
    pb[ks, 1, :] += pb[ks, nj + 1, :]
    pb[ks, nj + 1, :] = 0.0
    pb[ks, nj, :] += pb[ks, 0, :]
    pb[ks, 0, :] = 0.0
    pb[ks, :, 1] += pb[ks, :, ni + 1]
    pb[ks, :, ni + 1] = 0.0
    pb[ks, :, ni] += pb[ks, :, 0]
    pb[ks, :, 0] = 0.0


def _family_adjoint(ab, kind, *, ni, nj, nk):
    """Adjoint of the periodic family fill (oracle boundary.fill, all sides periodic)."""
    ng = NG
    kfirst, kcount = (2, nk - 1) if kind == 3 else (1, nk)
    ks = slice(kfirst, kfirst + kcount)
    stag_x = kind == 1
    stag_y = kind == 2

    def we():
        # forward: for layer 0..2: west dst ng-1-layer <- src ni+ng-1-layer; east dst (stag: ni+ng+1+layer <- ng+1+layer,
        # else ni+ng+layer <- ng+layer); all rows.  Reverse the copies in reverse order.
        for layer in reversed(range(ng)):
            if stag_x:
                dst, src = ni + ng + 1 + layer, ng + 1 + layer
            else:
                dst, src = ni + ng + layer, ng + layer
            ab[ks, :, src] += ab[ks, :, dst]
            ab[ks, :, dst] = 0.0
            dst, src = ng - 1 - layer, ni + ng - 1 - layer
            ab[ks, :, src] += ab[ks, :, dst]
            ab[ks, :, dst] = 0.0

    def sn():
        for layer in reversed(range(ng)):
            if stag_y:
                dst, src = nj + ng + 1 + layer, ng + 1 + layer
            else:
                dst, src = nj + ng + layer, ng + layer
            ab[ks, src, :] += ab[ks, dst, :]
            ab[ks, dst, :] = 0.0
            dst, src = ng - 1 - layer, nj + ng - 1 - layer
            ab[ks, src, :] += ab[ks, dst, :]
            ab[ks, dst, :] = 0.0

    # forward order: bcv is south/north then west/east; the others west/east then south/north
    if kind == 2:
        we()
        sn()
    else:
        sn()
        we()


def _thomas_transpose(aa, bb, cc, rhs):
    """Solve T^T x = rhs for the rows k = 2..nk of T (aa below, bb on, cc above the diagonal)."""
    # T^T has aa(k+1) above and cc(k-1) below the diagonal of row k
    nrow = len(bb)
    lo = [np.zeros_like(bb[0])] + [cc[r - 1] for r in range(1, nrow)]
    up = [aa[r + 1] for r in range(nrow - 1)] + [np.zeros_like(bb[0])]
    cp = [None] * nrow
    dp = [None] * nrow
    for r in range(nrow):
        piv = bb[r] if r == 0 else bb[r] - lo[r] * cp[r - 1]
        cp[r] = up[r] / piv
        dp[r] = (rhs[r] if r == 0 else rhs[r] - lo[r] * dp[r - 1]) / piv
    x = [None] * nrow
    x[-1] = dp[-1]
    for r in range(nrow - 2, -1, -1):
        x[r] = dp[r] - cp[r] * x[r + 1]
    return x


def _small_step_adjoint(b, coef, m, qk, pk1, pk2, mm, *, ni, nj, nk, dts, rdx, rdy, nrk, n, nloop, alph,
                        beta, kdiv):
    ng = NG
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    ks = slice(1, nk + 1)
    last = n == nloop
    tavg = float(F(1.0) / F(nloop))
    thv, ppterm, pia = (coef[k].astype(D) for k in ("thv", "ppterm", "piadv"))
    rho = m["rho0_3d"][:, 0, 0].astype(D)
    rr0 = m["rr0"].astype(D)
    halfcp = float(F(CP) * F(0.5))
    temx, temy = float((F(dts) * F(0.5)) * F(rdx)), float((F(dts) * F(0.5)) * F(rdy))
    ub, vb, wb, pb, pdb = b["u3d"], b["v3d"], b["w3d"], b["pp3d"], b["ppd"]
    # A8: time averages over yy = ng..ng+nj, xx = ng..ng+ni
    yy, xx = slice(ng, ng + nj + 1), slice(ng, ng + ni + 1)
    if last:
        b["rru"][ks, yy, xx] *= tavg
        b["rrv"][ks, yy, xx] *= tavg
        b["rrw"][ks, yy, xx] *= tavg
        ub[ks, yy, xx] += b["rru"][ks, yy, xx]
        vb[ks, yy, xx] += b["rrv"][ks, yy, xx]
        wb[ks, yy, xx] += alph * b["rrw"][ks, yy, xx]
    else:
        ub[ks, yy, xx] += b["rru"][ks, yy, xx]
        vb[ks, yy, xx] += b["rrv"][ks, yy, xx]
        wb[ks, yy, xx] += b["rrw"][ks, yy, xx]
    # A7: the ppd halo fill between small steps
    if n < nloop:
        _bcp_adjoint(pdb, ni=ni, nj=nj, nk=nk)
    # A6: pressure update, filter, ppx
    s1 = pk1 + qk
    s2 = pk2 - qk
    fpkb = np.zeros((nk + 2, nj, ni), D)
    w3 = nrk == 3 and last
    for k in range(1, nk + 1):
        upb = pb[k, J, I] + (1.0 + kdiv) * pdb[k, J, I]
        oldb = -kdiv * pdb[k, J, I]
        if w3:
            upb = upb + kdiv * b["ppx"][k, J, I]
            oldb = oldb - kdiv * b["ppx"][k, J, I]
            b["ppx"][k, J, I] = 0.0
        pdb[k, J, I] = 0.0
        pb[k, J, I] = oldb + upb
        fpkb[k] = upb
        if k >= 2:
            wb[k, ys, xs] += alph * s1[k] * upb
        if k <= nk - 1:
            wb[k + 1, ys, xs] += alph * s2[k] * upb
    # A5: transposed Thomas solve for w(2..nk)
    asq = float(F(alph) * F(alph))
    aa, bb, cc = {}, {}, {}
    for k in range(2, nk + 1):
        aa[k] = (-asq) * (mm[k] * (pk1[k - 1] + qk[k - 1])) if k >= 3 else np.zeros((nj, ni), D)
        cc[k] = (-asq) * ((-mm[k]) * (pk2[k] - qk[k])) if k < nk else np.zeros((nj, ni), D)
        bb[k] = 1.0 - asq * (mm[k] * ((-(pk1[k] + qk[k])) + (pk2[k - 1] - qk[k - 1])))
    rows = list(range(2, nk + 1))
    ddb = _thomas_transpose([aa[k] for k in rows], [bb[k] for k in rows], [cc[k] for k in rows],
                            [wb[k, ys, xs].copy() for k in rows])
    for r, k in enumerate(rows):
        d = ddb[r]
        wb[k, ys, xs] = d  # w(k) before the solve enters dd(k) with weight 1
        b["wten"][k, ys, xs] += dts * d
        pb[k, J, I] += (-mm[k]) * d
        pb[k - 1, J, I] -= (-mm[k]) * d
        fpkb[k] += alph * (-mm[k]) * d
        fpkb[k - 1] -= alph * (-mm[k]) * d
    # A4: pressure advection (w after bcwsfc)
    urow = np.zeros((nk + 2, nj, ni + 1), D)  # adjoints of u' on faces 0..ni of rows ys
    vrow = np.zeros((nk + 2, nj + 1, ni), D)
    for k in range(1, nk + 1):
        f = fpkb[k]
        hzb = f * rr0[k - 1]
        # hz = -(xa + ya) * rr0 ; xa = temx*(rho*u_e)*(pe - pc) + temx*(rho*u_w)*(pc - pw)
        pc = pia[k, J, I]
        urow[k, :, 1:] += -(temx * rho[k - 1]) * (pia[k, J, 2 : ni + 2] - pc) * hzb
        urow[k, :, :ni] += -(temx * rho[k - 1]) * (pc - pia[k, J, 0:ni]) * hzb
        vrow[k, 1:, :] += -(temy * rho[k - 1]) * (pia[k, 2 : nj + 2, I] - pc) * hzb
        vrow[k, :nj, :] += -(temy * rho[k - 1]) * (pc - pia[k, 0:nj, I]) * hzb
        wb[k, ys, xs] += beta * pk1[k] * f
        wb[k + 1, ys, xs] += beta * pk2[k] * f
    # A3: bcwsfc on flat terrain sets w(1) = 0
    wb[1, ys, xs] = 0.0
    # A2: fpk = dts*(ppten - ppterm*div) + beta*((-qk)*(w2 - w1)); div from u', v'
    for k in range(1, nk + 1):
        f = fpkb[k]
        b["ppten"][k, J, I] += dts * f
        divb = -dts * ppterm[k, J, I] * f
        if k < nk:
            wb[k + 1, ys, xs] += -beta * qk[k] * f
        if k > 1:
            wb[k, ys, xs] += beta * qk[k] * f
        urow[k, :, 1:] += rdx * divb
        urow[k, :, :ni] -= rdx * divb
        vrow[k, 1:, :] += rdy * divb
        vrow[k, :nj, :] -= rdy * divb
    ub[1 : nk + 1, ys, ng : ng + ni + 1] += urow[1 : nk + 1]
    vb[1 : nk + 1, ng : ng + nj + 1, xs] += vrow[1 : nk + 1]
    # A1: pressure gradient on faces 0..ni (x) and 0..nj (y): u' = u + dts*uten - dts*cu*(ppd_r - ppd_l)
    for k in range(1, nk + 1):
        ubf = ub[k, ys, ng : ng + ni + 1]  # faces 0..ni, rows J
        b["uten"][k, ys, ng : ng + ni + 1] += dts * ubf
        cu = (halfcp * (thv[k, J, 1 : ni + 2] + thv[k, J, 0 : ni + 1])) * rdx
        pdb[k, J, 1 : ni + 2] += -dts * cu * ubf
        pdb[k, J, 0 : ni + 1] += dts * cu * ubf
        vbf = vb[k, ng : ng + nj + 1, xs]
        b["vten"][k, ng : ng + nj + 1, xs] += dts * vbf
        cv = (halfcp * (thv[k, 1 : nj + 2, I] + thv[k, 0 : nj + 1, I])) * rdy
        pdb[k, 1 : nj + 2, I] += -dts * cv * vbf
        pdb[k, 0 : nj + 1, I] += dts * cv * vbf


def stage_adjoint(bar, coef, m, *, ni, nj, nk, dt, nsound, nrk, dx, dy, dz, prof=None):
    """Input adjoints of one acoustic RK stage from its output adjoints (binary64).

    ``bar`` holds adjoints of the stage's outputs (u3d, v3d, w3d, pp3d, ppd,
    ppx, rru, rrv, rrw; full padded shapes).  Returned: adjoints of the
    stage's inputs -- ``u3d, v3d, w3d, pp3d`` (the state it starts from at
    nrk = 1, or what it leaves untouched at nrk > 1), ``ua, va, wa, ppi``
    (nrk > 1), ``wa`` also through rrw's start, ``ppx`` (through ppd's start,
    and ppx itself where the stage does not write it), ``uten, vten, wten,
    ppten``.
    """
    ng = NG
    # the restatement's constants, in its arithmetic type (binary32 as on the
    # device, or binary64 for the yardstick copy of this package)
    nloop, dts_v = sound3_stage_schedule_f32(dt, nsound, nrk)
    dts = float(F(dts_v))
    alph, kdiv = float(F(ALPH)), float(F(KDIV))
    beta = float(F(1.0) - F(ALPH))
    rdx = float(F(1.0) / F(dx))
    rdy = float(F(1.0) / F(dy))
    prof = prof if prof is not None else terrain_coordinate_profiles(nk, dz)
    qk, pk1, pk2, mm = _coefficients(coef, m, prof, ni=ni, nj=nj, nk=nk, dts=dts)
    b = {k: bar[k].astype(D).copy() for k in STATE}
    shp = {k: bar[k].shape for k in STATE}
    for name, like in (("uten", "u3d"), ("vten", "v3d"), ("wten", "w3d"), ("ppten", "pp3d")):
        b[name] = np.zeros(shp[like], D)
    # stage end
    if nrk < 3:
        _bcp_adjoint(b["pp3d"], ni=ni, nj=nj, nk=nk)
    b["w3d"][1, ng:ng + nj, ng:ng + ni] = 0.0  # bcwsfc (flat)
    b["w3d"][1, ng - 1 : ng + nj + 1, ng - 1 : ng + ni + 1] = 0.0  # bcw_rigid
    b["w3d"][nk + 1, ng - 1 : ng + nj + 1, ng - 1 : ng + ni + 1] = 0.0
    _family_adjoint(b["w3d"], 3, ni=ni, nj=nj, nk=nk)
    _family_adjoint(b["v3d"], 2, ni=ni, nj=nj, nk=nk)
    _family_adjoint(b["u3d"], 1, ni=ni, nj=nj, nk=nk)
    for n in range(nloop, 0, -1):
        _small_step_adjoint(b, coef, m, qk, pk1, pk2, mm, ni=ni, nj=nj, nk=nk, dts=dts, rdx=rdx, rdy=rdy, nrk=nrk,
                            n=n, nloop=nloop, alph=alph, beta=beta, kdiv=kdiv)
    # stage start
    ks = slice(1, nk + 1)
    yy, xx = slice(ng - 1, ng + nj + 1), slice(ng - 1, ng + ni + 1)
    out = {k: b[k] for k in ("u3d", "v3d", "w3d", "pp3d", "uten", "vten", "wten", "ppten")}
    out["wa"] = np.zeros(shp["w3d"], D)
    out["wa"][ks, yy, xx] += beta * b["rrw"][ks, yy, xx]
    if nrk != 1:
        out["ua"] = np.zeros(shp["u3d"], D)
        out["va"] = np.zeros(shp["v3d"], D)
        for name, src in (("ua", "u3d"), ("va", "v3d"), ("wa", "w3d")):
            out[name][ks, yy, xx] += b[src][ks, yy, xx]
            b[src][ks, yy, xx] = 0.0
        out["ppi"] = np.zeros(shp["pp3d"], D)
        out["ppi"][ks] += b["pp3d"][ks]
        b["pp3d"][ks] = 0.0
    out["ppx"] = b["ppx"].copy()
    out["ppx"][ks] += b["ppd"][ks]
    return out


__all__ = ["STATE", "stage_adjoint"]
