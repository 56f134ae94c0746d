"""Fifth-order flux-form momentum advection oracle (test infrastructure only).

Restates ``cuda_momentum_advection5.py`` of the reference:
  * mass fluxes and divx    -> cpu_terrain_open_momentum_advection5_inplace :753-868
  * advu / advv / advw      -> _cpu_momentum_transport5 :871-1011 and the face
                               helpers :358-750 (``_u_fx`` ... ``_w_vertical_flux``)
  * lateral family          -> ``_LateralFamily`` :167-207, :103-152
Vectorised over whole index ranges; each output word is formed by the same
sequence of binary32 operations as the reference's scalar loop.
"""

from __future__ import annotations

import numpy as np

from .common import F, NG, flx3, flx5

ZERO = F(0.0)
HALF = F(0.5)


class Lateral:
    """Wrapped reads across periodic sides (cuda_momentum_advection5.py:167-207)."""

    def __init__(self, ni, nj, sides):
        self.ni, self.nj, self.sides = ni, nj, tuple(sides)

    def periodic(self, s):
        return self.sides[s] == 1

    def open(self, s):
        return self.sides[s] == 2

    def cx(self, i):
        i = np.asarray(i)
        out = i.copy()
        if self.periodic(0):
            out = np.where(i < 0, i + self.ni, out)
        if self.periodic(1):
            out = np.where(i >= self.ni, i - self.ni, out)
        return out

    def fx(self, i):
        i = np.asarray(i)
        out = i.copy()
        if self.periodic(0):
            out = np.where(i < 0, i + self.ni, out)
        if self.periodic(1):
            out = np.where(i > self.ni, i - self.ni, out)
        return out

    def cy(self, j):
        j = np.asarray(j)
        out = j.copy()
        if self.periodic(2):
            out = np.where(j < 0, j + self.nj, out)
        if self.periodic(3):
            out = np.where(j >= self.nj, j - self.nj, out)
        return out

    def fy(self, j):
        j = np.asarray(j)
        out = j.copy()
        if self.periodic(2):
            out = np.where(j < 0, j + self.nj, out)
        if self.periodic(3):
            out = np.where(j > self.nj, j - self.nj, out)
        return out


def mass_flux_and_divergence(a, m, divx, *, ni, nj, nk, dx, dy, dz, flat=None):
    """cuda_momentum_advection5.py:787-863 (every face, column and cell).

    ``flat`` (a flat stretched column: ``mh``, ``mf``, ``rdz``) selects the
    terrain_flag=.false. statements of cuda_gm1_child_stretched_momentum.py
    :215-262 (solve2.F:493-520, 605-617): rru = rho0(1,1,k)*u, rrv likewise,
    rrw = rf0(1,1,k)*w, and divx's vertical term ((d*rdz)*mh(k)) with
    smeps 1e-30.
    """
    ng = NG
    if flat is not None:
        return _flat_mass_flux_and_divergence(a, m, divx, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, flat=flat)
    rho0 = m["rho0_3d"]
    ks = slice(1, nk + 1)
    f = np.arange(ni + 1)
    left, right = np.maximum(f - 1, 0), np.minimum(f, ni - 1)
    dens = rho0[:, :, left] + rho0[:, :, right]
    a["rru"][ks, ng : ng + nj, ng : ng + ni + 1] = (HALF * dens) * (
        a["u3d"][ks, ng : ng + nj, ng : ng + ni + 1] * m["rgzu"][None])
    g = np.arange(nj + 1)
    south, north = np.maximum(g - 1, 0), np.minimum(g, nj - 1)
    dens = rho0[:, south, :] + rho0[:, north, :]
    a["rrv"][ks, ng : ng + nj + 1, ng : ng + ni] = (HALF * dens) * (
        a["v3d"][ks, ng : ng + nj + 1, ng : ng + ni] * m["rgzv"][None])
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    a["rrw"][1, ys, xs] = ZERO
    a["rrw"][nk + 1, ys, xs] = ZERO
    dz32 = F(dz)
    zt = F(nk) * dz32
    rzt = F(1.0) / zt
    rru, rrv = a["rru"], a["rrv"]
    for k in range(2, nk + 1):
        sigmaf = F(k - 1) * dz32
        sigma = F(k - 1.5) * dz32
        r2 = (sigmaf - sigma) * (F(1.0) / dz32)
        r1 = HALF * (F(1.0) - r2)
        r2 = HALF * r2
        ux = r2 * (rru[k, ys, xs] + rru[k, ys, ng + 1 : ng + ni + 1]) + r1 * (
            rru[k - 1, ys, xs] + rru[k - 1, ys, ng + 1 : ng + ni + 1])
        vy = r2 * (rrv[k, ys, xs] + rrv[k, ng + 1 : ng + nj + 1, xs]) + r1 * (
            rrv[k - 1, ys, xs] + rrv[k - 1, ng + 1 : ng + nj + 1, xs])
        t = ux * m["dzdx"] + vy * m["dzdy"]
        t = t * (sigmaf - zt)
        t = t * m["gz"]
        t = t * rzt
        a["rrw"][k, ys, xs] = m["rf0"][k - 1] * a["w3d"][k, ys, xs] + t
    rdx = F(1.0) / F(dx)
    rdy = F(1.0) / F(dy)
    rdz = F(1.0) / dz32
    rrw = a["rrw"]
    v = (rru[ks, ys, ng + 1 : ng + ni + 1] - rru[ks, ys, xs]) * rdx + (
        rrv[ks, ng + 1 : ng + nj + 1, xs] - rrv[ks, ys, xs]) * rdy
    v = v + (rrw[2 : nk + 2, ys, xs] - rrw[1 : nk + 1, ys, xs]) * rdz
    divx[...] = np.where(np.abs(v) < F(1.0e-20), ZERO, v)
    return rdx, rdy, rdz


def _flat_mass_flux_and_divergence(a, m, divx, *, ni, nj, nk, dx, dy, flat):
    ng = NG
    ks = slice(1, nk + 1)
    rho_c = m["rho0_3d"][:, 0, 0]
    a["rru"][ks, ng : ng + nj, ng : ng + ni + 1] = rho_c[:, None, None] * a["u3d"][ks, ng : ng + nj, ng : ng + ni + 1]
    a["rrv"][ks, ng : ng + nj + 1, ng : ng + ni] = rho_c[:, None, None] * a["v3d"][ks, ng : ng + nj + 1, ng : ng + ni]
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    a["rrw"][1, ys, xs] = ZERO
    a["rrw"][nk + 1, ys, xs] = ZERO
    for k in range(2, nk + 1):
        a["rrw"][k, ys, xs] = m["rf0"][k - 1, 0, 0] * a["w3d"][k, ys, xs]
    rdx = F(1.0) / F(dx)
    rdy = F(1.0) / F(dy)
    rdz = F(flat.rdz)
    rru, rrv, rrw = a["rru"], a["rrv"], a["rrw"]
    v = (rru[ks, ys, ng + 1 : ng + ni + 1] - rru[ks, ys, xs]) * rdx + (
        rrv[ks, ng + 1 : ng + nj + 1, xs] - rrv[ks, ys, xs]) * rdy
    v = v + ((rrw[2 : nk + 2, ys, xs] - rrw[1 : nk + 1, ys, xs]) * rdz) * flat.mh[1 : nk + 1][:, None, None]
    divx[...] = np.where(np.abs(v) < F(1.0e-30), ZERO, v)
    return rdx, rdy, rdz


def _c2_pair(c2, kface, j0, i0, j1, i1, ni, nj):
    """0.5*(c2[clamped j0,i0] + c2[clamped j1,i1]) at face level kface (:358-359)."""
    cj0, ci0 = np.clip(j0, 0, nj - 1), np.clip(i0, 0, ni - 1)
    cj1, ci1 = np.clip(j1, 0, nj - 1), np.clip(i1, 0, ni - 1)
    return HALF * (c2[kface - 1][cj0, ci0] + c2[kface - 1][cj1, ci1])


def _uvz(q, rrw_pair_sum, c2h, kface, nk, ysel, xsel):
    """Vertical flux of u or v at w-level kface (:468-567); rrw pair summed by caller."""
    mass = HALF * rrw_pair_sum
    c1 = F(1.0) - c2h
    Q = lambda lev: q[lev][ysel, xsel]  # noqa: E731
    if kface == 2:
        r = np.where(mass >= ZERO, c1 * Q(1) + c2h * Q(2), flx3(Q(3), Q(2), Q(1)))
    elif kface == nk:
        r = np.where(mass >= ZERO, flx3(Q(nk - 2), Q(nk - 1), Q(nk)), c1 * Q(nk - 1) + c2h * Q(nk))
    elif kface == 3:
        r = np.where(mass >= ZERO, flx3(Q(1), Q(2), Q(3)), flx5(Q(5), Q(4), Q(3), Q(2), Q(1)))
    elif kface == nk - 1:
        k = kface
        r = np.where(mass >= ZERO, flx5(Q(k - 3), Q(k - 2), Q(k - 1), Q(k), Q(k + 1)), flx3(Q(k + 1), Q(k), Q(k - 1)))
    else:
        k = kface
        r = np.where(mass >= ZERO, flx5(Q(k - 3), Q(k - 2), Q(k - 1), Q(k), Q(k + 1)),
                     flx5(Q(k + 2), Q(k + 1), Q(k), Q(k - 1), Q(k - 2)))
    return mass * r


def _wz(w, rrw, k, nk, ys, xs):
    """Vertical flux of w at scalar level k (:708-750)."""
    mass = HALF * (rrw[k, ys, xs] + rrw[k + 1, ys, xs])
    Q = lambda lev: w[lev, ys, xs]  # noqa: E731
    if k == 1:
        r = np.where(mass >= ZERO, HALF * (Q(1) + Q(2)), flx3(Q(3), Q(2), Q(1)))
    elif k == 2:
        r = np.where(mass >= ZERO, flx3(Q(1), Q(2), Q(3)), flx5(Q(5), Q(4), Q(3), Q(2), Q(1)))
    elif k == nk - 1:
        r = np.where(mass >= ZERO, flx5(Q(k - 2), Q(k - 1), Q(k), Q(k + 1), Q(k + 2)), flx3(Q(k + 2), Q(k + 1), Q(k)))
    elif k == nk:
        r = np.where(mass >= ZERO, flx3(Q(k - 1), Q(k), Q(k + 1)), HALF * (Q(k) + Q(k + 1)))
    else:
        r = np.where(mass >= ZERO, flx5(Q(k - 2), Q(k - 1), Q(k), Q(k + 1), Q(k + 2)),
                     flx5(Q(k + 3), Q(k + 2), Q(k + 1), Q(k), Q(k - 1)))
    return mass * r


def momentum_transport5(a, m, divx, *, ni, nj, nk, rdx, rdy, rdz, lateral_sides=(2, 2, 2, 2), flat=None):
    """advu/advv/advw increments (cuda_momentum_advection5.py:871-1011).

    ``flat`` selects the flat stretched column's statements
    (cuda_gm1_child_stretched_momentum.py:264-318, adv.F:1380-1485,
    2321-2416, 3299-3356): vertical differences ((d*rdz)*mh(k)) for u and v
    and ((d*rdz)*mf(k)) for w, tendencies scaled by rr0(1,1,k) / rrf0(1,1,k),
    the open-side corrections grouped left to right (((q*0.5)*c)*rd)*h, and
    w's horizontal differences ((d*rdx)*uh).
    """
    ng = NG
    lat = Lateral(ni, nj, lateral_sides)
    u, v, w = a["u3d"], a["v3d"], a["w3d"]
    rru, rrv, rrw = a["rru"], a["rrv"], a["rrw"]
    uh, vh = a["uh"], a["vh"]
    rho0, c2, gzu, gzv, rrf0, gz = m["rho0_3d"], m["c2"], m["gzu"], m["gzv"], m["rrf0"], m["gz"]

    # ---------------- advu: faces first_face..last_face, rows 0..nj-1
    f0 = 0 if lat.periodic(0) else 1
    f1 = ni if lat.periodic(1) else ni - 1
    faces = np.arange(f0, f1 + 1)
    J = np.arange(nj)
    west, east = lat.cx(faces - 1), lat.cx(faces)

    def u_fx(k, fcs):  # (nj, len) at level k   (:362-376)
        x = fcs + ng
        mass = HALF * (rru[k][ng : ng + nj][:, lat.fx(fcs) + ng] + rru[k][ng : ng + nj][:, lat.fx(fcs + 1) + ng])
        U = u[k][ng : ng + nj]
        pos = flx5(U[:, x - 2], U[:, x - 1], U[:, x], U[:, x + 1], U[:, x + 2])
        neg = flx5(U[:, x + 3], U[:, x + 2], U[:, x + 1], U[:, x], U[:, x - 1])
        return mass * np.where(mass >= ZERO, pos, neg)

    def u_fy_base(k, p, fcs):  # p: row-face index array (len P) -> (P, len)   (:379-393)
        y = p + ng
        x = fcs + ng
        mass = HALF * (rrv[k][y][:, lat.cx(fcs) + ng] + rrv[k][y][:, lat.cx(fcs - 1) + ng])
        U = u[k]
        pos = flx5(U[y - 3][:, x], U[y - 2][:, x], U[y - 1][:, x], U[y][:, x], U[y + 1][:, x])
        neg = flx5(U[y + 2][:, x], U[y + 1][:, x], U[y][:, x], U[y - 1][:, x], U[y - 2][:, x])
        return mass * np.where(mass >= ZERO, pos, neg)

    def u_fy(k, fcs):  # rows 0..nj  (:396-412)
        p = np.arange(nj + 1)
        out = u_fy_base(k, p, fcs)
        mt = rrv[k][p + ng][:, lat.cx(fcs) + ng] + rrv[k][p + ng][:, lat.cx(fcs - 1) + ng]
        if lat.open(2):
            sel = mt[0] >= ZERO
            out[0] = np.where(sel, u_fy_base(k, np.array([1]), fcs)[0], out[0])
        if lat.open(3):
            sel = mt[nj] <= ZERO
            out[nj] = np.where(sel, u_fy_base(k, np.array([nj - 1]), fcs)[0], out[nj])
        return out

    def u_z(kface, fcs):  # (:468-567, field u3d)
        ysel = np.arange(ng, ng + nj)[:, None]
        xsel = (fcs + ng)[None, :]
        pair = rrw[kface][ysel, lat.cx(fcs)[None, :] + ng] + rrw[kface][ysel, lat.cx(fcs - 1)[None, :] + ng]
        jj = J[:, None]
        c2h = _c2_pair(c2, kface, jj, fcs[None, :] - 1, jj, fcs[None, :], ni, nj)
        return _uvz(u, pair, c2h, kface, nk, ysel, xsel)

    xz = {kf: u_z(kf, faces) for kf in range(2, nk + 1)}
    for k in range(1, nk + 1):
        k0 = k - 1
        fxall = u_fx(k, np.arange(f0 - 1, f1 + 1))
        advx = (-(fxall[:, 1:] - fxall[:, :-1])) * rdx
        fy = u_fy(k, faces)
        advy = (-(fy[1:] - fy[:-1])) * rdy
        uc = u[k][ng : ng + nj][:, faces + ng]
        if lat.open(2):
            mass = rrv[k, ng, east + ng] + rrv[k, ng, west + ng]
            corr = (rrv[k, ng + 1, west + ng] - rrv[k, ng, west + ng]) + (rrv[k, ng + 1, east + ng] - rrv[k, ng, east + ng])
            upd = (advy[0] - (((uc[0] * HALF) * corr) * rdy) * vh[ng] if flat is not None
                   else advy[0] - (uc[0] * HALF) * (corr * (rdy * vh[ng])))
            advy[0] = np.where(mass >= ZERO, upd, advy[0])
        if lat.open(3) and not (nj - 1 == 0 and lat.open(2)):
            y = nj - 1 + ng
            mass = rrv[k, ng + nj, east + ng] + rrv[k, ng + nj, west + ng]
            corr = (rrv[k, y + 1, west + ng] - rrv[k, y, west + ng]) + (rrv[k, y + 1, east + ng] - rrv[k, y, east + ng])
            upd = (advy[nj - 1] - (((uc[nj - 1] * HALF) * corr) * rdy) * vh[y] if flat is not None
                   else advy[nj - 1] - (uc[nj - 1] * HALF) * (corr * (rdy * vh[y])))
            advy[nj - 1] = np.where(mass <= ZERO, upd, advy[nj - 1])
        if k == 1:
            advz = (-xz[2]) * rdz
        elif k == nk:
            advz = xz[nk] * rdz
        else:
            advz = (-(xz[k + 1] - xz[k])) * rdz
        if flat is not None:
            advz = advz * flat.mh[k]
        total = (advx + advy) + advz
        total = total + uc * (HALF * (divx[k0][:, east] + divx[k0][:, west]))
        if flat is not None:
            scale = m["rr0"][k0, 0, 0]
        else:
            scale = gzu[:, faces] / (HALF * (rho0[k0][:, west] + rho0[k0][:, east]))
        a["uten"][k][ng : ng + nj, faces + ng] = a["uten"][k][ng : ng + nj, faces + ng] + total * scale

    # ---------------- advv: rows first_row..last_row, columns 0..ni-1
    g0 = 0 if lat.periodic(2) else 1
    g1 = nj if lat.periodic(3) else nj - 1
    rows = np.arange(g0, g1 + 1)
    I = np.arange(ni)
    south, north = lat.cy(rows - 1), lat.cy(rows)

    def v_fx_base(k, rws, p):  # rws rows (R,), p faces (P,) -> (R, P)   (:415-429)
        y = rws + ng
        x = p + ng
        mass = HALF * (rru[k][lat.cy(rws) + ng][:, x] + rru[k][lat.cy(rws - 1) + ng][:, x])
        V = v[k][y]
        pos = flx5(V[:, x - 3], V[:, x - 2], V[:, x - 1], V[:, x], V[:, x + 1])
        neg = flx5(V[:, x + 2], V[:, x + 1], V[:, x], V[:, x - 1], V[:, x - 2])
        return mass * np.where(mass >= ZERO, pos, neg)

    def v_fx(k, rws):  # faces 0..ni   (:432-448)
        p = np.arange(ni + 1)
        out = v_fx_base(k, rws, p)
        mt = rru[k][lat.cy(rws) + ng][:, p + ng] + rru[k][lat.cy(rws - 1) + ng][:, p + ng]
        if lat.open(0):
            out[:, 0] = np.where(mt[:, 0] >= ZERO, v_fx_base(k, rws, np.array([1]))[:, 0], out[:, 0])
        if lat.open(1):
            out[:, ni] = np.where(mt[:, ni] <= ZERO, v_fx_base(k, rws, np.array([ni - 1]))[:, 0], out[:, ni])
        return out

    def v_fy(k, fcs):  # fcs row faces -> (len, ni)   (:451-465)
        y = fcs + ng
        x = I + ng
        mass = HALF * (rrv[k][lat.fy(fcs) + ng][:, x] + rrv[k][lat.fy(fcs + 1) + ng][:, x])
        V = v[k]
        pos = flx5(V[y - 2][:, x], V[y - 1][:, x], V[y][:, x], V[y + 1][:, x], V[y + 2][:, x])
        neg = flx5(V[y + 3][:, x], V[y + 2][:, x], V[y + 1][:, x], V[y][:, x], V[y - 1][:, x])
        return mass * np.where(mass >= ZERO, pos, neg)

    def v_z(kface, rws):
        ysel = (rws + ng)[:, None]
        xsel = (I + ng)[None, :]
        pair = rrw[kface][lat.cy(rws)[:, None] + ng, xsel] + rrw[kface][lat.cy(rws - 1)[:, None] + ng, xsel]
        ii = I[None, :]
        c2h = _c2_pair(c2, kface, rws[:, None] - 1, ii, rws[:, None], ii, ni, nj)
        return _uvz(v, pair, c2h, kface, nk, ysel, xsel)

    yz = {kf: v_z(kf, rows) for kf in range(2, nk + 1)}
    for k in range(1, nk + 1):
        k0 = k - 1
        fx = v_fx(k, rows)
        advx = (-(fx[:, 1:] - fx[:, :-1])) * rdx
        fyall = v_fy(k, np.arange(g0 - 1, g1 + 1))
        advy = (-(fyall[1:] - fyall[:-1])) * rdy
        vc = v[k][rows + ng][:, ng : ng + ni]
        if lat.open(0):
            mass = rru[k][north + ng, ng] + rru[k][south + ng, ng]
            corr = (rru[k][south + ng, ng + 1] - rru[k][south + ng, ng]) + (rru[k][north + ng, ng + 1] - rru[k][north + ng, ng])
            upd = (advx[:, 0] - (((vc[:, 0] * HALF) * corr) * rdx) * uh[ng] if flat is not None
                   else advx[:, 0] - (vc[:, 0] * HALF) * (corr * (rdx * uh[ng])))
            advx[:, 0] = np.where(mass >= ZERO, upd, advx[:, 0])
        if lat.open(1) and not (ni - 1 == 0 and lat.open(0)):
            x = ni - 1 + ng
            mass = rru[k][north + ng, ng + ni] + rru[k][south + ng, ng + ni]
            corr = (rru[k][south + ng, x + 1] - rru[k][south + ng, x]) + (rru[k][north + ng, x + 1] - rru[k][north + ng, x])
            upd = (advx[:, ni - 1] - (((vc[:, ni - 1] * HALF) * corr) * rdx) * uh[x] if flat is not None
                   else advx[:, ni - 1] - (vc[:, ni - 1] * HALF) * (corr * (rdx * uh[x])))
            advx[:, ni - 1] = np.where(mass <= ZERO, upd, advx[:, ni - 1])
        if k == 1:
            advz = (-yz[2]) * rdz
        elif k == nk:
            advz = yz[nk] * rdz
        else:
            advz = (-(yz[k + 1] - yz[k])) * rdz
        if flat is not None:
            advz = advz * flat.mh[k]
        total = (advx + advy) + advz
        total = total + vc * (HALF * (divx[k0][north] + divx[k0][south]))
        if flat is not None:
            scale = m["rr0"][k0, 0, 0]
        else:
            scale = gzv[rows] / (HALF * (rho0[k0][south] + rho0[k0][north]))
        a["vten"][k][rows + ng, ng : ng + ni] = a["vten"][k][rows + ng, ng : ng + ni] + total * scale

    # ---------------- advw: w levels 2..nk, every cell
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    Jc, Ic = np.arange(nj)[:, None], np.arange(ni)[None, :]

    def w_fx(kf):  # faces 0..ni -> (nj, ni+1)   (:570-636)
        p = np.arange(ni + 1)[None, :]

        def base(pp):
            c2h = HALF * (c2[kf - 1][Jc, np.clip(pp - 1, 0, ni - 1)] + c2[kf - 1][Jc, np.clip(pp, 0, ni - 1)])
            c1h = F(1.0) - c2h
            mass = c2h * rru[kf][Jc + ng, pp + ng] + c1h * rru[kf - 1][Jc + ng, pp + ng]
            W = w[kf]
            y = Jc + ng
            x = pp + ng
            pos = flx5(W[y, x - 3], W[y, x - 2], W[y, x - 1], W[y, x], W[y, x + 1])
            neg = flx5(W[y, x + 2], W[y, x + 1], W[y, x], W[y, x - 1], W[y, x - 2])
            return mass, mass * np.where(mass >= ZERO, pos, neg)

        mass, out = base(np.broadcast_to(p, (nj, ni + 1)))
        if lat.open(0):
            _, alt = base(np.full((nj, 1), 1))
            out[:, 0] = np.where(mass[:, 0] >= ZERO, alt[:, 0], out[:, 0])
        if lat.open(1):
            _, alt = base(np.full((nj, 1), ni - 1))
            out[:, ni] = np.where(mass[:, ni] <= ZERO, alt[:, 0], out[:, ni])
        return out

    def w_fy(kf):  # faces 0..nj -> (nj+1, ni)   (:639-705)
        p = np.arange(nj + 1)[:, None]

        def base(pp):
            c2h = HALF * (c2[kf - 1][np.clip(pp - 1, 0, nj - 1), Ic] + c2[kf - 1][np.clip(pp, 0, nj - 1), Ic])
            c1h = F(1.0) - c2h
            mass = c2h * rrv[kf][pp + ng, Ic + ng] + c1h * rrv[kf - 1][pp + ng, Ic + ng]
            W = w[kf]
            y = pp + ng
            x = Ic + ng
            pos = flx5(W[y - 3, x], W[y - 2, x], W[y - 1, x], W[y, x], W[y + 1, x])
            neg = flx5(W[y + 2, x], W[y + 1, x], W[y, x], W[y - 1, x], W[y - 2, x])
            return mass, mass * np.where(mass >= ZERO, pos, neg)

        mass, out = base(np.broadcast_to(p, (nj + 1, ni)))
        if lat.open(2):
            _, alt = base(np.full((1, ni), 1))
            out[0] = np.where(mass[0] >= ZERO, alt[0], out[0])
        if lat.open(3):
            _, alt = base(np.full((1, ni), nj - 1))
            out[nj] = np.where(mass[nj] <= ZERO, alt[0], out[nj])
        return out

    wz = {k: _wz(w, rrw, k, nk, ys, xs) for k in range(1, nk + 1)}
    for kf in range(2, nk + 1):
        fx = w_fx(kf)
        fy = w_fy(kf)
        if flat is not None:
            advx = ((-(fx[:, 1:] - fx[:, :-1])) * rdx) * uh[ng : ng + ni][None, :]
            advy = ((-(fy[1:] - fy[:-1])) * rdy) * vh[ng : ng + nj][:, None]
        else:
            advx = (-(fx[:, 1:] - fx[:, :-1])) * (rdx * uh[ng : ng + ni][None, :])
            advy = (-(fy[1:] - fy[:-1])) * (rdy * vh[ng : ng + nj][:, None])
        cc2 = c2[kf - 1]
        cc1 = F(1.0) - cc2
        wc = w[kf, ys, xs]
        if lat.open(0):
            mass = cc2[:, 0] * rru[kf, ys, ng] + cc1[:, 0] * rru[kf - 1, ys, ng]
            corr = cc1[:, 0] * (rru[kf - 1, ys, ng + 1] - rru[kf - 1, ys, ng]) + cc2[:, 0] * (rru[kf, ys, ng + 1] - rru[kf, ys, ng])
            upd = (advx[:, 0] - ((wc[:, 0] * corr) * rdx) * uh[ng] if flat is not None
                   else advx[:, 0] - wc[:, 0] * (corr * (rdx * uh[ng])))
            advx[:, 0] = np.where(mass >= ZERO, upd, advx[:, 0])
        if lat.open(1) and not (ni - 1 == 0 and lat.open(0)):
            x = ni - 1 + ng
            mass = cc2[:, ni - 1] * rru[kf, ys, ng + ni] + cc1[:, ni - 1] * rru[kf - 1, ys, ng + ni]
            corr = cc1[:, ni - 1] * (rru[kf - 1, ys, x + 1] - rru[kf - 1, ys, x]) + cc2[:, ni - 1] * (rru[kf, ys, x + 1] - rru[kf, ys, x])
            upd = (advx[:, ni - 1] - ((wc[:, ni - 1] * corr) * rdx) * uh[x] if flat is not None
                   else advx[:, ni - 1] - wc[:, ni - 1] * (corr * (rdx * uh[x])))
            advx[:, ni - 1] = np.where(mass <= ZERO, upd, advx[:, ni - 1])
        if lat.open(2):
            mass = cc2[0] * rrv[kf, ng, xs] + cc1[0] * rrv[kf - 1, ng, xs]
            corr = cc1[0] * (rrv[kf - 1, ng + 1, xs] - rrv[kf - 1, ng, xs]) + cc2[0] * (rrv[kf, ng + 1, xs] - rrv[kf, ng, xs])
            upd = (advy[0] - ((wc[0] * corr) * rdy) * vh[ng] if flat is not None
                   else advy[0] - wc[0] * (corr * (rdy * vh[ng])))
            advy[0] = np.where(mass >= ZERO, upd, advy[0])
        if lat.open(3) and not (nj - 1 == 0 and lat.open(2)):
            y = nj - 1 + ng
            mass = cc2[nj - 1] * rrv[kf, ng + nj, xs] + cc1[nj - 1] * rrv[kf - 1, ng + nj, xs]
            corr = cc1[nj - 1] * (rrv[kf - 1, y + 1, xs] - rrv[kf - 1, y, xs]) + cc2[nj - 1] * (rrv[kf, y + 1, xs] - rrv[kf, y, xs])
            upd = (advy[nj - 1] - ((wc[nj - 1] * corr) * rdy) * vh[y] if flat is not None
                   else advy[nj - 1] - wc[nj - 1] * (corr * (rdy * vh[y])))
            advy[nj - 1] = np.where(mass <= ZERO, upd, advy[nj - 1])
        advz = (-(wz[kf] - wz[kf - 1])) * rdz
        if flat is not None:
            advz = advz * flat.mf[kf]
        total = (advx + advy) + advz
        total = total + wc * (cc2 * divx[kf - 1] + cc1 * divx[kf - 2])
        scale = rrf0[kf - 1, 0, 0] if flat is not None else rrf0[kf - 1] * gz
        a["wten"][kf, ys, xs] = a["wten"][kf, ys, xs] + total * scale


def cpu_terrain_open_momentum_advection5_inplace(
    arrays, metrics, divx, *, ni, nj, nk, nrk, dx, dy, dz, lateral_sides=(2, 2, 2, 2), flat=None,
):
    """cuda_momentum_advection5.py:753-868 then :871-1011 (``flat``: the flat stretched column's texts)."""
    if nrk not in (1, 2, 3):
        raise ValueError("nrk must be an integer in 1..3")
    rdx, rdy, rdz = mass_flux_and_divergence(arrays, metrics, divx, ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz,
                                             flat=flat)
    momentum_transport5(arrays, metrics, divx, ni=ni, nj=nj, nk=nk, rdx=rdx, rdy=rdy, rdz=rdz,
                        lateral_sides=lateral_sides, flat=flat)
