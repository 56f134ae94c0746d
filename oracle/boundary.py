"""Lateral/vertical boundary fills and open-boundary operators (test infrastructure).

Restates, in binary32 NumPy:
  * family fills bcs/bcu/bcv/bcw     -> cuda_boundary_families.py:856-938
                                        (cpu_family_we/sn_inplace, _side_plan);
                                        mode 2 equals cpu_bc{s,u,v,w}_open_inplace
                                        (cuda_sound3_open_boundary.py:275-337)
  * pressure fills bcp               -> cuda_boundary_families.py:941-962;
                                        mode 2 equals cpu_bcp_open_inplace (:261-272)
  * rigid w lid/floor (bcw flag 1)   -> cuda_sound3_open_boundary.py:333-337
  * bcwsfc on physical bottom faces  -> cuda_terrain_open_boundary.py:641-665
  * terrain ssopenbc (open sides)    -> cuda_terrain_open_boundary.py:678-826
  * w-level interpolation            -> cuda_terrain_open_boundary.py:610-638
  * Durran-Klemp radbc (irbc=4)      -> cuda_lateral_closure.py:390-480 composed as
                                        cuda_sound3_open_boundary.py:463-483
  * periodic boundary-face pressure gradient (no CPU form in the reference;
    restates the kernel ``cm1_bc_face_pgrad_f32_v1``,
    cuda_boundary_families.py:727-783; parity unpinned against reference words)
  * bc2d w(k=1) open fill            -> cuda_momentum_rk.py:274-290
"""

from __future__ import annotations

import numpy as np

from .common import F, NG

PERIODIC, OPEN = 1, 2
KIND_SCALAR, KIND_U, KIND_V, KIND_W = 0, 1, 2, 3
ZERO = F(0.0)
HALF = F(0.5)


def _side_plan(mode, side, kind, stag_kind, layer, ng, n):
    """(dst, src, sign, wall) padded indices of one ghost plane (:856-884)."""
    stag = kind == stag_kind
    if side == 0:
        if mode == PERIODIC:
            return ng - 1 - layer, n + ng - 1 - layer, 1.0, None
        if mode == OPEN:
            return layer, ng, 1.0, None
        if stag:
            return ng - 1 - layer, ng + 1 + layer, -1.0, ng
        return ng - 1 - layer, ng + layer, 1.0, None
    if mode == PERIODIC:
        if stag:
            return n + ng + 1 + layer, ng + 1 + layer, 1.0, None
        return n + ng + layer, ng + layer, 1.0, None
    if mode == OPEN:
        if stag:
            return n + ng + 1 + layer, n + ng, 1.0, None
        return n + ng + layer, n + ng - 1, 1.0, None
    if stag:
        return n + ng + 1 + layer, n + ng - 1 - layer, -1.0, n + ng
    return n + ng + layer, n + ng - 1 - layer, 1.0, None


def family_we(arr, *, ni, nj, nk, kind, wbc, ebc):
    """West/east ghost planes of one padded carrier (cuda_boundary_families.py:887-911)."""
    ng = NG
    ny = arr.shape[1]
    kfirst, kcount = (2, nk - 1) if kind == KIND_W else (1, nk)
    jhi = nj + 2 if kind == KIND_V else nj + 1
    ks = slice(kfirst, kfirst + kcount)
    for side, mode in ((0, wbc), (1, ebc)):
        if mode == 0:
            continue
        rows = slice(0, ny) if mode == PERIODIC else slice(ng - 1, ng - 1 + jhi + 1)
        for layer in range(ng):
            dst, src, sign, wall = _side_plan(mode, side, kind, KIND_U, layer, ng, ni)
            value = arr[ks, rows, src]
            arr[ks, rows, dst] = -value if sign < 0 else value
            if wall is not None and layer == 0:
                arr[ks, rows, wall] = ZERO


def family_sn(arr, *, ni, nj, nk, kind, sbc, nbc):
    """South/north ghost planes (cuda_boundary_families.py:914-938)."""
    ng = NG
    nx = arr.shape[2]
    kfirst, kcount = (2, nk - 1) if kind == KIND_W else (1, nk)
    ihi = ni + 2 if kind == KIND_U else ni + 1
    ks = slice(kfirst, kfirst + kcount)
    for side, mode in ((0, sbc), (1, nbc)):
        if mode == 0:
            continue
        cols = slice(0, nx) if mode == PERIODIC else slice(ng - 1, ng - 1 + ihi + 1)
        for layer in range(ng):
            dst, src, sign, wall = _side_plan(mode, side, kind, KIND_V, layer, ng, nj)
            value = arr[ks, src, cols]
            arr[ks, dst, cols] = -value if sign < 0 else value
            if wall is not None and layer == 0:
                arr[ks, wall, cols] = ZERO


def fill(arr, kind, sides, *, ni, nj, nk):
    """bcs/bcu/bcw: west-east then south-north; bcv: south-north then west-east."""
    w, e, s, n = sides
    if kind == KIND_V:
        family_sn(arr, ni=ni, nj=nj, nk=nk, kind=kind, sbc=s, nbc=n)
        family_we(arr, ni=ni, nj=nj, nk=nk, kind=kind, wbc=w, ebc=e)
    else:
        family_we(arr, ni=ni, nj=nj, nk=nk, kind=kind, wbc=w, ebc=e)
        family_sn(arr, ni=ni, nj=nj, nk=nk, kind=kind, sbc=s, nbc=n)


def bcp(p, sides, *, ni, nj, nk):
    """Pressure-family one-cell halo (cuda_boundary_families.py:941-962)."""
    w, e, s, n = sides
    ks = slice(1, nk + 1)
    if w:
        p[ks, :, 0] = p[ks, :, ni if w == PERIODIC else 1]
    if e:
        p[ks, :, ni + 1] = p[ks, :, 1 if e == PERIODIC else ni]
    if s:
        p[ks, 0, :] = p[ks, nj if s == PERIODIC else 1, :]
    if n:
        p[ks, nj + 1, :] = p[ks, 1 if n == PERIODIC else nj, :]


def bcw_rigid(w, *, ni, nj, nk):
    """w = 0 on the floor and lid over rows/cols 0..n+1 (bcw flag 1)."""
    ng = NG
    w[1, ng - 1 : ng + nj + 1, ng - 1 : ng + ni + 1] = ZERO
    w[nk + 1, ng - 1 : ng + nj + 1, ng - 1 : ng + ni + 1] = ZERO


def bcwsfc(a, m, *, ni, nj, nk):
    """w(k=1) from the terrain slope on physical cells (cuda_terrain_open_boundary.py:641-665)."""
    ng = NG
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    u1, v1 = a["u3d"][1], a["v3d"][1]
    uterm = (u1[ys, xs] + u1[ys, ng + 1 : ng + ni + 1]) * m["dzdx"]
    vterm = (v1[ys, xs] + v1[ng + 1 : ng + nj + 1, xs]) * m["dzdy"]
    a["w3d"][1, ys, xs] = (HALF * (uterm + vterm)) * m["gz"]


def bc2d_open_w_surface(w, *, ni, nj):
    """w(k=1) halo in bc.F's W, E, S, N order (cuda_momentum_rk.py:274-290)."""
    ng = NG
    plane = w[1]
    plane[:, :ng] = plane[:, ng : ng + 1]
    plane[:, ng + ni : ng + ni + ng] = plane[:, ng + ni - 1 : ng + ni]
    plane[:ng, :] = plane[ng : ng + 1, :]
    plane[ng + nj : ng + nj + ng, :] = plane[ng + nj - 1 : ng + nj, :]


def interp_at_w(q, kface, y, x, nk, prof):
    """Value of a u/v column at w level kface (cuda_terrain_open_boundary.py:610-638)."""
    if kface == 1:
        c = prof["cgs"]
        return (c[0] * q[1, y, x] + c[1] * q[2, y, x]) + c[2] * q[3, y, x]
    if kface == nk + 1:
        c = prof["cgt"]
        return (c[0] * q[nk, y, x] + c[1] * q[nk - 1, y, x]) + c[2] * q[nk - 2, y, x]
    r2 = (prof["sigmaf"][kface] - prof["sigma"][kface - 1]) * prof["rds"][kface]
    r1 = F(1.0) - r2
    return r1 * q[kface - 1, y, x] + r2 * q[kface, y, x]


def _at_w_all(q, y, x, nk, prof):
    return np.stack([interp_at_w(q, kf, y, x, nk, prof) for kf in range(1, nk + 2)])


def ssopen_terrain(a, m, prof, sides, *, ni, nj, nk, dts, dx, dy):
    """Open normal-velocity small step on open sides (cuda_terrain_open_boundary.py:678-826)."""
    ng = NG
    dts = F(dts)
    rdx = F(1.0) / F(dx)
    rdy = F(1.0) / F(dy)
    u, v = a["u3d"], a["v3d"]
    rdsf = prof["rdsf"]
    ys = np.arange(ng, ng + nj)
    xs = np.arange(ng, ng + ni)
    if sides[0] == OPEN:  # west, bc.F:2392-2459
        s1 = _at_w_all(u, ys, ng, nk, prof)
        s2 = _at_w_all(u, ys, ng + 1, nk, prof)
        for k in range(1, nk + 1):
            h = m["gz"][:, 0] * (u[k, ys, ng + 1] * m["rgzu"][:, 1] - u[k, ys, ng] * m["rgzu"][:, 0])
            h = (h * rdx) * a["uh"][ng]
            vert = m["gx"][k, :, 0] * (s2[k] + s1[k]) - m["gx"][k - 1, :, 0] * (s2[k - 1] + s1[k - 1])
            vert = (HALF * vert) * rdsf[k]
            rad = (-a["radbcw"][k, ys]) * (h + vert)
            t = rad + a["uten"][k, ys, ng]
            u[k, ys, ng] = u[k, ys, ng] + dts * t
    if sides[1] == OPEN:  # east, bc.F:2467-2534
        s0 = _at_w_all(u, ys, ng + ni - 1, nk, prof)
        s1 = _at_w_all(u, ys, ng + ni, nk, prof)
        for k in range(1, nk + 1):
            h = m["gz"][:, ni - 1] * (u[k, ys, ng + ni] * m["rgzu"][:, ni] - u[k, ys, ng + ni - 1] * m["rgzu"][:, ni - 1])
            h = (h * rdx) * a["uh"][ng + ni - 1]
            vert = m["gx"][k, :, ni - 1] * (s1[k] + s0[k]) - m["gx"][k - 1, :, ni - 1] * (s1[k - 1] + s0[k - 1])
            vert = (HALF * vert) * rdsf[k]
            rad = (-a["radbce"][k, ys]) * (h + vert)
            t = rad + a["uten"][k, ys, ng + ni]
            u[k, ys, ng + ni] = u[k, ys, ng + ni] + dts * t
    if sides[2] == OPEN:  # south, bc.F:2542-2609
        s1 = _at_w_all(v, ng, xs, nk, prof)
        s2 = _at_w_all(v, ng + 1, xs, nk, prof)
        for k in range(1, nk + 1):
            h = m["gz"][0, :] * (v[k, ng + 1, xs] * m["rgzv"][1, :] - v[k, ng, xs] * m["rgzv"][0, :])
            h = (h * rdy) * a["vh"][ng]
            vert = m["gy"][k, 0, :] * (s2[k] + s1[k]) - m["gy"][k - 1, 0, :] * (s2[k - 1] + s1[k - 1])
            vert = (HALF * vert) * rdsf[k]
            rad = (-a["radbcs"][k, xs]) * (h + vert)
            t = rad + a["vten"][k, ng, xs]
            v[k, ng, xs] = v[k, ng, xs] + dts * t
    if sides[3] == OPEN:  # north, bc.F:2617-2684
        s0 = _at_w_all(v, ng + nj - 1, xs, nk, prof)
        s1 = _at_w_all(v, ng + nj, xs, nk, prof)
        for k in range(1, nk + 1):
            h = m["gz"][nj - 1, :] * (v[k, ng + nj, xs] * m["rgzv"][nj, :] - v[k, ng + nj - 1, xs] * m["rgzv"][nj - 1, :])
            h = (h * rdy) * a["vh"][ng + nj - 1]
            vert = m["gy"][k, nj - 1, :] * (s1[k] + s0[k]) - m["gy"][k - 1, nj - 1, :] * (s1[k - 1] + s0[k - 1])
            vert = (HALF * vert) * rdsf[k]
            rad = (-a["radbcn"][k, xs]) * (h + vert)
            t = rad + a["vten"][k, ng + nj, xs]
            v[k, ng + nj, xs] = v[k, ng + nj, xs] + dts * t


def radiation4(a, *, ni, nj, nk, dx, dy, dttmp, limit=350.0):
    """Durran-Klemp phase speeds averaged over k (cuda_lateral_closure.py:390-480).

    Packed index p of the reference maps to padded index p + ng - 1.
    """
    ng = NG
    off = ng - 1
    dx32, dy32, dt32 = F(dx), F(dy), F(dttmp)
    eps = F(1.0e-10)
    lim = F(limit)
    ua, u3d, va, v3d = a["ua"], a["u3d"], a["va"], a["v3d"]
    ks = slice(1, nk + 1)
    rows = np.arange(1, nj + 1) + off  # packed j = 1..nj
    cols = np.arange(1, ni + 1) + off

    def speed(foo_hi, foo_lo, qa_n, q3_n, rf_n, d32, west):
        foo = foo_hi - foo_lo
        num = (d32 * rf_n) * (qa_n - q3_n)
        signed = np.copysign(eps, foo).astype(F)
        den = dt32 * (signed + foo)
        with np.errstate(all="ignore"):
            c = num / den
        if west:
            return np.maximum(np.minimum(c, ZERO), -lim)
        return np.minimum(np.maximum(c, ZERO), lim)

    U, U3, V, V3 = ua[ks], u3d[ks], va[ks], v3d[ks]
    cw = speed(U[:, rows, 3 + off], U[:, rows, 2 + off], U[:, rows, 2 + off], U3[:, rows, 2 + off],
               a["ruf"][2 + off], dx32, True)
    ce = speed(U[:, rows, ni + off], U[:, rows, ni - 1 + off], U[:, rows, ni + off], U3[:, rows, ni + off],
               a["ruf"][ni + off], dx32, False)
    cs = speed(V[:, 3 + off, cols], V[:, 2 + off, cols], V[:, 2 + off, cols], V3[:, 2 + off, cols],
               a["rvf"][2 + off], dy32, True)
    cn = speed(V[:, nj + off, cols], V[:, nj - 1 + off, cols], V[:, nj + off, cols], V3[:, nj + off, cols],
               a["rvf"][nj + off], dy32, False)

    def kavg(c):
        acc = np.zeros(c.shape[1:], F)
        for k in range(nk):
            acc = acc + c[k]
        return acc / F(nk)

    aw, ae, as_, an = kavg(cw), kavg(ce), kavg(cs), kavg(cn)
    a["radbcw"][1 : nk + 1, ng : ng + nj] = aw[None, :]
    a["radbce"][1 : nk + 1, ng : ng + nj] = ae[None, :]
    a["radbcs"][1 : nk + 1, ng : ng + ni] = as_[None, :]
    a["radbcn"][1 : nk + 1, ng : ng + ni] = an[None, :]


def periodic_face_pgrad(a, m, prof, sides, *, ni, nj, nk, dts, rdx, rdy, cp):
    """Pressure gradient on periodic boundary faces 1 and n+1 (kernel text only)."""
    ng = NG
    if 1 not in sides:
        return
    halfcp = F(cp) * HALF
    eighth = F(0.125)
    rds = prof["rds"]
    ppd, thv = a["ppd"], a["thv"]
    dts = F(dts)

    def vert(k, jj, ii):
        dph = (ppd[k + 1, jj, ii] - ppd[k, jj, ii]) * rds[k + 1] if k < nk else np.zeros_like(ppd[k, jj, ii])
        dpl = (ppd[k, jj, ii] - ppd[k - 1, jj, ii]) * rds[k] if k > 1 else np.zeros_like(ppd[k, jj, ii])
        return dph, dpl

    J = np.arange(1, nj + 1)
    I = np.arange(1, ni + 1)
    for k in range(1, nk + 1):
        for half, on in ((0, sides[0] == 1), (1, sides[1] == 1)):
            if not on:
                continue
            ff = 1 if half == 0 else ni + 1
            h = (ppd[k, J, ff] - ppd[k, J, ff - 1]) * F(rdx)
            dph, dpl = vert(k, J, ff)
            dphl, dpll = vert(k, J, ff - 1)
            vs = (dph + dphl) + (dpl + dpll)
            i0 = ff - 1
            tr = (eighth * vs) * (m["gxu"][k - 1, J - 1, i0] + m["gxu"][k, J - 1, i0])
            pg = (halfcp * (thv[k, J, ff] + thv[k, J, ff - 1])) * (h + tr)
            y = J - 1 + ng
            x = i0 + ng
            a["u3d"][k, y, x] = a["u3d"][k, y, x] + dts * (a["uten"][k, y, x] - pg)
        for half, on in ((0, sides[2] == 1), (1, sides[3] == 1)):
            if not on:
                continue
            ff = 1 if half == 0 else nj + 1
            h = (ppd[k, ff, I] - ppd[k, ff - 1, I]) * F(rdy)
            dph, dpl = vert(k, ff, I)
            dphl, dpll = vert(k, ff - 1, I)
            vs = (dph + dphl) + (dpl + dpll)
            j0 = ff - 1
            tr = (eighth * vs) * (m["gyv"][k - 1, j0, I - 1] + m["gyv"][k, j0, I - 1])
            pg = (halfcp * (thv[k, ff, I] + thv[k, ff - 1, I])) * (h + tr)
            y = j0 + ng
            x = I - 1 + ng
            a["v3d"][k, y, x] = a["v3d"][k, y, x] + dts * (a["vten"][k, y, x] - pg)
