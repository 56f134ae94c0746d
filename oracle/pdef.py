"""Final-stage moisture advection through the positive-definite chain (test infrastructure).

Restates, whole-domain and vectorised in binary32 NumPy, the chain
``launch_cuda_scalar_moist_rk_stage`` runs for each moisture row at nrk=3
(``cuda_scalar_moist_rk.py:1311-1499``, CM1 adv.F:421-445 with pdef=1 and
pdef.F:15-581):

  1. WENO5 face fluxes with the magnitude gate (fluxes are 0 where every
     stencil value is <= weps), and the vertical lower closures of
     ``cm1_rk_pdef_vertical_flux_v1`` with ``pdefweno = (species != qv)``;
  2. directional tendencies ax, ay, az with the open-side inflow edits
     (``cm1_rk_pdef_directional_v1``);
  3. for each direction independently: candidate ``rho0*A + dt*(adv_d +
     s*div_d)*gz`` and the neighbour-donor repair of negative candidates
     (``cm1_rk_pdef_candidate_*``/``correct_*``; the row is extended past the
     domain by the lateral family as bc.F fills it, the column by zeros);
  4. total ``((ax+ay)+az + s*divx)*rr0*gz``.

Pinned by the reference's CPU transcription ``cpu_pdef_direction_inplace``
(:352-520) on its exact interior (golden vectors), and by closed forms: a
non-negative candidate is left alone, a repaired row conserves its sum.
The gated fluxes and the directional/total kernels have no CPU form in the
reference (parity unpinned against reference words beyond those checks).
"""

from __future__ import annotations

import numpy as np

from .common import F, NG, upstrpd, weno5

SMEPS = F(1.0e-30)
ZERO = F(0.0)


def _gate(stencil, weps):
    mx = np.maximum(np.maximum(np.abs(stencil[0]), np.abs(stencil[1])),
                    np.maximum(np.abs(stencil[2]), np.maximum(np.abs(stencil[3]), np.abs(stencil[4]))))
    return mx.astype(np.float64) > weps


def pdef_fluxes(s, rru, rrv, rrw, c1, c2, *, ni, nj, nk, weps, pdefweno):
    """fx (u shape), fy (v shape), fz (w shape) with only the written faces set."""
    ng = NG
    fx = np.zeros(rru.shape, F)
    fy = np.zeros(rrv.shape, F)
    fz = np.zeros(rrw.shape, F)
    S = s[1 : nk + 1, ng : ng + nj, :]
    x = np.arange(ni + 1) + ng
    m = rru[1 : nk + 1, ng : ng + nj, ng : ng + ni + 1]
    pos = [S[..., x - 3], S[..., x - 2], S[..., x - 1], S[..., x], S[..., x + 1]]
    neg = [S[..., x + 2], S[..., x + 1], S[..., x], S[..., x - 1], S[..., x - 2]]
    st = [np.where(m >= ZERO, p, q) for p, q in zip(pos, neg)]
    fx[1 : nk + 1, ng : ng + nj, ng : ng + ni + 1] = np.where(_gate(st, weps), m * weno5(*st, weps), ZERO)
    S = s[1 : nk + 1, :, ng : ng + ni]
    y = np.arange(nj + 1) + ng
    m = rrv[1 : nk + 1, ng : ng + nj + 1, ng : ng + ni]
    pos = [S[:, y - 3], S[:, y - 2], S[:, y - 1], S[:, y], S[:, y + 1]]
    neg = [S[:, y + 2], S[:, y + 1], S[:, y], S[:, y - 1], S[:, y - 2]]
    st = [np.where(m >= ZERO, p, q) for p, q in zip(pos, neg)]
    fy[1 : nk + 1, ng : ng + nj + 1, ng : ng + ni] = np.where(_gate(st, weps), m * weno5(*st, weps), ZERO)
    ys, xs = slice(ng, ng + nj), slice(ng, ng + ni)
    wf = F(weps)
    for k in range(2, nk + 1):
        m = rrw[k, ys, xs]
        sk = lambda kk: s[kk, ys, xs]  # noqa: E731
        interp = c1[k - 1] * sk(k - 1) + c2[k - 1] * sk(k)
        if k == 2:
            r = np.where(m >= ZERO, interp, upstrpd(sk(k + 1), sk(k), sk(k - 1), wf))
            fz[k, ys, xs] = m * r
            continue
        if k == nk:
            r = np.where(m > ZERO, upstrpd(sk(k - 2), sk(k - 1), sk(k), wf), interp)
            fz[k, ys, xs] = m * r
            continue
        low = np.zeros(m.shape, bool)
        if k == 3:
            low = (m > ZERO) if pdefweno else (m >= ZERO)
        if k == nk - 1:
            low = low | (m < ZERO)
        lowval = upstrpd(sk(k - 2), sk(k - 1), sk(k), wf) if k == 3 else upstrpd(sk(k + 1), sk(k), sk(k - 1), wf)
        if k == 3 and k == nk - 1:
            lowval = np.where(m < ZERO, upstrpd(sk(k + 1), sk(k), sk(k - 1), wf), lowval)
        pos = [sk(k - 3), sk(k - 2), sk(k - 1), sk(k), sk(k + 1)]
        neg = [sk(k + 2), sk(k + 1), sk(k), sk(k - 1), sk(k - 2)]
        st = [np.where(m >= ZERO, p, q) for p, q in zip(pos, neg)]
        wen = m * weno5(*st, weps)
        if pdefweno:
            wen = np.where(_gate(st, weps), wen, ZERO)
        fz[k, ys, xs] = np.where(low, m * lowval, wen)
    return fx, fy, fz


def directional(fx, fy, fz, s, rru, rrv, *, ni, nj, nk, dx, dy, dz, sides, flat=None):
    """ax, ay, az (cells) with the open-side inflow edits (cm1_rk_pdef_directional_v1).

    ``flat``: -(fe-fw)*rdx and ((d*rdz)*mh(k)) (cuda_gm1_child_stretched_scalar.py:355-378).
    """
    ng = NG
    rdx, rdy, rdz = F(1.0) / F(dx), F(1.0) / F(dy), F(1.0) / F(dz)
    K, J, I = slice(1, nk + 1), slice(ng, ng + nj), slice(ng, ng + ni)
    fw, fe = fx[K, J, ng : ng + ni].copy(), fx[K, J, ng + 1 : ng + ni + 1].copy()
    fs, fn = fy[K, ng : ng + nj, I].copy(), fy[K, ng + 1 : ng + nj + 1, I].copy()
    uw, ue = rru[K, J, ng : ng + ni], rru[K, J, ng + 1 : ng + ni + 1]
    vs, vn = rrv[K, ng : ng + nj, I], rrv[K, ng + 1 : ng + nj + 1, I]
    w, e, so, no = sides
    ii = np.arange(ni)[None, None, :]
    jj = np.arange(nj)[None, :, None]
    iw = (ii == 0) & (w == 2) & (uw >= ZERO)
    ie = (ii == ni - 1) & (e == 2) & (ue <= ZERO)
    js = (jj == 0) & (so == 2) & (vs >= ZERO)
    jn = (jj == nj - 1) & (no == 2) & (vn <= ZERO)
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
    sv = s[K, J, I]
    tx = (sv * (ue - uw)) * rdx
    ty = (sv * (vn - vs)) * rdy
    ax = np.where(iw, ax - tx, ax)
    ax = np.where(ie, ax - tx, ax)
    ay = np.where(js, ay - ty, ay)
    ay = np.where(jn, ay - ty, ay)
    az = np.empty((nk, nj, ni), F)
    if flat is not None:
        rdz = F(flat.rdz)
        az[0] = (-fz[2, J, I]) * rdz
        az[nk - 1] = fz[nk, J, I] * rdz
        az[1 : nk - 1] = (-(fz[3 : nk + 1, J, I] - fz[2:nk, J, I])) * rdz
        return ax, ay, az * flat.mh[1 : nk + 1][:, None, None]
    az[0] = (-fz[2, J, I]) * rdz
    az[nk - 1] = fz[nk, J, I] * rdz
    if nk > 2:
        az[1 : nk - 1] = (fz[2:nk, J, I] - fz[3 : nk + 1, J, I]) * rdz
    return ax, ay, az


def _family_index(n, lo, hi):
    """Row index map for cells c = -2..n+1 as bc.F fills the halo (pdef_cell)."""
    c = np.arange(-2, n + 2)
    out = c.copy()
    low, high = c < 0, c >= n
    out[low] = (c[low] + n) if lo == 1 else (0 if lo == 2 else -c[low] - 1)
    out[high] = (c[high] - n) if hi == 1 else (n - 1 if hi == 2 else 2 * n - 1 - c[high])
    return out


def _row_mass(values, ext):
    """Donor-repair mass and flags along the last axis.

    ``values`` (..., n); ``ext`` maps extended cells -2..n+1 to row cells, or
    None for zero extension (the vertical).
    """
    n = values.shape[-1]
    if ext is None:
        pad = np.zeros(values.shape[:-1] + (2,), F)
        v = np.concatenate([pad, values, pad], axis=-1)
    else:
        v = values[..., ext]
    # extended index e = c + 2
    def at(c):
        return v[..., c + 2]

    def ratio(c):  # c may be an index array in -1..n
        cv = at(c)
        f1 = np.maximum(ZERO, at(c - 1))
        f2 = np.maximum(ZERO, at(c + 1))
        s = f1 + f2
        with np.errstate(all="ignore"):
            r = np.where(s > SMEPS, np.maximum(cv, -s) / s, ZERO)
        return np.where(cv >= ZERO, ZERO, r).astype(F), s

    c = np.arange(n)
    vc = at(c)
    r0, s0 = ratio(c)
    rl, _ = ratio(c - 1)
    rr, _ = ratio(c + 1)
    mass = np.zeros(values.shape, F)
    flag = np.zeros(values.shape, bool)
    neg = (vc < ZERO) & (r0 != ZERO)
    mass = np.where(neg, ZERO - s0 * r0, mass)
    flag |= neg
    pos = vc > SMEPS
    a1 = pos & (rl != ZERO)
    mass = np.where(a1, mass + vc * rl, mass)
    a2 = pos & (rr != ZERO)
    mass = np.where(a2, mass + vc * rr, mass)
    flag |= a1 | a2
    return mass, flag


def repair_direction(qa, s, rho0, gz, rgz, flux, adv, *, axis, ni, nj, nk, dt, spacing, sides, flat=None):
    """Candidate + repair of one direction; returns the corrected directional tendency.

    ``flat``: the flat column's candidate rho0(1,1,k)*s0 + dt*(adv + div) and
    correction ((dum+mass) - rho0(1,1,k)*s0)*rdt with no gz/rgz, and the
    vertical ((s*d)*rdz)*mh(k) (cuda_gm1_child_stretched_scalar.py:380-448).
    """
    ng = NG
    dt = F(dt)
    rd = F(1.0) / F(spacing)
    rdt = F(1.0) / dt
    K, J, I = slice(1, nk + 1), slice(ng, ng + nj), slice(ng, ng + ni)
    sv = s[K, J, I]
    if axis == "x":
        lo_f, hi_f = flux[K, J, ng : ng + ni], flux[K, J, ng + 1 : ng + ni + 1]
    elif axis == "y":
        lo_f, hi_f = flux[K, ng : ng + nj, I], flux[K, ng + 1 : ng + nj + 1, I]
    else:
        lo_f, hi_f = flux[1 : nk + 1, J, I], flux[2 : nk + 2, J, I]
    d = hi_f - lo_f
    if flat is not None:
        if axis == "z":
            div = ((sv * d) * F(flat.rdz)) * flat.mh[1 : nk + 1][:, None, None]
        else:
            div = (sv * d) * rd
        r0a = rho0[:, 0, 0][:, None, None] * qa
        cand = r0a + dt * (adv + div)
    else:
        div = (sv * d) * rd
        r0a = rho0 * qa
        cand = r0a + (dt * (adv + div)) * gz
    if axis != "z":
        w, e, so, no = sides
        if axis == "x":
            ii = np.arange(ni)[None, None, :]
            infl = ((ii == 0) & (w == 2) & (lo_f >= ZERO)) | ((ii == ni - 1) & (e == 2) & (hi_f <= ZERO))
        else:
            jj = np.arange(nj)[None, :, None]
            infl = ((jj == 0) & (so == 2) & (lo_f >= ZERO)) | ((jj == nj - 1) & (no == 2) & (hi_f <= ZERO))
        cand = np.where(infl, r0a, cand)
    if axis == "x":
        mass, flag = _row_mass(cand, _family_index(ni, sides[0], sides[1]))
    elif axis == "y":
        mt, ft = _row_mass(np.swapaxes(cand, 1, 2), _family_index(nj, sides[2], sides[3]))
        mass, flag = np.swapaxes(mt, 1, 2), np.swapaxes(ft, 1, 2)
    else:
        mt, ft = _row_mass(np.moveaxis(cand, 0, -1), None)
        mass, flag = np.moveaxis(mt, -1, 0), np.moveaxis(ft, -1, 0)
    first = ((cand + mass) - r0a) * rdt
    if flat is None:
        first = first * rgz
    return np.where(flag, first - div, adv), cand


def total(s, ax, ay, az, divx, rr0, gz, *, ni, nj, nk, flat=None):
    ng = NG
    sv = s[1 : nk + 1, ng : ng + nj, ng : ng + ni]
    t = ((ax + ay) + az) + sv * divx
    if flat is not None:
        return t * rr0[:, 0, 0][:, None, None]
    return (t * rr0) * gz


def pdef_advection(qa, s, a, m, divx, *, ni, nj, nk, dt, dx, dy, dz, weps, pdefweno, sides, flat=None):
    """The moisture row's final-stage advection (cells) from the whole chain (``flat``: flat stretched column)."""
    fx, fy, fz = pdef_fluxes(s, a["rru"], a["rrv"], a["rrw"], m["c1"], m["c2"], ni=ni, nj=nj, nk=nk, weps=weps,
                             pdefweno=pdefweno)
    ax, ay, az = directional(fx, fy, fz, s, a["rru"], a["rrv"], ni=ni, nj=nj, nk=nk, dx=dx, dy=dy, dz=dz,
                             sides=sides, flat=flat)
    common = dict(ni=ni, nj=nj, nk=nk, dt=dt, sides=sides, flat=flat)
    ax, _ = repair_direction(qa, s, m["rho0_3d"], m["gz"], m["rgz"], a["rru"], ax, axis="x", spacing=dx, **common)
    ay, _ = repair_direction(qa, s, m["rho0_3d"], m["gz"], m["rgz"], a["rrv"], ay, axis="y", spacing=dy, **common)
    az, _ = repair_direction(qa, s, m["rho0_3d"], m["gz"], m["rgz"], a["rrw"], az, axis="z", spacing=dz, **common)
    return total(s, ax, ay, az, divx, m["rr0"], m["gz"], ni=ni, nj=nj, nk=nk, flat=flat)
