"""Adjoint of the dry large step on flat terrain with periodic sides, in binary64 (test infrastructure).

The forward is ``large_step.large_step`` (cuda_gm1_forecast.py:1569-1627,
rk stage :1241-1345, solve3 :1414-1457) for a dry state (numq = 0, no held
closure tendencies) on flat, horizontally uniform metrics (gz = rgzu = rgzv =
1, zero slopes, one base column) with CM1's (1, 1, 1, 1) lateral sides.  On
that geometry every lateral stencil is a periodic shift, so the state is kept
here without halos: u, v and scalars (nk, nj, ni), w (nk+1, nj, ni) holding
levels 1..nk+1, with u face f the west face of cell f and v face g the south
face of row g.  ``Step.forward`` restates the large step statement by
statement in that form (pinned against ``large_step.large_step`` in
tests/test_adjoint_large_step.py) and ``Step.adjoint`` is its transpose at a
recorded trajectory:

  * bilinear terms (mass flux x transported field, coefficient x pressure)
    transpose into both factors;
  * the upwind selections, the WENO5 weights and the positive-definite
    limiter of the final RK stage's scalar fluxes are frozen at their forward
    values (a ``Tape`` records them; replaying it makes the forward the
    linear map the adjoint transposes, which is what the finite-difference
    checks differentiate);
  * the smeps clamps are taken as the identity;
  * the Thomas solve's transpose is a solve with T^T, and its dependence on
    the coefficients (thv -> mm, ppterm -> qk, piadv -> pk1/pk2) is carried
    through as -T^T-solve(xbar) . (dT w).

The loss of the adjoint configuration is ``storm_intensity``: the
base-density-weighted mean of w^2 over the w levels between 2 and 8 km.
"""

from __future__ import annotations

import numpy as np

from . import stage as stg
from .common import (C1_6, C2_6, C2_60, C5_6, C11_6, C27_60, C47_60, CM1_6, CM3_60, CM7_6, CM13_60, C13_12, F,
                     flx3, flx5)
from .large_step import THETA_WEPS
from .sound import ALPH, KDIV, SMEPS, sound3_stage_schedule_f32

D = np.float64
FLX5 = tuple(float(c) for c in (C2_60, CM13_60, C47_60, C27_60, CM3_60))
FLX3 = tuple(float(c) for c in (CM1_6, C5_6, C2_6))
STATE = ("u", "v", "w", "th", "pp", "ppx")


def ex(a, s):
    """a at i+s (periodic in x)."""
    return np.roll(a, -s, axis=-1)


def ny(a, s):
    """a at j+s (periodic in y)."""
    return np.roll(a, -s, axis=-2)


def _st_T(bar, offs, coefs, shift):
    """Transpose of sum_i coefs[i] * shift(q, offs[i]) applied to bar."""
    out = np.zeros_like(bar)
    for o, c in zip(offs, coefs):
        out += shift(c * bar, -o)
    return out


class Tape:
    """Upwind selections, frozen WENO5/limiter coefficients and trajectory of one forward pass.

    Recording evaluates the forward's own nonlinear choices and stores them;
    replaying reuses them (and drops the smeps clamps), turning the forward
    into the linear map that the adjoint transposes.
    """

    def __init__(self, exact: bool = True):
        self.record = True
        self.exact = exact
        self.sel = {}
        self.coef = {}
        self.jac = {}
        self.branch = {}
        self.traj = {}

    def replay(self):
        self.record = False
        return self

    def mask(self, key, cond):
        if self.record:
            self.sel[key] = np.asarray(cond)
        return self.sel[key]

    def clamp(self, x, eps):
        return np.where(np.abs(x) < eps, 0.0, x) if self.record else x

    def weno(self, key, s1, s2, s3, s4, s5, weps):
        """WENO5 (common.weno5); on replay its normalised weights stay frozen unless ``exact``."""
        if self.record or self.exact:
            if self.record:
                self.jac[key] = _weno_jacobian(s1, s2, s3, s4, s5, weps)
            d1 = (s1 - 2.0 * s2) + s3
            e1 = (s1 - 4.0 * s2) + 3.0 * s3
            b1 = float(C13_12) * (d1 * d1) + 0.25 * (e1 * e1)
            d2 = (s2 - 2.0 * s3) + s4
            e2 = s2 - s4
            b2 = float(C13_12) * (d2 * d2) + 0.25 * (e2 * e2)
            d3 = (s3 - 2.0 * s4) + s5
            e3 = (3.0 * s3 - 4.0 * s4) + s5
            b3 = float(C13_12) * (d3 * d3) + 0.25 * (e3 * e3)
            tau = np.abs(b1 - b3)
            with np.errstate(all="ignore"):
                q1 = np.fmin(1.0e30, tau / (b1 + weps))
                q2 = np.fmin(1.0e30, tau / (b2 + weps))
                q3 = np.fmin(1.0e30, tau / (b3 + weps))
            w1 = float(F(0.1)) * (1.0 + q1 * q1)
            w2 = float(F(0.6)) * (1.0 + q2 * q2)
            w3 = float(F(0.3)) * (1.0 + q3 * q3)
            den = (w1 + w2) + w3
            if self.record:
                self.coef[key] = (w1 / den, w2 / den, w3 / den)
            p1, p2, p3 = self._parts(s1, s2, s3, s4, s5)
            return ((w1 * p1 + w2 * p2) + w3 * p3) / den
        o1, o2, o3 = self.coef[key]
        p1, p2, p3 = self._parts(s1, s2, s3, s4, s5)
        return (o1 * p1 + o2 * p2) + o3 * p3

    @staticmethod
    def _parts(s1, s2, s3, s4, s5):
        p1 = (float(C2_6) * s1 + float(CM7_6) * s2) + float(C11_6) * s3
        p2 = (float(CM1_6) * s2 + float(C5_6) * s3) + float(C2_6) * s4
        p3 = (float(C2_6) * s3 + float(C5_6) * s4) + float(CM1_6) * s5
        return p1, p2, p3

    def weno_coefs(self, key):
        """d(reconstruction)/d(s1..s5): exact, or with the weights frozen."""
        if self.exact:
            return self.jac[key]
        o1, o2, o3 = self.coef[key]
        c = lambda x: float(x)  # noqa: E731
        return (o1 * c(C2_6), o1 * c(CM7_6) + o2 * c(CM1_6), (o1 * c(C11_6) + o2 * c(C5_6)) + o3 * c(C2_6),
                o2 * c(C2_6) + o3 * c(C5_6), o3 * c(CM1_6))

    def upstrpd(self, key, s1, s2, s3, weps):
        """Positive-definite closure (common.upstrpd); on replay its limiter stays frozen, or with ``exact``
        keeps its recorded branch (0, 1, rr or 1/6 + rr/3) evaluated at the new values."""
        dd = s2 - s1
        se = np.copysign(abs(weps), dd)
        with np.errstate(all="ignore"):
            rr = (s3 - s2) / (se + dd)
        cand = float(C1_6) + float(C2_6) * rr
        if self.record:
            phi = np.maximum(0.0, np.minimum(rr, np.minimum(cand, 1.0)))
            inner = np.minimum(cand, 1.0)
            mn = np.minimum(rr, inner)
            br = np.where(~(mn > 0.0), 0, np.where(rr <= inner, 2, np.where(cand <= 1.0, 3, 1)))
            self.coef[key] = phi
            self.branch[key] = br
            den = se + dd
            with np.errstate(all="ignore"):
                g1 = (s3 - s2) / (den * den)
                g2 = -1.0 / den - (s3 - s2) / (den * den)
                g3 = 1.0 / den
            scale = np.where(br == 2, 1.0, np.where(br == 3, float(C2_6), 0.0))
            dphi = [scale * g for g in (g1, g2, g3)]
            self.jac[key] = (-phi + dd * dphi[0], 1.0 + phi + dd * dphi[1], dd * dphi[2])
            return s2 + phi * (s2 - s1)
        if self.exact:
            br = self.branch[key]
            phi = np.where(br == 0, 0.0, np.where(br == 1, 1.0, np.where(br == 2, rr, cand)))
        else:
            phi = self.coef[key]
        return s2 + phi * (s2 - s1)

    def upstrpd_coefs(self, key):
        """d(closure)/d(s1, s2, s3): exact, or with the limiter frozen."""
        if self.exact:
            return self.jac[key]
        phi = self.coef[key]
        return (-phi, 1.0 + phi, 0.0)



def _weno_jacobian(s1, s2, s3, s4, s5, weps):
    """d(WENO5 reconstruction)/d(s1..s5) with the WENO-Z weights differentiated (zero where q is clamped)."""
    c13 = float(C13_12)
    d1 = (s1 - 2.0 * s2) + s3
    e1 = (s1 - 4.0 * s2) + 3.0 * s3
    d2 = (s2 - 2.0 * s3) + s4
    e2 = s2 - s4
    d3 = (s3 - 2.0 * s4) + s5
    e3 = (3.0 * s3 - 4.0 * s4) + s5
    b = [c13 * (d1 * d1) + 0.25 * (e1 * e1), c13 * (d2 * d2) + 0.25 * (e2 * e2), c13 * (d3 * d3) + 0.25 * (e3 * e3)]
    z = np.zeros_like(d1)
    db = [[2 * c13 * d1 + 0.5 * e1, -4 * c13 * d1 - 2.0 * e1, 2 * c13 * d1 + 1.5 * e1, z, z],
          [z, 2 * c13 * d2 + 0.5 * e2, -4 * c13 * d2, 2 * c13 * d2 - 0.5 * e2, z],
          [z, z, 2 * c13 * d3 + 1.5 * e3, -4 * c13 * d3 - 2.0 * e3, 2 * c13 * d3 + 0.5 * e3]]
    diff = b[0] - b[2]
    tau = np.abs(diff)
    sg = np.sign(diff)
    dtau = [sg * (db[0][n] - db[2][n]) for n in range(5)]
    a = (float(F(0.1)), float(F(0.6)), float(F(0.3)))
    wv, dw = [], []
    for i in range(3):
        den = b[i] + weps
        with np.errstate(all="ignore"):
            q = tau / den
        live = q < 1.0e30
        q = np.fmin(1.0e30, q)
        wv.append(a[i] * (1.0 + q * q))
        with np.errstate(all="ignore"):
            dq = [np.where(live, dtau[n] / den - tau * db[i][n] / (den * den), 0.0) for n in range(5)]
        dw.append([2.0 * a[i] * q * dq[n] for n in range(5)])
    W = (wv[0] + wv[1]) + wv[2]
    om = [x / W for x in wv]
    p = Tape._parts(s1, s2, s3, s4, s5)
    dp = [(float(C2_6), float(CM7_6), float(C11_6), 0.0, 0.0), (0.0, float(CM1_6), float(C5_6), float(C2_6), 0.0),
          (0.0, 0.0, float(C2_6), float(C5_6), float(CM1_6))]
    jac = []
    for n in range(5):
        dW = (dw[0][n] + dw[1][n]) + dw[2][n]
        g = 0.0
        for i in range(3):
            g = g + om[i] * dp[i][n] + p[i] * ((dw[i][n] - om[i] * dW) / W)
        jac.append(g)
    return tuple(jac)


class Params:
    """The static inputs of the dry periodic step, in binary64 per level."""

    def __init__(self, metrics, base, prof, cfg):
        self.ni, self.nj, self.nk = cfg["ni"], cfg["nj"], cfg["nk"]
        self.dx, self.dy, self.dz, self.dt, self.nsound = cfg["dx"], cfg["dy"], cfg["dz"], cfg["dt"], cfg["nsound"]
        col = lambda n: np.asarray(metrics[n], D)[:, 0, 0]  # noqa: E731
        self.rho, self.rr0, self.rf0, self.rrf0 = col("rho0_3d"), col("rr0"), col("rf0"), col("rrf0")
        self.c1, self.c2 = col("c1"), col("c2")
        self.th0 = np.asarray(base["th0"], D)
        self.pi0 = np.asarray(base["pi0"], D)
        self.qv0 = np.asarray(base["qv0"], D)
        self.rds = np.asarray(prof["rds"], D)
        self.rdsf = np.asarray(prof["rdsf"], D)
        f = lambda x: float(F(x))  # noqa: E731
        self.rdx = float(F(1.0) / F(self.dx))
        self.rdy = float(F(1.0) / F(self.dy))
        self.rdz = float(F(1.0) / F(self.dz))
        self.alph, self.kdiv = f(ALPH), f(KDIV)
        self.beta = float(F(1.0) - F(ALPH))
        self.halfcp = float(F(1005.7) * F(0.5))
        self.cp = f(1005.7)
        rd, rv, cp, cpv = stg.RD, stg.RV, stg.CP, stg.CPV
        self.rth = (F(1.0) / np.asarray(base["th0"], F)).astype(D)
        self.rovcp = float(rd / cp)
        self.reps = float(rv / rd)
        self.repsm1 = float(F(rv / rd) - F(1.0))
        self.g = float(stg.G)
        self.cpm = float(cp + cpv * F(0.0))
        self.cvm = float(F(1.0) / (F(cp - rd) + F(0.0)))
        self.kterm = float((rd + rv * F(0.0)) - F(rd / cp) * F(self.cpm))

    def lev(self, x):
        return x[:, None, None]


def zeros_like_state(p):
    s = (p.nk, p.nj, p.ni)
    return {"u": np.zeros(s), "v": np.zeros(s), "w": np.zeros((p.nk + 1, p.nj, p.ni)), "th": np.zeros(s),
            "pp": np.zeros(s), "ppx": np.zeros(s)}


# --------------------------------------------------------------------------- stage terms, buoyancy

def stage_terms(p, th, pp):
    """thv, dum8, thterm, ppterm, piadv of a dry state (stage.stage_terms)."""
    L = p.lev
    theta = L(p.th0) + th
    b = th * L(p.rth) + p.repsm1 * (0.0 - L(p.qv0))
    b = b - ((0.0 - 0.0) - 0.0)
    thv = (theta * (1.0 + p.reps * 0.0)) / ((1.0 + 0.0) + 0.0)
    thterm = (theta * p.kterm) * p.cvm
    pi = L(p.pi0) + pp
    ppterm = ((pi * p.rovcp) * p.cpm) * p.cvm
    return {"thv": thv, "dum8": p.g * b, "thterm": thterm, "ppterm": ppterm, "piadv": L(p.pi0) + pp}


def stage_terms_T(p, bar):
    """(thb, ppb) from the adjoints of thv, dum8, thterm, ppterm, piadv."""
    L = p.lev
    thb = bar["thv"] * (1.0 / 1.0) + bar["dum8"] * (p.g * L(p.rth)) + bar["thterm"] * (p.kterm * p.cvm)
    ppb = bar["ppterm"] * ((p.rovcp * p.cpm) * p.cvm) + bar["piadv"]
    return thb, ppb


def buoyancy(p, dum8):
    """w-level buoyancy added to wten (levels 2..nk; stage.buoyancy on flat metrics)."""
    out = np.zeros((p.nk + 1, p.nj, p.ni))
    for kf in range(2, p.nk + 1):
        out[kf - 1] = p.c1[kf - 1] * dum8[kf - 2] + p.c2[kf - 1] * dum8[kf - 1]
    return out


def buoyancy_T(p, wtb):
    d8b = np.zeros((p.nk, p.nj, p.ni))
    for kf in range(2, p.nk + 1):
        d8b[kf - 2] += p.c1[kf - 1] * wtb[kf - 1]
        d8b[kf - 1] += p.c2[kf - 1] * wtb[kf - 1]
    return d8b


# --------------------------------------------------------------------------- momentum advection

def mass_fluxes(p, T, u, v, w):
    """rru, rrv, rrw and divx (momentum5.mass_flux_and_divergence on flat periodic metrics)."""
    L = p.lev
    den = L(0.5 * (p.rho + p.rho))
    rru = den * u
    rrv = den * v
    rrw = np.zeros_like(w)
    for kf in range(2, p.nk + 1):
        rrw[kf - 1] = p.rf0[kf - 1] * w[kf - 1]
    d = (ex(rru, 1) - rru) * p.rdx + (ny(rrv, 1) - rrv) * p.rdy
    d = d + (rrw[1:] - rrw[:-1]) * p.rdz
    return rru, rrv, rrw, T.clamp(d, 1.0e-20)


def mass_fluxes_T(p, rrub, rrvb, rrwb, divxb):
    L = p.lev
    rrub, rrvb, rrwb = rrub.copy(), rrvb.copy(), rrwb.copy()
    rrub += ex(divxb * p.rdx, -1) - divxb * p.rdx
    rrvb += ny(divxb * p.rdy, -1) - divxb * p.rdy
    rrwb[1:] += divxb * p.rdz
    rrwb[:-1] -= divxb * p.rdz
    den = L(0.5 * (p.rho + p.rho))
    ub = den * rrub
    vb = den * rrvb
    wb = np.zeros_like(rrwb)
    for kf in range(2, p.nk + 1):
        wb[kf - 1] = p.rf0[kf - 1] * rrwb[kf - 1]
    return ub, vb, wb


def _uvz_branch(kf, nk):
    """(pos, neg) stencils of the u/v vertical flux at w level kf: (kind, levels) with kind 'lin', 'f3', 'f5'."""
    if kf == 2:
        return ("lin", (1, 2)), ("f3", (3, 2, 1))
    if kf == nk:
        return ("f3", (nk - 2, nk - 1, nk)), ("lin", (nk - 1, nk))
    if kf == 3:
        return ("f3", (1, 2, 3)), ("f5", (5, 4, 3, 2, 1))
    if kf == nk - 1:
        return ("f5", (kf - 3, kf - 2, kf - 1, kf, kf + 1)), ("f3", (kf + 1, kf, kf - 1))
    return ("f5", (kf - 3, kf - 2, kf - 1, kf, kf + 1)), ("f5", (kf + 2, kf + 1, kf, kf - 1, kf - 2))


def _wz_branch(k, nk):
    """(pos, neg) stencils of w's vertical flux at scalar level k ('half' = 0.5*(a+b))."""
    if k == 1:
        return ("half", (1, 2)), ("f3", (3, 2, 1))
    if k == 2:
        return ("f3", (1, 2, 3)), ("f5", (5, 4, 3, 2, 1))
    if k == nk - 1:
        return ("f5", (k - 2, k - 1, k, k + 1, k + 2)), ("f3", (k + 2, k + 1, k))
    if k == nk:
        return ("f3", (k - 1, k, k + 1)), ("half", (k, k + 1))
    return ("f5", (k - 2, k - 1, k, k + 1, k + 2)), ("f5", (k + 3, k + 2, k + 1, k, k - 1))


def _eval_branch(br, Q, c2h):
    kind, levs = br
    if kind == "lin":
        return (1.0 - c2h) * Q(levs[0]) + c2h * Q(levs[1])
    if kind == "half":
        return 0.5 * (Q(levs[0]) + Q(levs[1]))
    if kind == "f3":
        return flx3(*(Q(lv) for lv in levs))
    return flx5(*(Q(lv) for lv in levs))


def _branch_T(br, bar, qb, c2h, index):
    """Accumulate the transpose of a branch into qb (levels mapped by ``index``)."""
    kind, levs = br
    if kind == "lin":
        coefs = (1.0 - c2h, c2h)
    elif kind == "half":
        coefs = (0.5, 0.5)
    elif kind == "f3":
        coefs = FLX3
    else:
        coefs = FLX5
    for lv, c in zip(levs, coefs):
        qb[index(lv)] += c * bar


def adv_u(p, T, key, u, rru, rrv, rrw, divx):
    """uten increment of momentum5.momentum_transport5 (advu, periodic flat)."""
    nk = p.nk
    mx = 0.5 * (rru + ex(rru, 1))
    sx = T.mask(key + ".x", mx >= 0.0)
    fx = mx * np.where(sx, flx5(ex(u, -2), ex(u, -1), u, ex(u, 1), ex(u, 2)),
                       flx5(ex(u, 3), ex(u, 2), ex(u, 1), u, ex(u, -1)))
    advx = (-(fx - ex(fx, -1))) * p.rdx
    my = 0.5 * (rrv + ex(rrv, -1))
    sy = T.mask(key + ".y", my >= 0.0)
    fy = my * np.where(sy, flx5(ny(u, -3), ny(u, -2), ny(u, -1), u, ny(u, 1)),
                       flx5(ny(u, 2), ny(u, 1), u, ny(u, -1), ny(u, -2)))
    advy = (-(ny(fy, 1) - fy)) * p.rdy
    pair = rrw + ex(rrw, -1)
    xz = _vertical_uv(p, T, key + ".z", u, pair)
    advz = _advz(p, xz)
    total = (advx + advy) + advz
    total = total + u * (0.5 * (divx + ex(divx, -1)))
    return total * p.lev(1.0 / (0.5 * (p.rho + p.rho)))


def adv_v(p, T, key, v, rru, rrv, rrw, divx):
    """vten increment (advv, periodic flat)."""
    mx = 0.5 * (rru + ny(rru, -1))
    sx = T.mask(key + ".x", mx >= 0.0)
    fx = mx * np.where(sx, flx5(ex(v, -3), ex(v, -2), ex(v, -1), v, ex(v, 1)),
                       flx5(ex(v, 2), ex(v, 1), v, ex(v, -1), ex(v, -2)))
    advx = (-(ex(fx, 1) - fx)) * p.rdx
    my = 0.5 * (rrv + ny(rrv, 1))
    sy = T.mask(key + ".y", my >= 0.0)
    fy = my * np.where(sy, flx5(ny(v, -2), ny(v, -1), v, ny(v, 1), ny(v, 2)),
                       flx5(ny(v, 3), ny(v, 2), ny(v, 1), v, ny(v, -1)))
    advy = (-(fy - ny(fy, -1))) * p.rdy
    pair = rrw + ny(rrw, -1)
    yz = _vertical_uv(p, T, key + ".z", v, pair)
    advz = _advz(p, yz)
    total = (advx + advy) + advz
    total = total + v * (0.5 * (divx + ny(divx, -1)))
    return total * p.lev(1.0 / (0.5 * (p.rho + p.rho)))


def _vertical_uv(p, T, key, q, pair):
    """Vertical fluxes of u or v at w levels 2..nk (index kf-1 of an (nk+1) array)."""
    nk = p.nk
    out = np.zeros((nk + 1,) + q.shape[1:])
    Q = lambda lv: q[lv - 1]  # noqa: E731
    for kf in range(2, nk + 1):
        mass = 0.5 * pair[kf - 1]
        c2h = 0.5 * (p.c2[kf - 1] + p.c2[kf - 1])
        pos, neg = _uvz_branch(kf, nk)
        s = T.mask(f"{key}.{kf}", mass >= 0.0)
        out[kf - 1] = mass * np.where(s, _eval_branch(pos, Q, c2h), _eval_branch(neg, Q, c2h))
    return out


def _advz(p, xz):
    nk = p.nk
    advz = np.zeros((nk,) + xz.shape[1:])
    advz[0] = (-xz[1]) * p.rdz
    advz[nk - 1] = xz[nk - 1] * p.rdz
    for k in range(2, nk):
        advz[k - 1] = (-(xz[k] - xz[k - 1])) * p.rdz
    return advz


def _advz_T(p, advzb):
    nk = p.nk
    xzb = np.zeros((nk + 1,) + advzb.shape[1:])
    xzb[1] -= advzb[0] * p.rdz
    xzb[nk - 1] += advzb[nk - 1] * p.rdz
    for k in range(2, nk):
        xzb[k] -= advzb[k - 1] * p.rdz
        xzb[k - 1] += advzb[k - 1] * p.rdz
    return xzb


def _vertical_uv_T(p, T, key, q, pair, xzb, qb, pairb):
    nk = p.nk
    Q = lambda lv: q[lv - 1]  # noqa: E731
    for kf in range(2, nk + 1):
        mass = 0.5 * pair[kf - 1]
        c2h = 0.5 * (p.c2[kf - 1] + p.c2[kf - 1])
        pos, neg = _uvz_branch(kf, nk)
        s = T.sel[f"{key}.{kf}"]
        r = np.where(s, _eval_branch(pos, Q, c2h), _eval_branch(neg, Q, c2h))
        fb = xzb[kf - 1]
        pairb[kf - 1] += 0.5 * (fb * r)
        rb = fb * mass
        _branch_T(pos, np.where(s, rb, 0.0), qb, c2h, lambda lv: lv - 1)
        _branch_T(neg, np.where(s, 0.0, rb), qb, c2h, lambda lv: lv - 1)


def adv_u_T(p, T, key, u, rru, rrv, rrw, divx, tb):
    """Adjoints (ub, rrub, rrvb, rrwb, divxb) of adv_u from the adjoint tb of its output."""
    totb = tb * p.lev(1.0 / (0.5 * (p.rho + p.rho)))
    ub = totb * (0.5 * (divx + ex(divx, -1)))
    h = 0.5 * (totb * u)
    divxb = h + ex(h, 1)
    # x fluxes
    advxb = totb
    fxb = -(advxb * p.rdx) + ex(advxb * p.rdx, 1)
    mx = 0.5 * (rru + ex(rru, 1))
    sx = T.sel[key + ".x"]
    r = np.where(sx, flx5(ex(u, -2), ex(u, -1), u, ex(u, 1), ex(u, 2)),
                 flx5(ex(u, 3), ex(u, 2), ex(u, 1), u, ex(u, -1)))
    mxb = fxb * r
    rb = fxb * mx
    ub += _st_T(np.where(sx, rb, 0.0), (-2, -1, 0, 1, 2), FLX5, ex)
    ub += _st_T(np.where(sx, 0.0, rb), (3, 2, 1, 0, -1), FLX5, ex)
    rrub = 0.5 * mxb + ex(0.5 * mxb, -1)
    # y fluxes
    fyb = -ny(totb * p.rdy, -1) + totb * p.rdy
    my = 0.5 * (rrv + ex(rrv, -1))
    sy = T.sel[key + ".y"]
    r = np.where(sy, flx5(ny(u, -3), ny(u, -2), ny(u, -1), u, ny(u, 1)),
                 flx5(ny(u, 2), ny(u, 1), u, ny(u, -1), ny(u, -2)))
    myb = fyb * r
    rb = fyb * my
    ub += _st_T(np.where(sy, rb, 0.0), (-3, -2, -1, 0, 1), FLX5, ny)
    ub += _st_T(np.where(sy, 0.0, rb), (2, 1, 0, -1, -2), FLX5, ny)
    rrvb = 0.5 * myb + ex(0.5 * myb, 1)
    # z fluxes
    pair = rrw + ex(rrw, -1)
    pairb = np.zeros_like(rrw)
    _vertical_uv_T(p, T, key + ".z", u, pair, _advz_T(p, totb), ub, pairb)
    rrwb = pairb + ex(pairb, 1)
    return ub, rrub, rrvb, rrwb, divxb


def adv_v_T(p, T, key, v, rru, rrv, rrw, divx, tb):
    totb = tb * p.lev(1.0 / (0.5 * (p.rho + p.rho)))
    vb = totb * (0.5 * (divx + ny(divx, -1)))
    h = 0.5 * (totb * v)
    divxb = h + ny(h, 1)
    fxb = -ex(totb * p.rdx, -1) + totb * p.rdx
    mx = 0.5 * (rru + ny(rru, -1))
    sx = T.sel[key + ".x"]
    r = np.where(sx, flx5(ex(v, -3), ex(v, -2), ex(v, -1), v, ex(v, 1)),
                 flx5(ex(v, 2), ex(v, 1), v, ex(v, -1), ex(v, -2)))
    mxb = fxb * r
    rb = fxb * mx
    vb += _st_T(np.where(sx, rb, 0.0), (-3, -2, -1, 0, 1), FLX5, ex)
    vb += _st_T(np.where(sx, 0.0, rb), (2, 1, 0, -1, -2), FLX5, ex)
    rrub = 0.5 * mxb + ny(0.5 * mxb, 1)
    fyb = -(totb * p.rdy) + ny(totb * p.rdy, 1)
    my = 0.5 * (rrv + ny(rrv, 1))
    sy = T.sel[key + ".y"]
    r = np.where(sy, flx5(ny(v, -2), ny(v, -1), v, ny(v, 1), ny(v, 2)),
                 flx5(ny(v, 3), ny(v, 2), ny(v, 1), v, ny(v, -1)))
    myb = fyb * r
    rb = fyb * my
    vb += _st_T(np.where(sy, rb, 0.0), (-2, -1, 0, 1, 2), FLX5, ny)
    vb += _st_T(np.where(sy, 0.0, rb), (3, 2, 1, 0, -1), FLX5, ny)
    rrvb = 0.5 * myb + ny(0.5 * myb, -1)
    pair = rrw + ny(rrw, -1)
    pairb = np.zeros_like(rrw)
    _vertical_uv_T(p, T, key + ".z", v, pair, _advz_T(p, totb), vb, pairb)
    rrwb = pairb + ny(pairb, 1)
    return vb, rrub, rrvb, rrwb, divxb


def adv_w(p, T, key, w, rru, rrv, rrw, divx):
    """wten increment at w levels 2..nk (advw, periodic flat); (nk+1) array."""
    nk = p.nk
    out = np.zeros_like(w)
    wz = np.zeros((nk + 1,) + w.shape[1:])  # index k-1: flux at scalar level k = 1..nk
    Q = lambda lv: w[lv - 1]  # noqa: E731
    for k in range(1, nk + 1):
        mass = 0.5 * (rrw[k - 1] + rrw[k])
        pos, neg = _wz_branch(k, nk)
        s = T.mask(f"{key}.wz.{k}", mass >= 0.0)
        wz[k - 1] = mass * np.where(s, _eval_branch(pos, Q, None), _eval_branch(neg, Q, None))
    for kf in range(2, nk + 1):
        c2h = 0.5 * (p.c2[kf - 1] + p.c2[kf - 1])
        c1h = 1.0 - c2h
        W = w[kf - 1]
        mx = c2h * rru[kf - 1] + c1h * rru[kf - 2]
        sx = T.mask(f"{key}.x.{kf}", mx >= 0.0)
        fx = mx * np.where(sx, flx5(ex(W, -3), ex(W, -2), ex(W, -1), W, ex(W, 1)),
                           flx5(ex(W, 2), ex(W, 1), W, ex(W, -1), ex(W, -2)))
        my = c2h * rrv[kf - 1] + c1h * rrv[kf - 2]
        sy = T.mask(f"{key}.y.{kf}", my >= 0.0)
        fy = my * np.where(sy, flx5(ny(W, -3), ny(W, -2), ny(W, -1), W, ny(W, 1)),
                           flx5(ny(W, 2), ny(W, 1), W, ny(W, -1), ny(W, -2)))
        advx = (-(ex(fx, 1) - fx)) * (p.rdx * 1.0)
        advy = (-(ny(fy, 1) - fy)) * (p.rdy * 1.0)
        advz = (-(wz[kf - 1] - wz[kf - 2])) * p.rdz
        cc2 = p.c2[kf - 1]
        cc1 = 1.0 - cc2
        total = (advx + advy) + advz
        total = total + W * (cc2 * divx[kf - 1] + cc1 * divx[kf - 2])
        out[kf - 1] = total * (p.rrf0[kf - 1] * 1.0)
    return out


def adv_w_T(p, T, key, w, rru, rrv, rrw, divx, tb):
    nk = p.nk
    wb = np.zeros_like(w)
    rrub = np.zeros_like(rru)
    rrvb = np.zeros_like(rrv)
    rrwb = np.zeros_like(rrw)
    divxb = np.zeros_like(divx)
    wzb = np.zeros((nk + 1,) + w.shape[1:])
    for kf in range(2, nk + 1):
        c2h = 0.5 * (p.c2[kf - 1] + p.c2[kf - 1])
        c1h = 1.0 - c2h
        W = w[kf - 1]
        cc2 = p.c2[kf - 1]
        cc1 = 1.0 - cc2
        totb = tb[kf - 1] * (p.rrf0[kf - 1] * 1.0)
        wb[kf - 1] += totb * (cc2 * divx[kf - 1] + cc1 * divx[kf - 2])
        h = totb * W
        divxb[kf - 1] += cc2 * h
        divxb[kf - 2] += cc1 * h
        wzb[kf - 1] -= totb * p.rdz
        wzb[kf - 2] += totb * p.rdz
        fxb = -ex(totb * p.rdx, -1) + totb * p.rdx
        fyb = -ny(totb * p.rdy, -1) + totb * p.rdy
        mx = c2h * rru[kf - 1] + c1h * rru[kf - 2]
        sx = T.sel[f"{key}.x.{kf}"]
        r = np.where(sx, flx5(ex(W, -3), ex(W, -2), ex(W, -1), W, ex(W, 1)),
                     flx5(ex(W, 2), ex(W, 1), W, ex(W, -1), ex(W, -2)))
        mxb = fxb * r
        rb = fxb * mx
        wb[kf - 1] += _st_T(np.where(sx, rb, 0.0), (-3, -2, -1, 0, 1), FLX5, ex)
        wb[kf - 1] += _st_T(np.where(sx, 0.0, rb), (2, 1, 0, -1, -2), FLX5, ex)
        rrub[kf - 1] += c2h * mxb
        rrub[kf - 2] += c1h * mxb
        my = c2h * rrv[kf - 1] + c1h * rrv[kf - 2]
        sy = T.sel[f"{key}.y.{kf}"]
        r = np.where(sy, flx5(ny(W, -3), ny(W, -2), ny(W, -1), W, ny(W, 1)),
                     flx5(ny(W, 2), ny(W, 1), W, ny(W, -1), ny(W, -2)))
        myb = fyb * r
        rb = fyb * my
        wb[kf - 1] += _st_T(np.where(sy, rb, 0.0), (-3, -2, -1, 0, 1), FLX5, ny)
        wb[kf - 1] += _st_T(np.where(sy, 0.0, rb), (2, 1, 0, -1, -2), FLX5, ny)
        rrvb[kf - 1] += c2h * myb
        rrvb[kf - 2] += c1h * myb
    Q = lambda lv: w[lv - 1]  # noqa: E731
    for k in range(1, nk + 1):
        mass = 0.5 * (rrw[k - 1] + rrw[k])
        pos, neg = _wz_branch(k, nk)
        s = T.sel[f"{key}.wz.{k}"]
        r = np.where(s, _eval_branch(pos, Q, None), _eval_branch(neg, Q, None))
        fb = wzb[k - 1]
        mb = 0.5 * (fb * r)
        rrwb[k - 1] += mb
        rrwb[k] += mb
        rb = fb * mass
        _branch_T(pos, np.where(s, rb, 0.0), wb, None, lambda lv: lv - 1)
        _branch_T(neg, np.where(s, 0.0, rb), wb, None, lambda lv: lv - 1)
    return wb, rrub, rrvb, rrwb, divxb


# --------------------------------------------------------------------------- acoustic stage

def coefficients(p, dts, thv, ppterm, piadv):
    """qk, pk1, pk2, mm at levels 1..nk (sound.prepare_coefficients on flat metrics)."""
    nk = p.nk
    qk = np.zeros_like(thv)
    pk1 = np.zeros_like(thv)
    pk2 = np.zeros_like(thv)
    mm = np.zeros_like(thv)
    for k in range(1, nk + 1):
        tem = dts * p.rdsf[k]
        qk[k - 1] = (tem * ppterm[k - 1]) * 1.0
        tem2 = ((dts * 1.0) * p.rdsf[k]) * p.rr0[k - 1]
        if k < nk:
            b2 = (-p.c2[k]) * piadv[k] + (1.0 - p.c1[k]) * piadv[k - 1]
            pk2[k - 1] = (tem2 * p.rf0[k]) * b2
        if k > 1:
            b1 = p.c1[k - 1] * piadv[k - 2] + (p.c2[k - 1] - 1.0) * piadv[k - 1]
            pk1[k - 1] = (tem2 * p.rf0[k - 1]) * b1
            tem1 = (dts * p.rds[k]) * p.cp
            th = p.c2[k - 1] * thv[k - 1] + p.c1[k - 1] * thv[k - 2]
            mm[k - 1] = (tem1 * th) * 1.0
    return qk, pk1, pk2, mm


def coefficients_T(p, dts, qkb, pk1b, pk2b, mmb):
    nk = p.nk
    thvb = np.zeros_like(qkb)
    ppb = np.zeros_like(qkb)
    piab = np.zeros_like(qkb)
    for k in range(1, nk + 1):
        tem = dts * p.rdsf[k]
        ppb[k - 1] += tem * qkb[k - 1]
        tem2 = ((dts * 1.0) * p.rdsf[k]) * p.rr0[k - 1]
        if k < nk:
            b2b = (tem2 * p.rf0[k]) * pk2b[k - 1]
            piab[k] += (-p.c2[k]) * b2b
            piab[k - 1] += (1.0 - p.c1[k]) * b2b
        if k > 1:
            b1b = (tem2 * p.rf0[k - 1]) * pk1b[k - 1]
            piab[k - 2] += p.c1[k - 1] * b1b
            piab[k - 1] += (p.c2[k - 1] - 1.0) * b1b
            thb = ((dts * p.rds[k]) * p.cp) * mmb[k - 1]
            thvb[k - 1] += p.c2[k - 1] * thb
            thvb[k - 2] += p.c1[k - 1] * thb
    return thvb, ppb, piab


def _wlo(w, nk):
    """w(k) at scalar levels k = 1..nk, zero at k = 1."""
    out = w[:nk].copy()
    out[0] = 0.0
    return out


def _whi(w, nk):
    """w(k+1) at scalar levels k = 1..nk, zero at k = nk."""
    out = w[1 : nk + 1].copy()
    out[nk - 1] = 0.0
    return out


def _thomas_rows(p, qk, pk1, pk2, mm):
    """aa, bb, cc of rows k = 2..nk (index k-2)."""
    nk = p.nk
    asq = float(F(p.alph) * F(p.alph))
    rows = []
    for k in range(2, nk + 1):
        aa = (-asq) * (mm[k - 1] * (pk1[k - 2] + qk[k - 2])) if k >= 3 else np.zeros_like(mm[0])
        cc = (-asq) * ((-mm[k - 1]) * (pk2[k - 1] - qk[k - 1])) if k < nk else np.zeros_like(mm[0])
        diag = (-(pk1[k - 1] + qk[k - 1])) + (pk2[k - 2] - qk[k - 2])
        bb = 1.0 - asq * (mm[k - 1] * diag)
        rows.append((aa, bb, cc))
    return rows, asq


def _tridiag_solve(lo, dg, up, rhs):
    """x with lo[r] x[r-1] + dg[r] x[r] + up[r] x[r+1] = rhs[r] (the reference's elimination order)."""
    n = len(dg)
    ec = [None] * n
    ed = [None] * n
    x = [None] * n
    for r in range(n):
        piv = dg[r] if r == 0 else lo[r] * ec[r - 1] + dg[r]
        inv = 1.0 / piv
        if r < n - 1:
            ec[r] = (-up[r]) * inv
            ed[r] = (rhs[r] if r == 0 else rhs[r] - lo[r] * ed[r - 1]) * inv
        else:
            x[r] = (rhs[r] - lo[r] * ed[r - 1]) * inv if r > 0 else rhs[r] * inv
    for r in range(n - 2, -1, -1):
        x[r] = ec[r] * x[r + 1] + ed[r]
    return x


def small_step(p, T, key, s, c, last, nrk, nloop, dts):
    """One small step (sound.acoustic_small_step, flat periodic); ``s`` is updated in place."""
    nk = p.nk
    L = p.lev
    thv, ppterm, pia = c["thv"], c["ppterm"], c["piadv"]
    qk, pk1, pk2, mm = c["qk"], c["pk1"], c["pk2"], c["mm"]
    alph, beta, kdiv = p.alph, p.beta, p.kdiv
    if T.record:
        T.traj[key] = {n: s[n].copy() for n in ("u", "v", "w", "pp", "ppd")}
    ppd = s["ppd"]
    pgx = (p.halfcp * (thv + ex(thv, -1))) * ((ppd - ex(ppd, -1)) * p.rdx)
    u = s["u"] + dts * (s["uten"] - pgx)
    pgy = (p.halfcp * (thv + ny(thv, -1))) * ((ppd - ny(ppd, -1)) * p.rdy)
    v = s["v"] + dts * (s["vten"] - pgy)
    hz = (ex(u, 1) - u) * p.rdx + (ny(v, 1) - v) * p.rdy
    div = T.clamp(1.0 * (hz + 0.0), float(F(SMEPS)))
    w = s["w"]
    src = dts * (s["ppten"] - ppterm * div)
    qp = (-qk) * (_whi(w, nk) - _wlo(w, nk))
    fpk = src + beta * qp
    xu = L(p.rho) * u
    yv = L(p.rho) * v
    temx = (dts * 0.5) * p.rdx
    temy = (dts * 0.5) * p.rdy
    ww2 = 0.0 + beta * w[1 : nk + 1]
    ww1 = 0.0 + beta * _wlo(w, nk)
    xa = (temx * ex(xu, 1)) * (ex(pia, 1) - pia) + (temx * xu) * (pia - ex(pia, -1))
    ya = (temy * ny(yv, 1)) * (ny(pia, 1) - pia) + (temy * yv) * (pia - ny(pia, -1))
    hz2 = ((-(xa + ya)) * L(p.rr0)) * 1.0
    fpk = fpk + (hz2 + (ww1 * pk1 + ww2 * pk2))
    rows, asq = _thomas_rows(p, qk, pk1, pk2, mm)
    pp = s["pp"]
    rhs = []
    for k in range(2, nk + 1):
        pdelta = pp[k - 1] - pp[k - 2]
        fwk = dts * s["wten"][k - 1] + beta * ((-mm[k - 1]) * pdelta)
        udelta = (pp[k - 1] + fpk[k - 1]) - (pp[k - 2] + fpk[k - 2])
        rhs.append((w[k - 1] + fwk) + alph * ((-mm[k - 1]) * udelta))
    x = _tridiag_solve([r[0] for r in rows], [r[1] for r in rows], [r[2] for r in rows], rhs)
    wn = w.copy()
    wn[0] = 0.0
    for k in range(2, nk + 1):
        wn[k - 1] = x[k - 2]
    vt = _wlo(wn, nk) * (pk1 + qk) + _whi(wn, nk) * (pk2 - qk)
    up = T.clamp((pp + fpk) + alph * vt, float(F(SMEPS)))
    tend = kdiv * (up - pp)
    if T.record:
        T.traj[key].update({"u1": u, "v1": v, "div": div, "fpk": fpk, "wn": wn})
    s["ppd"] = up + tend
    s["pp"] = up
    if nrk == 3 and last:
        s["ppx"] = tend
    tavg = float(F(1.0) / F(nloop))
    rw = s["rrw"].copy()
    if not last:
        s["rru"] = s["rru"] + u
        s["rrv"] = s["rrv"] + v
        rw[:nk] = rw[:nk] + wn[:nk]
    else:
        s["rru"] = (s["rru"] + u) * tavg
        s["rrv"] = (s["rrv"] + v) * tavg
        rw[:nk] = (rw[:nk] + wn[:nk] * alph) * tavg
    s["rrw"] = rw
    s["u"], s["v"], s["w"] = u, v, wn


def small_step_T(p, T, key, b, c, cb, last, nrk, nloop, dts):
    """Transpose of small_step: ``b`` holds the adjoints of the step's outputs and becomes those of its inputs;
    ``cb`` accumulates the adjoints of uten, vten, wten, ppten, thv, ppterm, piadv, qk, pk1, pk2, mm."""
    nk = p.nk
    L = p.lev
    tr = T.traj[key]
    thv, ppterm, pia = c["thv"], c["ppterm"], c["piadv"]
    qk, pk1, pk2, mm = c["qk"], c["pk1"], c["pk2"], c["mm"]
    alph, beta, kdiv = p.alph, p.beta, p.kdiv
    w, pp, ppd = tr["w"], tr["pp"], tr["ppd"]
    u1, v1, div, fpk, wn = tr["u1"], tr["v1"], tr["div"], tr["fpk"], tr["wn"]
    tavg = float(F(1.0) / F(nloop))
    # averages
    ub = b["u"].copy()
    vb = b["v"].copy()
    wnb = b["w"].copy()
    rwb = b["rrw"].copy()
    if last:
        ub += b["rru"] * tavg
        vb += b["rrv"] * tavg
        wnb[:nk] += (rwb[:nk] * tavg) * alph
        b["rru"] = b["rru"] * tavg
        b["rrv"] = b["rrv"] * tavg
        rwb[:nk] = rwb[:nk] * tavg
    else:
        ub += b["rru"]
        vb += b["rrv"]
        wnb[:nk] += rwb[:nk]
    b["rrw"] = rwb
    # pressure update and filter
    upb = b["pp"] + (1.0 + kdiv) * b["ppd"]
    ppb = -kdiv * b["ppd"]
    if nrk == 3 and last:
        upb = upb + kdiv * b["ppx"]
        ppb = ppb - kdiv * b["ppx"]
        b["ppx"] = np.zeros_like(b["ppx"])
    ppb = ppb + upb
    fpkb = upb.copy()
    vtb = alph * upb
    wlo, whi = _wlo(wn, nk), _whi(wn, nk)
    cb["pk1"] += vtb * wlo
    cb["qk"] += vtb * (wlo - whi)
    cb["pk2"] += vtb * whi
    wnb[1:nk] += (vtb * (pk1 + qk))[1:nk]
    wnb[1:nk] += (vtb * (pk2 - qk))[0 : nk - 1]
    # Thomas solve: wn(2..nk) = T^-1 rhs
    rows, asq = _thomas_rows(p, qk, pk1, pk2, mm)
    aa = [r[0] for r in rows]
    bb = [r[1] for r in rows]
    cc = [r[2] for r in rows]
    nrow = nk - 1
    xb = [wnb[k - 1].copy() for k in range(2, nk + 1)]
    lam = _tridiag_solve([cc[r - 1] if r > 0 else np.zeros_like(bb[0]) for r in range(nrow)], bb,
                         [aa[r + 1] if r < nrow - 1 else np.zeros_like(bb[0]) for r in range(nrow)], xb)
    wb = np.zeros_like(w)
    wb[nk] += wnb[nk]  # the lid level passes through
    for r, k in enumerate(range(2, nk + 1)):
        lb = lam[r]
        x_k = wn[k - 1]
        # rhs(k) = (w(k) + fwk) + alph*((-mm)*udelta), fwk = dts*wten + beta*((-mm)*pdelta)
        pdelta = pp[k - 1] - pp[k - 2]
        udelta = (pp[k - 1] + fpk[k - 1]) - (pp[k - 2] + fpk[k - 2])
        wb[k - 1] += lb
        cb["wten"][k - 1] += dts * lb
        cb["mm"][k - 1] += (-(beta * pdelta + alph * udelta)) * lb
        pdb = beta * (-mm[k - 1]) * lb
        udb = alph * (-mm[k - 1]) * lb
        ppb[k - 1] += pdb + udb
        ppb[k - 2] -= pdb + udb
        fpkb[k - 1] += udb
        fpkb[k - 2] -= udb
        # dT: aa x(k-1) + bb x(k) + cc x(k+1)
        bbb = -lb * x_k
        diag = (-(pk1[k - 1] + qk[k - 1])) + (pk2[k - 2] - qk[k - 2])
        cb["mm"][k - 1] += (-asq * diag) * bbb
        diagb = (-asq * mm[k - 1]) * bbb
        cb["pk1"][k - 1] -= diagb
        cb["qk"][k - 1] -= diagb
        cb["pk2"][k - 2] += diagb
        cb["qk"][k - 2] -= diagb
        if k >= 3:
            aab = -lb * wn[k - 2]
            cb["mm"][k - 1] += (-asq) * (pk1[k - 2] + qk[k - 2]) * aab
            cb["pk1"][k - 2] += (-asq) * mm[k - 1] * aab
            cb["qk"][k - 2] += (-asq) * mm[k - 1] * aab
        if k < nk:
            ccb = -lb * wn[k]
            cb["mm"][k - 1] += asq * (pk2[k - 1] - qk[k - 1]) * ccb
            cb["pk2"][k - 1] += asq * mm[k - 1] * ccb
            cb["qk"][k - 1] -= asq * mm[k - 1] * ccb
    # pressure advection: fpk += hz2 + (ww1*pk1 + ww2*pk2)
    cb["pk1"] += fpkb * (beta * _wlo(w, nk))
    cb["pk2"] += fpkb * (beta * w[1 : nk + 1])
    wb[1:nk] += (beta * (fpkb * pk1))[1:nk]
    wb[1 : nk + 1] += beta * (fpkb * pk2)
    hz2b = fpkb * L(p.rr0)
    xab = -hz2b
    yab = -hz2b
    temx = (dts * 0.5) * p.rdx
    temy = (dts * 0.5) * p.rdy
    xu = L(p.rho) * u1
    yv = L(p.rho) * v1
    t2 = xab * (temx * ex(xu, 1))
    t1 = xab * (temx * xu)
    cb["piadv"] += ex(t2, -1) - t2 + t1 - ex(t1, 1)
    xub = ex(temx * (xab * (ex(pia, 1) - pia)), -1) + temx * (xab * (pia - ex(pia, -1)))
    t2 = yab * (temy * ny(yv, 1))
    t1 = yab * (temy * yv)
    cb["piadv"] += ny(t2, -1) - t2 + t1 - ny(t1, 1)
    yvb = ny(temy * (yab * (ny(pia, 1) - pia)), -1) + temy * (yab * (pia - ny(pia, -1)))
    ub += L(p.rho) * xub
    vb += L(p.rho) * yvb
    # fpk = dts*(ppten - ppterm*div) + beta*((-qk)*(whi - wlo))
    qpb = beta * fpkb
    cb["qk"] -= qpb * (_whi(w, nk) - _wlo(w, nk))
    wb[1 : nk + 1] += (-qk * qpb) * np.r_[np.ones(nk - 1), 0.0][:, None, None]
    wb[1:nk] += (qk * qpb)[1:nk]
    cb["ppten"] += dts * fpkb
    cb["ppterm"] -= dts * (div * fpkb)
    divb = -dts * (ppterm * fpkb)
    ub += ex(divb * p.rdx, -1) - divb * p.rdx
    vb += ny(divb * p.rdy, -1) - divb * p.rdy
    # u1 = u + dts*(uten - pgx)
    cb["uten"] += dts * ub
    cb["vten"] += dts * vb
    pgxb = -dts * ub
    pgyb = -dts * vb
    A = p.halfcp * (thv + ex(thv, -1))
    B = (ppd - ex(ppd, -1)) * p.rdx
    t = p.halfcp * (pgxb * B)
    cb["thv"] += t + ex(t, 1)
    t = p.rdx * (A * pgxb)
    ppdb = t - ex(t, 1)
    A = p.halfcp * (thv + ny(thv, -1))
    B = (ppd - ny(ppd, -1)) * p.rdy
    t = p.halfcp * (pgyb * B)
    cb["thv"] += t + ny(t, 1)
    t = p.rdy * (A * pgyb)
    ppdb += t - ny(t, 1)
    b["u"], b["v"], b["w"], b["pp"], b["ppd"] = ub, vb, wb, ppb, ppdb


def acoustic_stage(p, T, key, nrk, inp):
    """One acoustic RK stage (sound.cpu_psolver3_stage_inplace, flat periodic)."""
    nk = p.nk
    nloop, dts_v = sound3_stage_schedule_f32(p.dt, p.nsound, nrk)
    dts = float(F(dts_v))
    c = {n: inp[n] for n in ("thv", "ppterm", "piadv")}
    c["qk"], c["pk1"], c["pk2"], c["mm"] = coefficients(p, dts, c["thv"], c["ppterm"], c["piadv"])
    s = {n: inp[n] for n in ("uten", "vten", "wten", "ppten")}
    s["ppd"] = inp["ppx"].copy()
    s["ppx"] = inp["ppx"]
    if nrk != 1:
        s["pp"] = inp["ppi"].copy()
        s["u"], s["v"] = inp["ua"].copy(), inp["va"].copy()
        s["w"] = inp["w"].copy()
        s["w"][:nk] = inp["wa"][:nk]
    else:
        s["pp"], s["u"], s["v"], s["w"] = (inp[n].copy() for n in ("pp", "u", "v", "w"))
    s["rru"] = np.zeros_like(s["u"])
    s["rrv"] = np.zeros_like(s["v"])
    s["rrw"] = np.zeros_like(s["w"])
    s["rrw"][:nk] = inp["wa"][:nk] * p.beta
    for n in range(1, nloop + 1):
        small_step(p, T, f"{key}.ss{n}", s, c, n == nloop, nrk, nloop, dts)
    return s


def acoustic_stage_T(p, T, key, nrk, inp, bar):
    """Adjoints of the acoustic stage's inputs from those of its outputs (u v w pp ppd ppx rru rrv rrw)."""
    nk = p.nk
    nloop, dts_v = sound3_stage_schedule_f32(p.dt, p.nsound, nrk)
    dts = float(F(dts_v))
    c = {n: inp[n] for n in ("thv", "ppterm", "piadv")}
    c["qk"], c["pk1"], c["pk2"], c["mm"] = coefficients(p, dts, c["thv"], c["ppterm"], c["piadv"])
    cb = {n: np.zeros_like(inp[n]) for n in ("uten", "vten", "wten", "ppten", "thv", "ppterm", "piadv")}
    for n in ("qk", "pk1", "pk2", "mm"):
        cb[n] = np.zeros_like(c["qk"])
    b = {n: bar[n].copy() for n in ("u", "v", "w", "pp", "ppd", "ppx", "rru", "rrv", "rrw")}
    ppx_pass = b["ppx"] if nrk != 3 else None
    for n in range(nloop, 0, -1):
        small_step_T(p, T, f"{key}.ss{n}", b, c, cb, n == nloop, nrk, nloop, dts)
    thvb, ppterb, piab = coefficients_T(p, dts, cb["qk"], cb["pk1"], cb["pk2"], cb["mm"])
    out = {"uten": cb["uten"], "vten": cb["vten"], "wten": cb["wten"], "ppten": cb["ppten"],
           "thv": cb["thv"] + thvb, "ppterm": cb["ppterm"] + ppterb, "piadv": cb["piadv"] + piab}
    out["ppx"] = b["ppd"] + (ppx_pass if ppx_pass is not None else 0.0)
    wa = np.zeros_like(b["w"])
    wa[:nk] += b["rrw"][:nk] * p.beta
    zero = lambda n: np.zeros_like(b[n])  # noqa: E731
    if nrk != 1:
        out["ppi"], out["ua"], out["va"] = b["pp"], b["u"], b["v"]
        wa[:nk] += b["w"][:nk]
        wpass = np.zeros_like(b["w"])
        wpass[nk] = b["w"][nk]
        out["w"], out["pp"], out["u"], out["v"] = wpass, zero("pp"), zero("u"), zero("v")
    else:
        out["ppi"], out["ua"], out["va"] = zero("pp"), zero("u"), zero("v")
        out["w"], out["pp"], out["u"], out["v"] = b["w"], b["pp"], b["u"], b["v"]
    out["wa"] = wa
    return out


# --------------------------------------------------------------------------- post-acoustic and scalar RK

def getdiv(p, T, rru, rrv, rrw, thterm):
    """div and theta_div of the stage-averaged winds (stage.getdiv, flat)."""
    nk = p.nk
    wt = np.zeros_like(rrw)
    for kf in range(2, nk + 1):
        wt[kf - 1] = rrw[kf - 1] + 0.0
    d = (ex(rru * 1.0, 1) - rru * 1.0) * p.rdx + (ny(rrv * 1.0, 1) - rrv * 1.0) * p.rdy
    d = d + (wt[1:] - wt[:-1]) * p.rdz
    d = T.clamp(d, float(stg.SMEPS))
    return d, d * thterm


def getdiv_T(p, db, rrwb_shape):
    nk = p.nk
    rrub = ex(db * p.rdx, -1) - db * p.rdx
    rrvb = ny(db * p.rdy, -1) - db * p.rdy
    wtb = np.zeros(rrwb_shape)
    wtb[1:] += db * p.rdz
    wtb[:-1] -= db * p.rdz
    rrwb = np.zeros(rrwb_shape)
    rrwb[1:nk] = wtb[1:nk]
    return rrub, rrvb, rrwb


def getdivx(p, T, rru, rrv, rrw):
    """Mass fluxes of the averaged winds and divx (stage.getdivx, flat)."""
    nk = p.nk
    L = p.lev
    den = L(0.5 * (p.rho + p.rho))
    mu = (den * rru) * 1.0
    mv = (den * rrv) * 1.0
    mw = np.zeros_like(rrw)
    for kf in range(2, nk + 1):
        mw[kf - 1] = p.rf0[kf - 1] * rrw[kf - 1] + 0.0
    d = (ex(mu, 1) - mu) * p.rdx + (ny(mv, 1) - mv) * p.rdy
    d = d + (mw[1:] - mw[:-1]) * p.rdz
    return mu, mv, mw, T.clamp(d, float(stg.SMEPS))


def getdivx_T(p, mub, mvb, mwb, divxb):
    nk = p.nk
    L = p.lev
    mub = mub + ex(divxb * p.rdx, -1) - divxb * p.rdx
    mvb = mvb + ny(divxb * p.rdy, -1) - divxb * p.rdy
    mwb = mwb.copy()
    mwb[1:] += divxb * p.rdz
    mwb[:-1] -= divxb * p.rdz
    den = L(0.5 * (p.rho + p.rho))
    rrwb = np.zeros_like(mwb)
    for kf in range(2, nk + 1):
        rrwb[kf - 1] = p.rf0[kf - 1] * mwb[kf - 1]
    return den * mub, den * mvb, rrwb


def _zflux_kind(k, nk, weno):
    """(pos kind, pos levels, neg kind, neg levels, strict) of the scalar z flux at face k."""
    up3 = "upd" if weno else "f3"
    r5 = "weno" if weno else "f5"
    if k == 2:
        return ("lin", (1, 2)), (up3, (3, 2, 1)), False
    if k == nk:
        return (up3, (nk - 2, nk - 1, nk)), ("lin", (nk - 1, nk)), weno
    if k == 3:
        return (up3, (1, 2, 3)), (r5, (5, 4, 3, 2, 1)), False
    if k == nk - 1:
        return (r5, (k - 3, k - 2, k - 1, k, k + 1)), (up3, (k + 1, k, k - 1)), False
    return (r5, (k - 3, k - 2, k - 1, k, k + 1)), (r5, (k + 2, k + 1, k, k - 1, k - 2)), False


def scalar_adv(p, T, key, sc, mu, mv, mw, divx, weno):
    """Advective tendency of a padded-free scalar (advection5.cpu_scalar_advection5_inplace, flat periodic)."""
    nk = p.nk
    weps = THETA_WEPS

    def recon(k2, a, b_, c_, d, e):
        return T.weno(k2, a, b_, c_, d, e, weps) if weno else flx5(a, b_, c_, d, e)

    sx = T.mask(key + ".x", mu >= 0.0)
    fx = mu * np.where(sx, recon(key + ".xp", ex(sc, -3), ex(sc, -2), ex(sc, -1), sc, ex(sc, 1)),
                       recon(key + ".xn", ex(sc, 2), ex(sc, 1), sc, ex(sc, -1), ex(sc, -2)))
    sy = T.mask(key + ".y", mv >= 0.0)
    fy = mv * np.where(sy, recon(key + ".yp", ny(sc, -3), ny(sc, -2), ny(sc, -1), sc, ny(sc, 1)),
                       recon(key + ".yn", ny(sc, 2), ny(sc, 1), sc, ny(sc, -1), ny(sc, -2)))
    fz = np.zeros_like(mw)
    S = lambda lv: sc[lv - 1]  # noqa: E731
    for k in range(2, nk + 1):
        mass = mw[k - 1]
        pos, neg, strict = _zflux_kind(k, nk, weno)
        vals = []
        for tag, (kind, levs) in (("p", pos), ("n", neg)):
            if kind == "lin":
                vals.append(p.c1[k - 1] * S(levs[0]) + p.c2[k - 1] * S(levs[1]))
            elif kind == "upd":
                vals.append(T.upstrpd(f"{key}.z{k}{tag}", *(S(lv) for lv in levs), weps))
            elif kind == "f3":
                vals.append(flx3(*(S(lv) for lv in levs)))
            elif kind == "weno":
                vals.append(T.weno(f"{key}.z{k}{tag}", *(S(lv) for lv in levs), weps))
            else:
                vals.append(flx5(*(S(lv) for lv in levs)))
        sel = T.mask(f"{key}.z{k}", (mass > 0.0) if strict else (mass >= 0.0))
        fz[k - 1] = mass * np.where(sel, vals[0], vals[1])
    ax = (fx - ex(fx, 1)) * p.rdx
    ay = (fy - ny(fy, 1)) * p.rdy
    az = np.zeros_like(sc)
    az[0] = (-fz[1]) * p.rdz
    az[nk - 1] = fz[nk - 1] * p.rdz
    for k in range(2, nk):
        az[k - 1] = (fz[k - 1] - fz[k]) * p.rdz
    total = (ax + ay) + az
    total = total + sc * divx
    return (total * p.lev(p.rr0)) * 1.0


def _recon_T(T, key, bar, offs, shift, weno, out):
    coefs = T.weno_coefs(key) if weno else FLX5
    for o, c in zip(offs, coefs):
        out += shift(c * bar, -o)


def scalar_adv_T(p, T, key, sc, mu, mv, mw, divx, weno, advb):
    """Adjoints (scb, mub, mvb, mwb, divxb) of scalar_adv."""
    nk = p.nk
    weps = THETA_WEPS
    totb = advb * p.lev(p.rr0)
    scb = totb * divx
    divxb = totb * sc
    fxb = totb * p.rdx - ex(totb * p.rdx, -1)
    fyb = totb * p.rdy - ny(totb * p.rdy, -1)
    fzb = np.zeros_like(mw)
    fzb[1] -= totb[0] * p.rdz
    fzb[nk - 1] += totb[nk - 1] * p.rdz
    for k in range(2, nk):
        fzb[k - 1] += totb[k - 1] * p.rdz
        fzb[k] -= totb[k - 1] * p.rdz

    def rc(k2, a, b_, c_, d, e):
        if weno:
            o1, o2, o3 = T.coef[k2]
            p1, p2, p3 = Tape._parts(a, b_, c_, d, e)
            return (o1 * p1 + o2 * p2) + o3 * p3
        return flx5(a, b_, c_, d, e)

    sx = T.sel[key + ".x"]
    r = np.where(sx, rc(key + ".xp", ex(sc, -3), ex(sc, -2), ex(sc, -1), sc, ex(sc, 1)),
                 rc(key + ".xn", ex(sc, 2), ex(sc, 1), sc, ex(sc, -1), ex(sc, -2)))
    mub = fxb * r
    rb = fxb * mu
    _recon_T(T, key + ".xp", np.where(sx, rb, 0.0), (-3, -2, -1, 0, 1), ex, weno, scb)
    _recon_T(T, key + ".xn", np.where(sx, 0.0, rb), (2, 1, 0, -1, -2), ex, weno, scb)
    sy = T.sel[key + ".y"]
    r = np.where(sy, rc(key + ".yp", ny(sc, -3), ny(sc, -2), ny(sc, -1), sc, ny(sc, 1)),
                 rc(key + ".yn", ny(sc, 2), ny(sc, 1), sc, ny(sc, -1), ny(sc, -2)))
    mvb = fyb * r
    rb = fyb * mv
    _recon_T(T, key + ".yp", np.where(sy, rb, 0.0), (-3, -2, -1, 0, 1), ny, weno, scb)
    _recon_T(T, key + ".yn", np.where(sy, 0.0, rb), (2, 1, 0, -1, -2), ny, weno, scb)
    mwb = np.zeros_like(mw)
    S = lambda lv: sc[lv - 1]  # noqa: E731
    for k in range(2, nk + 1):
        mass = mw[k - 1]
        pos, neg, strict = _zflux_kind(k, nk, weno)
        sel = T.sel[f"{key}.z{k}"]
        vals, coefs = [], []
        for tag, (kind, levs) in (("p", pos), ("n", neg)):
            if kind == "lin":
                cs = (p.c1[k - 1], p.c2[k - 1])
                vals.append(p.c1[k - 1] * S(levs[0]) + p.c2[k - 1] * S(levs[1]))
            elif kind == "upd":
                phi = T.coef[f"{key}.z{k}{tag}"]
                cs = T.upstrpd_coefs(f"{key}.z{k}{tag}")
                vals.append(S(levs[1]) + phi * (S(levs[1]) - S(levs[0])))
            elif kind == "f3":
                cs = FLX3
                vals.append(flx3(*(S(lv) for lv in levs)))
            elif kind == "weno":
                cs = T.weno_coefs(f"{key}.z{k}{tag}")
                vals.append(rc(f"{key}.z{k}{tag}", *(S(lv) for lv in levs)))
            else:
                cs = FLX5
                vals.append(flx5(*(S(lv) for lv in levs)))
            coefs.append((levs, cs))
        fb = fzb[k - 1]
        mwb[k - 1] += fb * np.where(sel, vals[0], vals[1])
        rb = fb * mass
        for (levs, cs), part in zip(coefs, (np.where(sel, rb, 0.0), np.where(sel, 0.0, rb))):
            for lv, cf in zip(levs, cs):
                scb[lv - 1] += cf * part
    return scb, mub, mvb, mwb, divxb


# --------------------------------------------------------------------------- RK stage and large step

def rk_stage(p, T, nrk, X, A, ppx):
    key = f"s{nrk}"
    if T.record:
        T.traj[key] = {"X": {n: X[n].copy() for n in X}, "ppx": ppx.copy()}
    terms = stage_terms(p, X["th"], X["pp"])
    rru, rrv, rrw, divx = mass_fluxes(p, T, X["u"], X["v"], X["w"])
    uten = adv_u(p, T, key + ".mu", X["u"], rru, rrv, rrw, divx)
    vten = adv_v(p, T, key + ".mv", X["v"], rru, rrv, rrw, divx)
    wten = adv_w(p, T, key + ".mw", X["w"], rru, rrv, rrw, divx) + buoyancy(p, terms["dum8"])
    inp = {"u": X["u"], "v": X["v"], "w": X["w"], "pp": X["pp"], "ua": A["u"], "va": A["v"], "wa": A["w"],
           "ppi": A["pp"], "ppx": ppx, "uten": uten, "vten": vten, "wten": wten, "ppten": np.zeros_like(X["pp"]),
           "thv": terms["thv"], "ppterm": terms["ppterm"], "piadv": terms["piadv"]}
    ac = acoustic_stage(p, T, key + ".ac", nrk, inp)
    div, theta_div = getdiv(p, T, ac["rru"], ac["rrv"], ac["rrw"], terms["thterm"])
    mu, mv, mw, divx2 = getdivx(p, T, ac["rru"], ac["rrv"], ac["rrw"])
    adv = scalar_adv(p, T, key + ".th", L0(p, X["th"]), mu, mv, mw, divx2, nrk == 3)
    dttmp = float(F(F(p.dt) / F(4 - nrk)))
    th = T.clamp(A["th"] + dttmp * ((0.0 - theta_div) + adv), float(stg.SMEPS_RK))
    return {"u": ac["u"], "v": ac["v"], "w": ac["w"], "pp": ac["pp"], "th": th}, ac["ppx"]


def L0(p, th):
    """theta packed for advection: (th0 - 300) + theta' (stage.pack_theta)."""
    return p.lev(p.th0 - float(stg.TH0R)) + th


def rk_stage_T(p, T, nrk, A, Xb, ppxb, Ab):
    """Transpose of rk_stage: returns (adjoint of X, adjoint of ppx); accumulates into Ab."""
    key = f"s{nrk}"
    X, ppx = T.traj[key]["X"], T.traj[key]["ppx"]
    Tr = T.record
    T.record = False  # re-evaluate the forward's intermediates with its frozen choices
    try:
        terms = stage_terms(p, X["th"], X["pp"])
        rru, rrv, rrw, divx = mass_fluxes(p, T, X["u"], X["v"], X["w"])
        uten = adv_u(p, T, key + ".mu", X["u"], rru, rrv, rrw, divx)
        vten = adv_v(p, T, key + ".mv", X["v"], rru, rrv, rrw, divx)
        wten = adv_w(p, T, key + ".mw", X["w"], rru, rrv, rrw, divx) + buoyancy(p, terms["dum8"])
        inp = {"u": X["u"], "v": X["v"], "w": X["w"], "pp": X["pp"], "ua": A["u"], "va": A["v"], "wa": A["w"],
               "ppi": A["pp"], "ppx": ppx, "uten": uten, "vten": vten, "wten": wten,
               "ppten": np.zeros_like(X["pp"]), "thv": terms["thv"], "ppterm": terms["ppterm"],
               "piadv": terms["piadv"]}
        s = acoustic_stage(p, T, key + ".ac", nrk, inp)
        div, theta_div = getdiv(p, T, s["rru"], s["rrv"], s["rrw"], terms["thterm"])
        mu, mv, mw, divx2 = getdivx(p, T, s["rru"], s["rrv"], s["rrw"])
    finally:
        T.record = Tr
    dttmp = float(F(F(p.dt) / F(4 - nrk)))
    thb = Xb["th"]
    Ab["th"] += thb
    tdb = -dttmp * thb
    advb = dttmp * thb
    scb, mub, mvb, mwb, divx2b = scalar_adv_T(p, T, key + ".th", L0(p, X["th"]), mu, mv, mw, divx2, nrk == 3, advb)
    rub, rvb, rwb = getdivx_T(p, mub, mvb, mwb, divx2b)
    db = tdb * terms["thterm"]
    thtermb = tdb * div
    r1, r2, r3 = getdiv_T(p, db, s["rrw"].shape)
    bar = {"u": Xb["u"], "v": Xb["v"], "w": Xb["w"], "pp": Xb["pp"], "ppd": np.zeros_like(Xb["pp"]),
           "ppx": ppxb, "rru": rub + r1, "rrv": rvb + r2, "rrw": rwb + r3}
    ab = acoustic_stage_T(p, T, key + ".ac", nrk, inp, bar)
    Ab["u"] += ab["ua"]
    Ab["v"] += ab["va"]
    Ab["w"] += ab["wa"]
    Ab["pp"] += ab["ppi"]
    ub, vb, wb, ppb = ab["u"], ab["v"], ab["w"], ab["pp"]
    t_ub, t_rru1, t_rrv1, t_rrw1, t_dx1 = adv_u_T(p, T, key + ".mu", X["u"], rru, rrv, rrw, divx, ab["uten"])
    t_vb, t_rru2, t_rrv2, t_rrw2, t_dx2 = adv_v_T(p, T, key + ".mv", X["v"], rru, rrv, rrw, divx, ab["vten"])
    t_wb, t_rru3, t_rrv3, t_rrw3, t_dx3 = adv_w_T(p, T, key + ".mw", X["w"], rru, rrv, rrw, divx, ab["wten"])
    d8b = buoyancy_T(p, ab["wten"])
    mu_b, mv_b, mw_b = mass_fluxes_T(p, t_rru1 + t_rru2 + t_rru3, t_rrv1 + t_rrv2 + t_rrv3,
                                     t_rrw1 + t_rrw2 + t_rrw3, t_dx1 + t_dx2 + t_dx3)
    th_b, pp_b = stage_terms_T(p, {"thv": ab["thv"], "dum8": d8b, "thterm": thtermb, "ppterm": ab["ppterm"],
                                   "piadv": ab["piadv"]})
    Xb_prev = {"u": ub + t_ub + mu_b, "v": vb + t_vb + mv_b, "w": wb + t_wb + mw_b, "pp": ppb + pp_b,
               "th": th_b + scb}
    return Xb_prev, ab["ppx"]


def acoustic_nloop(p, nrk):
    return sound3_stage_schedule_f32(p.dt, p.nsound, nrk)[0]


class Step:
    """The dry periodic large step and its adjoint on halo-free fields."""

    def __init__(self, p: Params):
        self.p = p

    def forward(self, S, T: Tape | None = None):
        T = T if T is not None else Tape()
        p = self.p
        A = {"u": S["u"], "v": S["v"], "w": S["w"], "th": S["th"], "pp": S["pp"]}
        X = dict(A)
        ppx = S["ppx"]
        for nrk in (1, 2, 3):
            X, ppx = rk_stage(p, T, nrk, X, A, ppx)
        out = dict(X)
        out["ppx"] = X["pp"] + ppx
        return out, T

    def adjoint(self, T: Tape, Sb):
        """Adjoint of the state before the step from that after it (tape of the step's forward)."""
        p = self.p
        A = T.traj["s1"]["X"]
        Xb = {n: Sb[n].copy() for n in ("u", "v", "w", "th", "pp")}
        Xb["pp"] = Xb["pp"] + Sb["ppx"]
        ppxb = Sb["ppx"].copy()
        Ab = {n: np.zeros_like(Xb[n]) for n in Xb}
        for nrk in (3, 2, 1):
            Xb, ppxb = rk_stage_T(p, T, nrk, A, Xb, ppxb, Ab)
        out = {n: Xb[n] + Ab[n] for n in Xb}
        out["ppx"] = ppxb
        return out


def storm_intensity(p, w, zlo=2000.0, zhi=8000.0):
    """Base-density-weighted mean of w^2 over the w levels with height in [zlo, zhi]; and dJ/dw."""
    z = (np.arange(p.nk + 1)) * float(p.dz)
    sel = (z >= zlo) & (z <= zhi)
    wt = np.where(sel, p.rf0, 0.0)
    norm = float(wt.sum()) * p.ni * p.nj
    J = float(np.sum(wt[:, None, None] * w * w)) / norm
    return J, (2.0 * wt[:, None, None] * w) / norm


def unpad(a, ni, nj, nk, ng=3):
    """Halo-free state from the padded carriers of a large-step state."""
    return {"u": np.asarray(a["u3d"][1 : nk + 1, ng : ng + nj, ng : ng + ni], D),
            "v": np.asarray(a["v3d"][1 : nk + 1, ng : ng + nj, ng : ng + ni], D),
            "w": np.asarray(a["w3d"][1 : nk + 2, ng : ng + nj, ng : ng + ni], D),
            "th": np.asarray(a["th3d"], D),
            "pp": np.asarray(a["pp3d"][1 : nk + 1, 1 : nj + 1, 1 : ni + 1], D),
            "ppx": np.asarray(a["ppx"][1 : nk + 1, 1 : nj + 1, 1 : ni + 1], D)}


__all__ = ["Params", "STATE", "Step", "Tape", "storm_intensity", "unpad"]
