"""Adjoint of the dry large step (flat terrain, periodic sides) in binary64.

``oracle/adjoint_large_step.py`` keeps the state without halos.  Its forward
is pinned here against the padded restatement ``oracle.large_step`` (both in
binary64) on a C5-like state, and each adjoint -- momentum advection, scalar
advection with frozen WENO5 weights, the acoustic stage with its coefficient
sensitivities, one RK stage, one large step -- is checked by the dot-product
identity against central differences of the same forward replaying its tape
(upwind choices, WENO5 weights and limiters frozen at the recorded values).
"""

from __future__ import annotations

import numpy as np
import pytest

from gm1_amd.large_step import host_initial_state
from test_adjoint_sound import c5_like
from test_binary64 import oracle_binary64, promote

FIELDS = ("u", "v", "w", "th", "pp", "ppx")


def setup(seed=3, ni=14, nj=12, nk=10):
    o64 = oracle_binary64()
    cfg = c5_like(ni, nj, nk, seed=seed)
    host = host_initial_state(cfg)
    a = promote({k: v.copy() for k, v in host["arrays"].items()})
    m = promote(host["metrics"])
    base = promote(host["base"])
    rng = np.random.default_rng(seed)
    ng = 3
    # winds and pressure everywhere, periodic-consistent (interior perturbed, halos filled)
    for name, scale in (("u3d", 3.0), ("v3d", 3.0), ("w3d", 0.5), ("th3d", 0.3), ("pp3d", 1e-4), ("ppx", 1e-4)):
        a[name] = a[name].copy()
        if name == "th3d":
            a[name] += scale * rng.standard_normal(a[name].shape)
        elif name in ("pp3d", "ppx"):
            a[name][1 : nk + 1, 1 : nj + 1, 1 : ni + 1] += scale * rng.standard_normal((nk, nj, ni))
        elif name == "w3d":
            a[name][2 : nk + 1, ng : ng + nj, ng : ng + ni] += scale * rng.standard_normal((nk - 1, nj, ni))
        else:
            a[name][1 : nk + 1, ng : ng + nj, ng : ng + ni] += scale * rng.standard_normal((nk, nj, ni))
    # the periodic image faces u(ni) = u(0), v(nj) = v(0), as an acoustic stage leaves them
    a["u3d"][:, :, ng + ni] = a["u3d"][:, :, ng]
    a["v3d"][:, ng + nj, :] = a["v3d"][:, ng, :]
    sides = (1, 1, 1, 1)
    for name, kind in (("u3d", o64.boundary.KIND_U), ("v3d", o64.boundary.KIND_V), ("w3d", o64.boundary.KIND_W)):
        o64.boundary.fill(a[name], kind, sides, ni=ni, nj=nj, nk=nk)
    # the state a large step starts from: A = current, ppi = pp3d
    for nm, src in (("ua", "u3d"), ("va", "v3d"), ("wa", "w3d"), ("tha", "th3d")):
        a[nm] = a[src].copy()
    for name in ("pp3d", "ppx"):
        o64.boundary.bcp(a[name], sides, ni=ni, nj=nj, nk=nk)
    a["ppi"] = a["pp3d"].copy()
    prof = o64.sound.terrain_coordinate_profiles(nk, cfg.dz)
    c = {"ni": ni, "nj": nj, "nk": nk, "dx": cfg.dx, "dy": cfg.dy, "dz": cfg.dz, "dt": cfg.dt,
         "nsound": cfg.nsound, "lateral_sides": sides}
    als = o64.adjoint_large_step
    p = als.Params(m, base, prof, c)
    return o64, als, p, a, m, base, prof, c


def test_halo_free_forward_matches_the_padded_restatement():
    o64, als, p, a, m, base, prof, c = setup()
    S = als.unpad(a, p.ni, p.nj, p.nk)
    out, _ = als.Step(p).forward(S)
    a1 = {k: v.copy() for k, v in a.items()}
    o64.large_step.large_step(a1, m, base, prof, c, zt=np.float64(p.nk * p.dz))
    ref = als.unpad(a1, p.ni, p.nj, p.nk)
    for n in FIELDS:
        err = np.abs(out[n] - ref[n]).max() / max(np.abs(ref[n]).max(), 1e-300)
        print(n, f"{err:.2e}")
        assert err <= 1e-12, (n, err)


def _dot(x, y):
    return sum(float(np.sum(x[n] * y[n])) for n in x)


def _fd_check(fwd, adj, x, ybar, rng, eps=1e-6, tol=1e-8, names=None):
    """<ybar, (F(x+eps d) - F(x-eps d))/(2 eps)> against <F^T ybar, d>."""
    names = names or list(x)
    d = {n: rng.standard_normal(x[n].shape) * (np.abs(x[n]).max() + 1e-30) for n in names}
    xp = {n: x[n] + (eps * d[n] if n in d else 0.0) for n in x}
    xm = {n: x[n] - (eps * d[n] if n in d else 0.0) for n in x}
    yp, ym = fwd(xp), fwd(xm)
    lhs = sum(float(np.sum(ybar[n] * (yp[n] - ym[n]))) for n in ybar) / (2.0 * eps)
    xb = adj(ybar)
    rhs = sum(float(np.sum(xb[n] * d[n])) for n in names)
    scale = sum(float(np.sum(np.abs(ybar[n] * (yp[n] - ym[n])))) for n in ybar) / (2.0 * eps)
    print(f"lhs={lhs:.12e} rhs={rhs:.12e} rel={abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs, scale)


def _state_and_tape(als, p, a):
    S = als.unpad(a, p.ni, p.nj, p.nk)
    out, T = als.Step(p).forward(S)
    return S, out, T


def test_momentum_advection_adjoint():
    o64, als, p, a, *_ = setup()
    S, _, T = _state_and_tape(als, p, a)
    T.replay()
    X = T.traj["s2"]["X"]
    rng = np.random.default_rng(1)

    def fwd(x):
        rru, rrv, rrw, divx = als.mass_fluxes(p, T, x["u"], x["v"], x["w"])
        return {"uten": als.adv_u(p, T, "s2.mu", x["u"], rru, rrv, rrw, divx),
                "vten": als.adv_v(p, T, "s2.mv", x["v"], rru, rrv, rrw, divx),
                "wten": als.adv_w(p, T, "s2.mw", x["w"], rru, rrv, rrw, divx)}

    def adj(yb):
        rru, rrv, rrw, divx = als.mass_fluxes(p, T, X["u"], X["v"], X["w"])
        u1 = als.adv_u_T(p, T, "s2.mu", X["u"], rru, rrv, rrw, divx, yb["uten"])
        v1 = als.adv_v_T(p, T, "s2.mv", X["v"], rru, rrv, rrw, divx, yb["vten"])
        w1 = als.adv_w_T(p, T, "s2.mw", X["w"], rru, rrv, rrw, divx, yb["wten"])
        mu, mv, mw = als.mass_fluxes_T(p, u1[1] + v1[1] + w1[1], u1[2] + v1[2] + w1[2], u1[3] + v1[3] + w1[3],
                                       u1[4] + v1[4] + w1[4])
        return {"u": u1[0] + mu, "v": v1[0] + mv, "w": w1[0] + mw}

    y0 = fwd(X)
    ybar = {n: rng.standard_normal(y0[n].shape) for n in y0}
    _fd_check(fwd, adj, {n: X[n] for n in ("u", "v", "w")}, ybar, rng)


@pytest.mark.parametrize("weno", [False, True])
def test_scalar_advection_adjoint(weno):
    o64, als, p, a, *_ = setup()
    S, _, T = _state_and_tape(als, p, a)
    T.replay()
    nrk = 3 if weno else 2
    key = f"s{nrk}.th"
    rng = np.random.default_rng(2)
    X = T.traj[f"s{nrk}"]["X"]
    sc0 = als.L0(p, X["th"])
    mu0 = rng.standard_normal(X["u"].shape) * 3.0
    mv0 = rng.standard_normal(X["v"].shape) * 3.0
    mw0 = np.zeros_like(X["w"])
    mw0[1 : p.nk] = rng.standard_normal((p.nk - 1, p.nj, p.ni))
    dx0 = rng.standard_normal(X["th"].shape) * 1e-3
    T2 = als.Tape()
    als.scalar_adv(p, T2, key, sc0, mu0, mv0, mw0, dx0, weno)
    T2.replay()

    def fwd(x):
        return {"adv": als.scalar_adv(p, T2, key, x["s"], x["mu"], x["mv"], x["mw"], x["divx"], weno)}

    def adj(yb):
        sb, mub, mvb, mwb, dxb = als.scalar_adv_T(p, T2, key, sc0, mu0, mv0, mw0, dx0, weno, yb["adv"])
        return {"s": sb, "mu": mub, "mv": mvb, "mw": mwb, "divx": dxb}

    x = {"s": sc0, "mu": mu0, "mv": mv0, "mw": mw0, "divx": dx0}
    ybar = {"adv": rng.standard_normal(sc0.shape)}
    # WENO5's weights are differentiated: the map is smooth but strongly curved, so the step is small
    _fd_check(fwd, adj, x, ybar, rng, eps=1e-8 if weno else 1e-6, tol=1e-7 if weno else 1e-8,
              names=["s", "mu", "mv", "divx"])


@pytest.mark.parametrize("nrk", [1, 2, 3])
def test_acoustic_stage_adjoint_with_coefficients(nrk):
    o64, als, p, a, *_ = setup()
    S, _, T = _state_and_tape(als, p, a)
    rng = np.random.default_rng(10 + nrk)
    X = T.traj[f"s{nrk}"]["X"]
    sh = X["u"].shape
    inp = {"u": X["u"], "v": X["v"], "w": X["w"], "pp": X["pp"], "ua": S["u"], "va": S["v"], "wa": S["w"],
           "ppi": S["pp"], "ppx": S["ppx"], "uten": 1e-3 * rng.standard_normal(sh),
           "vten": 1e-3 * rng.standard_normal(sh), "wten": 1e-3 * rng.standard_normal(X["w"].shape),
           "ppten": 1e-7 * rng.standard_normal(sh)}
    terms = als.stage_terms(p, X["th"], X["pp"])
    for n in ("thv", "ppterm", "piadv"):
        inp[n] = terms[n]
    T2 = als.Tape()
    als.acoustic_stage(p, T2, "ac", nrk, inp)
    T2.replay()
    outs = ("u", "v", "w", "pp", "ppd", "ppx", "rru", "rrv", "rrw")

    def fwd(x):
        s = als.acoustic_stage(p, T2, "ac", nrk, x)
        return {n: s[n] for n in outs}

    def adj(yb):
        return als.acoustic_stage_T(p, T2, "ac", nrk, inp, yb)

    y0 = fwd(inp)
    ybar = {n: rng.standard_normal(y0[n].shape) for n in outs}
    names = ["u", "v", "w", "pp", "ua", "va", "wa", "ppi", "ppx", "uten", "vten", "wten", "ppten", "thv",
             "ppterm", "piadv"]
    _fd_check(fwd, adj, inp, ybar, rng, names=names)


def test_rk_stage_adjoint():
    o64, als, p, a, *_ = setup()
    S, _, T = _state_and_tape(als, p, a)
    T.replay()
    rng = np.random.default_rng(5)
    nrk = 2
    A = T.traj["s1"]["X"]
    X = T.traj[f"s{nrk}"]["X"]
    ppx = T.traj[f"s{nrk}"]["ppx"]

    def fwd(x):
        Ax = {n: x["A" + n] for n in ("u", "v", "w", "th", "pp")}
        Xx = {n: x["X" + n] for n in ("u", "v", "w", "th", "pp")}
        Xn, ppxn = als.rk_stage(p, T, nrk, Xx, Ax, x["ppx"])
        return {**{"X" + n: Xn[n] for n in Xn}, "ppx": ppxn}

    def adj(yb):
        Ab = {n: np.zeros_like(A[n]) for n in A}
        Xb, ppxb = als.rk_stage_T(p, T, nrk, A, {n: yb["X" + n] for n in ("u", "v", "w", "th", "pp")}, yb["ppx"],
                                  Ab)
        return {**{"X" + n: Xb[n] for n in Xb}, **{"A" + n: Ab[n] for n in Ab}, "ppx": ppxb}

    x = {**{"X" + n: X[n] for n in X}, **{"A" + n: A[n] for n in A}, "ppx": ppx}
    y0 = fwd(x)
    ybar = {n: rng.standard_normal(y0[n].shape) for n in y0}
    _fd_check(fwd, adj, x, ybar, rng, eps=1e-8, tol=1e-6)


def test_large_step_adjoint_and_storm_intensity_gradient():
    o64, als, p, a, *_ = setup()
    S, out, T = _state_and_tape(als, p, a)
    T.replay()
    step = als.Step(p)
    rng = np.random.default_rng(6)

    def fwd(x):
        return step.forward(x, T)[0]

    ybar = {n: rng.standard_normal(out[n].shape) for n in FIELDS}
    _fd_check(fwd, lambda yb: step.adjoint(T, yb), S, ybar, rng, eps=1e-8, tol=1e-6)
    # the loss and its gradient through the adjoint: dJ/d(th, u, v) against central differences
    J, wb = als.storm_intensity(p, out["w"])
    Sb = {n: np.zeros_like(out[n]) for n in FIELDS}
    Sb["w"] = wb
    g = step.adjoint(T, Sb)
    for name in ("th", "u", "v"):
        d = rng.standard_normal(S[name].shape)
        eps = 1e-6 * np.abs(S[name]).max()
        Jp = als.storm_intensity(p, fwd({**S, name: S[name] + eps * d})["w"])[0]
        Jm = als.storm_intensity(p, fwd({**S, name: S[name] - eps * d})["w"])[0]
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.sum(g[name] * d))
        print(name, fd, ad)
        assert abs(fd - ad) <= 1e-6 * abs(fd), (name, fd, ad)
