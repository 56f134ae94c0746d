"""The flat, uniform, doubly periodic psolver=3 stage (test infrastructure).

A vectorised binary32 restatement of ``cpu_sound3_periodic_stage_inplace``
(``cuda_sound3_periodic.py:672-1064``) and its fills
(``fill_periodic_scalar_inplace`` :506-516, ``fill_periodic_uv_inplace``
:519-532).  Every statement keeps the reference's operand order; columns are
processed together (NumPy over (j, i)) and only k runs serially, which changes
no word because no statement reads another column's value written in the same
pass.  Layout (``array_shapes`` :449-491): the scalar family
``(nk+2, nj+2, ni+2)`` with a one-cell periodic halo, u ``(nk+2, nj+2, ni+1)``,
v ``(nk+2, nj+1, ni+2)``, w ``(nk+2, nj, ni)`` without lateral halo.

Pinned by golden vectors of the reference transcription (two fixtures, every
output the reference compares), by the reference's KATs (cadence, fill counts,
Thomas residual and pivots, duplicate faces, divergence sum), and by the
horizontally uniform limit equalling the column oracle of the terrain stage.
"""

from __future__ import annotations

import numpy as np

from .sound import sound3_stage_schedule_f32

F = np.float32

CP = F(1005.7)
ALPH = F(0.60)
KDIV = F(0.10)
SMEPS = F(1.0e-30)

SCALAR_NAMES = ("mh", "rr0", "rf0", "c1", "c2", "mf", "piadv", "ppterm", "thv", "ppten", "ppi", "pp3d", "ppx",
                "ppd", "qk", "pk1", "pk2", "mm", "fpk", "divh", "hadvp", "elim_c", "elim_d", "tri_aa", "tri_bb",
                "tri_cc", "tri_dd", "tri_pivot", "tri_residual")
W_NAMES = ("wten", "wa", "w3d", "rrw")
U_NAMES = ("ua", "u3d", "uten", "rru")
V_NAMES = ("va", "v3d", "vten", "rrv")
METRIC_NAMES = ("dtu0", "dtv0", "uf", "vf", "uh", "vh")
ALL_ARRAY_NAMES = SCALAR_NAMES + W_NAMES + U_NAMES + V_NAMES + METRIC_NAMES
ZEROED = ("qk", "pk1", "pk2", "mm", "fpk", "divh", "hadvp", "elim_c", "elim_d", "tri_aa", "tri_bb", "tri_cc",
          "tri_dd", "tri_pivot", "tri_residual")
COMPARE_OUTPUT_NAMES = ("u3d", "v3d", "w3d", "pp3d", "ppx", "ppd", "rru", "rrv", "rrw", "qk", "pk1", "pk2", "mm",
                        "fpk", "divh", "hadvp", "elim_c", "elim_d", "tri_aa", "tri_bb", "tri_cc", "tri_dd",
                        "tri_pivot", "tri_residual")


def array_shapes(ni: int, nj: int, nk: int) -> dict[str, tuple[int, ...]]:
    """cuda_sound3_periodic.py:449-491."""

    shapes = {n: (nk + 2, nj + 2, ni + 2) for n in SCALAR_NAMES}
    shapes.update({n: (nk + 2, nj, ni) for n in W_NAMES})
    shapes.update({n: (nk + 2, nj + 2, ni + 1) for n in U_NAMES})
    shapes.update({n: (nk + 2, nj + 1, ni + 2) for n in V_NAMES})
    shapes.update({"dtu0": (nj + 2, ni + 1), "dtv0": (nj + 1, ni + 2), "uf": (ni + 2,), "vf": (nj + 2,),
                   "uh": (ni + 2,), "vh": (nj + 2,)})
    return shapes


def default_periodic_face_weights(ni: int, nj: int) -> tuple[np.ndarray, np.ndarray]:
    """cuda_sound3_periodic.py:494-503: every live face weighs one."""

    dtu0 = np.zeros((nj + 2, ni + 1), np.float32)
    dtv0 = np.zeros((nj + 1, ni + 2), np.float32)
    dtu0[1 : nj + 1, :] = F(1.0)
    dtv0[:, 1 : ni + 1] = F(1.0)
    return dtu0, dtv0


def fill_periodic_scalar(value: np.ndarray, *, ni: int, nj: int, nk: int) -> None:
    """cuda_sound3_periodic.py:506-516: x halos, then y halos (corners from the x-filled rows)."""

    s = value[1 : nk + 1]
    s[:, :, 0] = s[:, :, ni]
    s[:, :, ni + 1] = s[:, :, 1]
    s[:, 0, :] = s[:, nj, :]
    s[:, nj + 1, :] = s[:, 1, :]


def fill_periodic_uv(u: np.ndarray, v: np.ndarray, *, ni: int, nj: int, nk: int) -> None:
    """cuda_sound3_periodic.py:519-532."""

    u[1 : nk + 1, 0, :] = u[1 : nk + 1, nj, :]
    u[1 : nk + 1, nj + 1, :] = u[1 : nk + 1, 1, :]
    v[1 : nk + 1, :, 0] = v[1 : nk + 1, :, ni]
    v[1 : nk + 1, :, ni + 1] = v[1 : nk + 1, :, 1]


def stage(a: dict, *, ni: int, nj: int, nk: int, dt: float, nsound: int, nrk: int, rdx: float, rdy: float,
          rdz: float, cp: float = CP, alph: float = ALPH, kdiv: float = KDIV, smeps: float = SMEPS) -> dict:
    """One source-ordered RK/acoustic stage in place (cuda_sound3_periodic.py:672-1064)."""

    nloop, dts = sound3_stage_schedule_f32(dt, nsound, nrk)
    dts = F(dts)
    rdx, rdy, rdz, cp, alph, kdiv, smeps = (F(x) for x in (rdx, rdy, rdz, cp, alph, kdiv, smeps))
    beta = F(F(1.0) - alph)
    asq = F(alph * alph)
    for name in ZEROED:
        a[name].fill(F(0.0))
    a["ppd"][1 : nk + 1] = a["ppx"][1 : nk + 1]
    a["rru"].fill(F(0.0))
    a["rrv"].fill(F(0.0))
    if nrk != 1:
        np.copyto(a["u3d"], a["ua"])
        np.copyto(a["v3d"], a["va"])
        a["pp3d"][1 : nk + 1] = a["ppi"][1 : nk + 1]

    # :728-784 stage coefficients on the physical columns.
    J, I = slice(1, nj + 1), slice(1, ni + 1)
    c = lambda name, k: a[name][k, J, I]  # noqa: E731
    dtsrdz = F(dts * rdz)
    for k in range(1, nk + 1):
        tem = dtsrdz * c("mh", k)
        r2 = dtsrdz * c("mh", k)
        r2 = r2 * c("rr0", k)
        r2 = r2 * c("rf0", k + 1)
        r1 = dtsrdz * c("mh", k)
        r1 = r1 * c("rr0", k)
        r1 = r1 * c("rf0", k)
        tem1 = F(dtsrdz * cp) * c("mf", k)
        a["qk"][k, J, I] = tem * c("ppterm", k)
        a["pk2"][k, J, I] = r2 * ((-c("c2", k + 1)) * c("piadv", k + 1) + (F(1.0) - c("c1", k + 1)) * c("piadv", k))
        a["pk1"][k, J, I] = r1 * (c("c1", k) * c("piadv", k - 1) + (c("c2", k) - F(1.0)) * c("piadv", k))
        a["mm"][k, J, I] = tem1 * (c("c2", k) * c("thv", k) + c("c1", k) * c("thv", k - 1))
        a["rrw"][k] = a["wa"][k] * beta
        if nrk != 1:
            a["w3d"][k] = a["wa"][k]
    a["pk1"][1, J, I] = F(0.0)
    a["pk2"][nk, J, I] = F(0.0)

    tem1 = F(F(rdx * cp) * F(0.5))
    tem2 = F(F(rdy * cp) * F(0.5))
    temx = F(F(dts * F(0.5)) * rdx)
    temy = F(F(dts * F(0.5)) * rdy)
    tavg = F(F(1.0) / F(nloop))
    u, v, w, ppd, thv = a["u3d"], a["v3d"], a["w3d"], a["ppd"], a["thv"]
    K = slice(1, nk + 1)
    for n in range(1, nloop + 1):
        # :787-828 horizontal pressure gradient on every live face.
        pg = tem1 * (ppd[K, 1 : nj + 2, 1 : ni + 2] - ppd[K, 1 : nj + 2, 0 : ni + 1])
        pg = pg * a["uf"][1 : ni + 2][None, None, :]
        pg = pg * (thv[K, 1 : nj + 2, 1 : ni + 2] + thv[K, 1 : nj + 2, 0 : ni + 1])
        uu = u[K, 1 : nj + 2, :]
        u[K, 1 : nj + 2, :] = uu + (dts * a["dtu0"][1 : nj + 2, :])[None] * (a["uten"][K, 1 : nj + 2, :] - pg)
        pg = tem2 * (ppd[K, 1 : nj + 2, 1 : ni + 2] - ppd[K, 0 : nj + 1, 1 : ni + 2])
        pg = pg * a["vf"][1 : nj + 2][None, :, None]
        pg = pg * (thv[K, 1 : nj + 2, 1 : ni + 2] + thv[K, 0 : nj + 1, 1 : ni + 2])
        vv = v[K, :, 1 : ni + 2]
        v[K, :, 1 : ni + 2] = vv + (dts * a["dtv0"][:, 1 : ni + 2])[None] * (a["vten"][K, :, 1 : ni + 2] - pg)

        _columns(a, ni=ni, nj=nj, nk=nk, nrk=nrk, n=n, nloop=nloop, dts=dts, rdx=rdx, rdy=rdy, temx=temx,
                 temy=temy, alph=alph, beta=beta, asq=asq, kdiv=kdiv, smeps=smeps)
        if n < nloop:
            fill_periodic_scalar(ppd, ni=ni, nj=nj, nk=nk)
        # :1013-1032 time averages of the face velocities.
        ru = a["rru"][K, 1 : nj + 2, :] + u[K, 1 : nj + 2, :]
        a["rru"][K, 1 : nj + 2, :] = ru if n < nloop else ru * tavg
        rv = a["rrv"][K, :, 1 : ni + 2] + v[K, :, 1 : ni + 2]
        a["rrv"][K, :, 1 : ni + 2] = rv if n < nloop else rv * tavg
    fill_periodic_uv(u, v, ni=ni, nj=nj, nk=nk)
    if nrk < 3:
        fill_periodic_scalar(a["pp3d"], ni=ni, nj=nj, nk=nk)
    return {"nrk": nrk, "nloop": nloop, "dts_f32": float(dts), "ppd_boundary_fills": nloop - 1,
            "stage_velocity_boundary_fills": 1, "stage_pressure_boundary_fills": 1 if nrk < 3 else 0,
            "ppx_written": nrk == 3}


def _columns(a, *, ni, nj, nk, nrk, n, nloop, dts, rdx, rdy, temx, temy, alph, beta, asq, kdiv, smeps):
    """:830-1010 divergence, pressure advection, fpk, Thomas solve, residual, pressure update."""

    J, I = slice(1, nj + 1), slice(1, ni + 1)
    u, v, w, piadv = a["u3d"], a["v3d"], a["w3d"], a["piadv"]
    uh = a["uh"][1 : ni + 1][None, :]
    vh = a["vh"][1 : nj + 1][:, None]
    P = lambda name, k: a[name][k, J, I]  # noqa: E731
    for k in range(1, nk + 1):
        du = ((u[k, J, 1 : ni + 1] - u[k, J, 0:ni]) * rdx) * uh
        dv = ((v[k, 1 : nj + 1, I] - v[k, 0:nj, I]) * rdy) * vh
        div = du + dv
        div = np.where(np.abs(div) < smeps, F(0.0), div).astype(np.float32)
        a["divh"][k, J, I] = div
        pc = piadv[k, J, I]
        gxw = pc - piadv[k, J, 0:ni]
        gxe = piadv[k, J, 2 : ni + 2] - pc
        gys = pc - piadv[k, 0:nj, I]
        gyn = piadv[k, 2 : nj + 2, I] - pc
        hx = u[k, J, 0:ni] * gxw + u[k, J, 1 : ni + 1] * gxe
        hx = ((-hx) * temx) * uh
        hy = v[k, 0:nj, I] * gys + v[k, 1 : nj + 1, I] * gyn
        hy = ((-hy) * temy) * vh
        hp = hx + hy
        a["hadvp"][k, J, I] = hp
        src = dts * (P("ppten", k) - P("ppterm", k) * div)
        vert = w[k] * (P("pk1", k) + P("qk", k)) + w[k + 1] * (P("pk2", k) - P("qk", k))
        a["fpk"][k, J, I] = (src + hp) + beta * vert
    for k in range(2, nk + 1):
        zero = np.zeros((nj, ni), np.float32)
        aa = (-asq) * (P("mm", k) * (P("pk1", k - 1) + P("qk", k - 1))) if k >= 3 else zero
        cc = (-asq) * ((-P("mm", k)) * (P("pk2", k) - P("qk", k))) if k < nk else zero
        diag = (-(P("pk1", k) + P("qk", k))) + (P("pk2", k - 1) - P("qk", k - 1))
        bb = F(1.0) - asq * (P("mm", k) * diag)
        pd = P("pp3d", k) - P("pp3d", k - 1)
        fwk = dts * a["wten"][k] + beta * ((-P("mm", k)) * pd)
        upd = (P("pp3d", k) + P("fpk", k)) - (P("pp3d", k - 1) + P("fpk", k - 1))
        dd = (w[k] + fwk) + alph * ((-P("mm", k)) * upd)
        pivot = bb if k < 3 else aa * P("elim_c", k - 1) + bb
        inv = F(1.0) / pivot
        for name, value in (("tri_aa", aa), ("tri_bb", bb), ("tri_cc", cc), ("tri_dd", dd), ("tri_pivot", pivot)):
            a[name][k, J, I] = value
        if k < nk:
            a["elim_c"][k, J, I] = (-cc) * inv
            num = dd if k < 3 else dd - aa * P("elim_d", k - 1)
            a["elim_d"][k, J, I] = num * inv
        else:
            w[k] = (dd - aa * P("elim_d", k - 1)) * inv
    for k in range(nk - 1, 1, -1):
        w[k] = P("elim_c", k) * w[k + 1] + P("elim_d", k)
    for k in range(2, nk + 1):
        lhs = P("tri_aa", k) * w[k - 1]
        lhs = lhs + P("tri_bb", k) * w[k]
        lhs = lhs + P("tri_cc", k) * w[k + 1]
        a["tri_residual"][k, J, I] = lhs - P("tri_dd", k)
    tavg = F(F(1.0) / F(nloop))
    for k in range(1, nk + 1):
        old = P("pp3d", k).copy()
        vert = w[k] * (P("pk1", k) + P("qk", k)) + w[k + 1] * (P("pk2", k) - P("qk", k))
        upd = (old + P("fpk", k)) + alph * vert
        upd = np.where(np.abs(upd) < smeps, F(0.0), upd).astype(np.float32)
        a["pp3d"][k, J, I] = upd
        a["elim_c"][k, J, I] = kdiv * (upd - old)
        a["ppd"][k, J, I] = upd + P("elim_c", k)
        if nrk == 3 and n == nloop:
            a["ppx"][k, J, I] = P("elim_c", k)
        if n < nloop:
            a["rrw"][k] = a["rrw"][k] + w[k]
        else:
            a["rrw"][k] = (a["rrw"][k] + w[k] * alph) * tavg


def three_stage(a: dict, **kwargs) -> list[dict]:
    """cpu_sound3_periodic_three_stage_inplace (cuda_sound3_periodic.py:1071-1079)."""

    return [stage(a, nrk=nrk, **kwargs) for nrk in (1, 2, 3)]
