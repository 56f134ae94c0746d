"""Kessler warm rain on the CM1 large step (test infrastructure).

Restates in binary32 NumPy, vectorised over cells and columns but in the
reference's statement order:

* ``kessler_driver``   -> ``cpu_gm1_kessler_driver_inplace``
                          (cuda_gm1_kessler_driver.py:431-559): pdefq of qv,
                          qc, qr (:300-334), fall speed, the surface-energy
                          budget, fallout in ``nfall`` substeps, the local
                          conversion (``cpu_gm1_kessler_local_inplace``,
                          cuda_gm1_kessler_local.py:251-426) and satadj
                          (:337-428), ptype 1, efall 0 or 1 (``efall``:
                          ``cpu_gm1_kessler_efall_inplace``,
                          cuda_gm1_kessler_efall.py:201-256), eqtset 1 or 2;
* ``large_step_transition`` -> ``cpu_gm1_kessler_large_step_inplace``
                          (cuda_gm1_kessler_large_step.py:183-292): eqtset=2
                          held remove/snapshot, pre-physics EOS, the dry-mass
                          hook, rr <- rho, the driver, held recapture;
* ``apmasscon``        -> ``cpu_apmasscon_inplace`` (cuda_apmasscon.py:256-327).

Binary64 budgets (asq, qbudget) and the dry-mass ledger accumulate in the
reference's loop order.  Transcendentals are NumPy's float32 ``exp``,
``power`` and ``sqrt``, as in the reference; their last-bit agreement with a
GPU's is not guaranteed (parity for those words is checked to a tolerance).
Pinned by golden vectors from the reference's own CPU functions and by
closed forms (total water changes only by surface rain; pdefq conserves the
column mass; post-satadj saturation).
"""

from __future__ import annotations

import numpy as np

F = np.float32
f64 = np.float64

RD, CP, CPV, RV = F(287.04), F(1005.7), F(1870.0), F(461.5)
XLV, CPL, CPI = F(2501000.0), F(4190.0), F(2106.0)
CV = CP - RD
CVV = CPV - RV
ROVCP = RD / CP
EPS = RD / RV
REPS = RV / RD
LV2 = CPL - CPV
LV1 = XLV + LV2 * F(273.15)
RP00 = F(1.0) / F(1.0e5)
P00 = F(1.0e5)
CPDCV = CP / CV
RDDCV = RD / CV
CPDRD = CP / RD
ZERO = F(0.0)


# Transcendental mode.  "numpy": NumPy's float32 exp/power, exactly what the
# reference's CPU transcription calls (pins the golden vectors).  "rounded":
# binary64 exp/pow rounded once to binary32, the GPU kernels' formulation; the
# two differ by at most one unit in the last place per call.
_MODE = {"value": "numpy"}


class transcendentals:
    """Context manager selecting the transcendental mode ("numpy" or "rounded")."""

    def __init__(self, mode: str):
        if mode not in ("numpy", "rounded"):
            raise ValueError(mode)
        self.mode = mode

    def __enter__(self):
        self.prev = _MODE["value"]
        _MODE["value"] = self.mode
        return self

    def __exit__(self, *exc):
        _MODE["value"] = self.prev
        return False


def _exp(x):
    x = np.asarray(x, F)
    if _MODE["value"] == "rounded":
        return np.exp(x.astype(f64)).astype(F)
    return np.exp(x).astype(F)


def _pow(b, e):
    b = np.asarray(b, F)
    if _MODE["value"] == "rounded":
        return np.power(b.astype(f64), f64(F(e))).astype(F)
    return np.power(b, F(e)).astype(F)


def saturation_qv(tnew, prs_cap, pdenom):
    """qvs = eps*esl/(p - esl) with esl = min(611.2 exp(17.67 (T-273.15)/(T-29.65)), 0.5 prs)."""
    esl = F(611.2) * _exp((F(17.67) * (tnew - F(273.15))) / (tnew - F(29.65)))
    esl = np.minimum(esl, prs_cap * F(0.5))
    return (EPS * esl) / (pdenom - esl)


def pdefq(q, rho, ruh, rvh, rmh, *, rmax, volume, asq, index):
    """Zero values below rmax, rescale to conserve column rho*q; asq[index] gains each row's budget."""
    nk, nj, ni = q.shape
    t1 = np.zeros((nj, ni), f64)
    t2 = np.zeros((nj, ni), f64)
    a1 = np.zeros((nj, ni), f64)
    a2 = np.zeros((nj, ni), f64)
    cellw = lambda k, qq: (((rho[k] * qq) * ruh[None, :]) * rvh[:, None]) * rmh[k]  # noqa: E731
    for k in range(nk):
        t1 += (rho[k] * q[k]).astype(f64)
        a1 += cellw(k, q[k]).astype(f64)
        q[k] = np.where(q[k] < F(rmax), ZERO, q[k])
        t2 += (rho[k] * q[k]).astype(f64)
    t3 = (t1 + 1.0e-20) / (t2 + 1.0e-20)
    t3 = np.where(t3 < 0.0, 1.0, t3)
    for k in range(nk):
        q[k] = (t3 * q[k].astype(f64)).astype(F)
        a2 += cellw(k, q[k]).astype(f64)
    for j in range(nj):
        budj = 0.0
        for i in range(ni):
            budj += float(a2[j, i]) - float(a1[j, i])
        asq[index] += budj * volume


def fall_speed(rho, qr):
    with np.errstate(all="ignore"):
        v = (F(14.34) * _pow(rho * qr, 0.1346)) * np.sqrt(F(1.15) / rho).astype(F)
    return np.where(qr > F(1.0e-12), v, ZERO).astype(F)


def fallout(a, *, dt, dz):
    """Flux-form sedimentation of qr per column (:510-543); qbudget[5] gains each row's budget."""
    nk, nj, ni = a["qr"].shape
    dt32 = F(dt)
    rdz = F(1.0) / F(dz)
    vt = a["fall_velocity"]
    crmax = np.zeros((nj, ni), np.float32)
    for k in range(nk):
        crmax = np.maximum(crmax, ((vt[k] * dt32) * rdz) * a["mh"][k])
    nfall = np.maximum(1, (crmax + F(1.0)).astype(np.int64))
    for k in range(nk - 2, -1, -1):
        vt[k] = np.where(a["qr"][k] < F(1.0e-8), vt[k + 1], vt[k])
    dtfall = (dt32 / nfall.astype(F)).astype(F)
    bud_terms = []
    for sub in range(int(nfall.max())):
        act = nfall > sub
        rq = np.zeros((nk + 1, nj, ni), np.float32)
        for k in range(nk):
            rq[k] = (a["rho"][k] * vt[k]) * np.maximum(ZERO, a["qr"][k])
        for k in range(nk):
            delta = (((dtfall * (rq[k + 1] - rq[k])) * rdz) * a["mh"][k]) / a["rho"][k]
            a["qr"][k] = np.where(act, a["qr"][k] + delta, a["qr"][k])
        a["rain"][0] = np.where(act, a["rain"][0] + (F(0.1) * dtfall) * rq[0], a["rain"][0])
        a["prate"][...] = np.where(act, rq[0], a["prate"])
        term = dtfall * a["rr"][0]
        term = term * vt[0]
        term = term * np.maximum(ZERO, a["qr"][0])
        term = term * a["ruh"][None, :]
        term = term * a["rvh"][:, None]
        bud_terms.append((act, term.astype(f64)))
    area = float(F(F(a["dx"]) * F(a["dy"])))
    for j in range(nj):
        bud = 0.0
        for i in range(ni):
            for act, term in bud_terms:
                if act[j, i]:
                    bud += float(term[j, i])
        a["qbudget"][5] += bud * area
    return int(nfall.max())


def local_conversion(a, *, dt, equation_set, temperature_override=None):
    """Autoconversion, accretion, rain evaporation and the theta/pi'/rho update (local.py:251-426)."""
    dt32 = F(dt)
    pi0, th0 = a["pi0"], a["th0"]
    pisum = pi0 + a["pp3d"]
    tmp = temperature_override if temperature_override is not None else (th0 + a["th3d"]) * pisum
    qv, qc, qr = a["qv"].copy(), a["qc"].copy(), a["qr"].copy()
    with np.errstate(all="ignore"):
        ar = np.where(qc > ZERO, np.maximum(F(0.001) * (qc - F(0.001)), ZERO) * dt32, ZERO)
        cr = np.where((qc > ZERO) & (qr > ZERO), ((F(2.2) * qc) * _pow(qr, 0.875)) * dt32, ZERO)
        prs, rho = a["prs"], a["rho"]
        qvs = saturation_qv(tmp, prs, prs)
        rhoqr = rho * qr
        vent = F(1.6) + F(30.3922) * _pow(rhoqr, 0.2046)
        deficit = F(1.0) - qv / qvs
        num = (vent * deficit) * _pow(rhoqr, 0.525)
        den = (F(2.03e4) + F(9.584e6) / (qvs * prs)) * rho
        er_c = np.minimum((num / den) * dt32, qr)
        er_c = np.where(qv + er_c > qvs, qvs - qv, er_c)
        er = np.where((qr > ZERO) & (qv < qvs), er_c, ZERO)
        over = (ar + cr) > qc
        term1 = ar + cr
        ar2 = np.where(over, (qc * ar) / term1, ar)
        cr2 = np.where(over, (qc * cr) / term1, cr)
    branches = {"qc_positive": qc > ZERO, "autoconverting": (qc > ZERO) & ((qc - F(0.001)) > ZERO),
                "accreting": (qc > ZERO) & (qr > ZERO), "evaporating": (qr > ZERO) & (qv < qvs),
                "evaporation_capped": (qr > ZERO) & (qv < qvs) & (qv + np.minimum((num / den) * dt32, qr) > qvs),
                "limited": over, "updated": er > F(1.0e-7)}
    ar, cr = ar2, cr2
    qvn = qv + er
    qcn = qc - (ar + cr)
    qrn = qr + ((ar + cr) - er)
    upd = er > F(1.0e-7)
    if equation_set == 2:
        cpml = (CP + CPV * qvn) + CPL * (qcn + qrn)
        cvml = (CV + CVV * qvn) + CPL * (qcn + qrn)
        rm = RD + RV * qvn
        latent = LV1 - LV2 * tmp
        first = latent / ((CPDCV * cvml) * pisum)
        second = ((th0 + a["th3d"]) * (RV / cvml)) * (F(1.0) - (ROVCP * cpml) / rm)
        th_new = a["th3d"] - er * (first - second)
        base = ((rho * rm) * (th0 + th_new)) * RP00
        pp_new = _pow(base, RDDCV) - pi0
        pis_new = pi0 + pp_new
        prs_new = P00 * _pow(pis_new, CPDRD)
        a["th3d"][...] = np.where(upd, th_new, a["th3d"])
        a["pp3d"][...] = np.where(upd, pp_new, a["pp3d"])
        a["prs"][...] = np.where(upd, prs_new, a["prs"])
    else:
        latent = LV1 - LV2 * tmp
        th_new = a["th3d"] - er * (latent / (CP * pisum))
        den2 = ((RD * (th0 + th_new)) * pisum) * (F(1.0) + qvn * REPS)
        rho_new = a["prs"] / den2
        a["th3d"][...] = np.where(upd, th_new, a["th3d"])
        a["rho"][...] = np.where(upd, rho_new, a["rho"])
    a["qv"][...], a["qc"][...], a["qr"][...] = qvn, qcn, qrn
    return {"tmp": tmp, "autoconversion": ar, "accretion": cr, "evaporation": er, "branches": branches}


def _budget_by_level(values, rr, ruh, rvh, rmh):
    """sum over (j, i) in loop order of rr*v*(ruh*rvh*rmh), per level, binary64."""
    nk, nj, ni = values.shape
    out = np.zeros(nk, f64)
    for k in range(nk):
        tem = (ruh[None, :] * rvh[:, None]) * rmh[k]
        terms = ((rr[k] * values[k]) * tem).astype(f64).ravel()
        acc = 0.0
        for t in terms:
            acc += float(t)
        out[k] = acc
    return out


def satadj(a, *, dx, dy, dz, equation_set):
    """Saturation adjustment (:337-428); returns iteration counts."""
    qv0, qc0 = a["qv"].copy(), a["qc"].copy()
    qvnew, qcnew = qv0.copy(), qc0.copy()
    pi0, th0, prs, rho = a["pi0"], a["th0"], a["prs"].copy(), a["rho"]
    tnew = (th0 + a["th3d"]) * (pi0 + a["pp3d"])
    with np.errstate(all="ignore"):
        qvs = saturation_qv(tnew, prs, prs)
    iterations = np.zeros(qv0.shape, np.uint32)
    if equation_set == 2:
        act = (qc0 > F(1.0e-12)) | (qv0 > qvs)
        ql = np.maximum(qc0 + a["qr"], ZERO)
        cvml = (CV + CVV * qvnew) + CPL * ql
        lhv = LV1 - LV2 * tnew
        t1 = (lhv - RV * tnew) / cvml
        d1 = (t1 * F(17.67)) * F(243.5)
        tlast = tnew.copy()
        pnew = prs.copy()
        running = act.copy()
        n = 0
        converge = F(2.0) * F(0.0001)
        while running.any():
            n += 1
            with np.errstate(all="ignore"):
                sq = (tnew - F(29.65)) * (tnew - F(29.65))
                denom = F(1.0) + (d1 * qvs) / sq
                dqv = (qvs - qvnew) / denom
                dqv = np.minimum(dqv, qcnew)
                dqv = np.where(qvnew + dqv < F(1.0e-20), F(1.0e-20) - qvnew, dqv)
                qv_n = qvnew + dqv
                qc_n = qcnew - dqv
                t_n = tnew - dqv * t1
                p_n = (rho * (RD + RV * qv_n)) * t_n
                conv = np.abs((t_n - tlast).astype(f64)) <= float(converge)
                qvs_n = saturation_qv(t_n, prs, p_n)
            if n == 100:
                raise RuntimeError("Kessler satadj failed to converge in 100 iterations")
            qvnew = np.where(running, qv_n, qvnew)
            qcnew = np.where(running, qc_n, qcnew)
            tnew = np.where(running, t_n, tnew)
            pnew = np.where(running, p_n, pnew)
            cont = running & ~conv
            tlast = np.where(cont, t_n, tlast)
            qvs = np.where(cont, qvs_n, qvs)
            iterations = np.where(running, np.uint32(n), iterations)
            running = cont
        a["prs"][...] = np.where(act, pnew, a["prs"])
        pp_new = _pow(pnew * RP00, ROVCP) - pi0
        a["pp3d"][...] = np.where(act, pp_new, a["pp3d"])
        a["th3d"][...] = np.where(act, tnew / (pi0 + a["pp3d"]) - th0, a["th3d"])
        a["qv"][...] = np.where(act, qvnew, a["qv"])
        a["qc"][...] = np.where(act, qcnew, a["qc"])
    else:
        lhv = LV1 - LV2 * tnew
        sq = (tnew - F(29.65)) * (tnew - F(29.65))
        num = ((lhv * qvs) * F(17.67)) * F(243.5)
        dqv = (qvs - qvnew) / (F(1.0) + num / (CP * sq))
        dqv = np.where(qvnew + dqv < F(1.0e-20), F(1.0e-20) - qvnew, dqv)
        dqv = np.minimum(dqv, np.maximum(ZERO, qcnew))
        qvnew = qvnew + dqv
        qcnew = qcnew - dqv
        iterations[...] = 1
        a["th3d"][...] = a["th3d"] - dqv * (lhv / (CP * (pi0 + a["pp3d"])))
        a["qv"][...], a["qc"][...] = qvnew, qcnew
        den = ((RD * (th0 + a["th3d"])) * (pi0 + a["pp3d"])) * (F(1.0) + qvnew * REPS)
        a["rho"][...] = a["prs"] / den
    cond = np.maximum(qcnew - qc0, ZERO)
    evac = np.maximum(qvnew - qv0, ZERO)
    vol = float(F(F(dx) * F(dy)) * F(dz))
    b1 = _budget_by_level(cond, a["rr"], a["ruh"], a["rvh"], a["rmh"])
    b2 = _budget_by_level(evac, a["rr"], a["ruh"], a["rvh"], a["rmh"])
    for k in range(len(b1)):
        a["qbudget"][0] += b1[k] * vol
    for k in range(len(b2)):
        a["qbudget"][1] += b2[k] * vol
    return iterations


GRAVITY = F(9.81)
EFALL_MIN = F(0.0001)


def efall(a, temperature, *, dz, dt, equation_set):
    """cpu_gm1_kessler_efall_inplace (cuda_gm1_kessler_efall.py:201-256; kessler.F:485-522,
    mp_driver.F:200-225): the mixture heat capacity, the energy-fallout tendency of every level below
    the top, and where |dt * tendency| >= 1e-4 the updated temperature, pressure, Exner perturbation and
    theta'.  In place on ``temperature`` (the driver's held temperature) and a's pp3d, th3d, prs; ``a``
    also holds mf (nk+1, nj, ni) and the fall speed.  Returns (cvm, tendency)."""
    qv, qc, qr, vt = a["qv"], a["qc"], a["qr"], a["fall_velocity"]
    if equation_set <= 1:
        cvm = np.full(qv.shape, CV, F)
    else:
        cvm = ((CV + CVV * qv) + CPL * qc) + CPL * qr
    rdz = F(1.0) / F(dz)
    vertical = CPL * (temperature[1:] - temperature[:-1])
    vertical = vertical * rdz
    vertical = vertical * a["mf"][1:-1]
    tendency = np.zeros(qv.shape, F)
    tendency[:-1] = ((qr[:-1] * vt[:-1]) * (vertical + GRAVITY)) / cvm[:-1]
    inc = F(dt) * tendency[:-1]
    hit = np.abs(inc) >= EFALL_MIN
    t_new = temperature[:-1] + inc
    prs = ((a["rho"][:-1] * RD) * t_new) * (F(1.0) + qv[:-1] * REPS)
    pp = _pow(prs * RP00, ROVCP) - a["pi0"][:-1]
    th = t_new / (a["pi0"][:-1] + pp) - a["th0"][:-1]
    for name, new in ((temperature, t_new), (a["prs"], prs), (a["pp3d"], pp), (a["th3d"], th)):
        name[:-1] = np.where(hit, new, name[:-1])
    return cvm, tendency


def kessler_driver(a, *, dx, dy, dz, dt, equation_set=2, budgets=True, efall_on=0):
    """Every phase of the ptype=1 driver in place on the mass arrays of ``a``.

    ``a`` holds (nk, nj, ni) float32 ruh/rvh/rmh/mh/pi0/th0/rho/rr/pp3d/th3d/
    prs/qv/qc/qr/fall_velocity, rain (1, nj, ni), prate (nj, ni) and float64
    asq (3,), qbudget (10,).  With ``budgets=False`` the binary64 diagnostics
    are skipped (they never feed back into the fields).  ``efall_on=1`` runs
    the energy fallout (``efall``, which needs a["mf"]) between the
    surface-energy budget and the mass fallout, and the local conversion then
    reads its held temperature (cuda_gm1_kessler_driver.py:482-551).
    """
    dx32, dy32, dz32, dt32 = F(dx), F(dy), F(dz), F(dt)
    a["dx"], a["dy"] = dx32, dy32
    temperature = (a["th0"] + a["th3d"]) * (a["pi0"] + a["pp3d"])
    vol = float(F(dx32 * dy32) * dz32)
    for q, rmax, idx in (("qv", 0.0, 0), ("qc", 1.0e-14, 1), ("qr", 1.0e-14, 2)):
        pdefq(a[q], a["rho"], a["ruh"], a["rvh"], a["rmh"], rmax=rmax, volume=vol, asq=a["asq"], index=idx)
    a["fall_velocity"][...] = fall_speed(a["rho"], a["qr"])
    if budgets:
        term = dt32 * a["fall_velocity"][0]
        term = term * a["rr"][0]
        term = term * a["qr"][0]
        term = term * (CPL * temperature[0] - LV1)
        term = term * a["ruh"][None, :]
        term = term * a["rvh"][:, None]
        nj, ni = term.shape
        for j in range(nj):
            bud = 0.0
            for i in range(ni):
                bud += float(term[j, i])
            a["qbudget"][6] += bud * float(dx32) * float(dy32)
    held = temperature.copy()
    energy = efall(a, held, dz=dz32, dt=dt32, equation_set=equation_set) if efall_on else None
    nfall = fallout(a, dt=dt32, dz=dz32)
    local = local_conversion(a, dt=dt32, equation_set=equation_set,
                             temperature_override=held if efall_on else None)
    if budgets:
        for p, name in enumerate(("autoconversion", "accretion", "evaporation")):
            lv = _budget_by_level(local[name], a["rr"], a["ruh"], a["rvh"], a["rmh"])
            for k in range(len(lv)):
                a["qbudget"][2 + p] = a["qbudget"][2 + p] + lv[k] * vol
    iters = satadj(a, dx=dx32, dy=dy32, dz=dz32, equation_set=equation_set)
    return {"temperature_initial": temperature, "temperature_held": held, "efall": energy, "local": local,
            "iterations": iters, "max_fall_substeps": nfall,
            "decisions": decision_words(local["branches"], iters, equation_set)}


# Bit layout of a cell's branch decisions (the device driver writes the same word).
DECISION_BITS = ("qc_positive", "autoconverting", "accreting", "evaporating", "evaporation_capped", "limited",
                 "updated")
DECISION_SATADJ = 1 << 7  # the saturation adjustment ran (eqtset 2: qc > 1e-12 or qv > qvs; eqtset 1: always)
DECISION_ITER_SHIFT = 8  # bits 8..15: its iteration count


def decision_words(branches, iterations, equation_set):
    """uint32 per cell: the local-conversion branches taken, whether satadj ran, and its iteration count."""
    word = np.zeros(iterations.shape, np.uint32)
    for b, name in enumerate(DECISION_BITS):
        word |= branches[name].astype(np.uint32) << np.uint32(b)
    ran = iterations > 0
    word |= np.where(ran, np.uint32(DECISION_SATADJ), np.uint32(0))
    word |= (iterations.astype(np.uint32) & np.uint32(0xFF)) << np.uint32(DECISION_ITER_SHIFT)
    return word


HELD = (("pp3d", "qpten"), ("th3d", "qtten"), ("qv", "qvten"), ("qc", "qcten"))


def eos(a):
    """prs = p00 (pi0+pp3d)^(cp/rd); rho = prs/((th0+th3d)(pi0+pp3d)(rd + max(qv,0) rv))."""
    work = a["pi0"] + a["pp3d"]
    a["prs"][...] = _pow(work, CPDRD) * P00
    rho = (a["th0"] + a["th3d"]) * work
    gas = np.maximum(a["qv"], ZERO) * RV + RD
    a["rho"][...] = a["prs"] / (rho * gas)


def large_step_transition(a, held, *, dx, dy, dz, dt, equation_set=2, hook=None, budgets=True, efall_on=0):
    """cpu_gm1_kessler_large_step_inplace: held remove/snapshot, EOS, hook, rr<-rho, driver, recapture."""
    dt32 = F(dt)
    if equation_set == 2:
        for s, h in HELD:
            a[s][...] = a[s] - held[h] * dt32
            held[h][...] = a[s]
    eos(a)
    if hook is not None:
        hook(a)
    a["rr"][...] = a["rho"]
    out = kessler_driver(a, dx=dx, dy=dy, dz=dz, dt=dt, equation_set=equation_set, budgets=budgets,
                         efall_on=efall_on)
    if equation_set == 2:
        for s, h in HELD:
            held[h][...] = (a[s] - held[h]) / dt32
    return out


def dry_mass_layers(rho, ruh, rvh, rmh):
    nk, nj, ni = rho.shape
    out = np.zeros(nk, f64)
    for k in range(nk):
        t = (((rho[k] * ruh[None, :]) * rvh[:, None]) * rmh[k]).astype(f64).ravel()
        acc = 0.0
        for v in t:
            acc += float(v)
        out[k] = acc
    return out


def dry_mass(rho, ruh, rvh, rmh):
    acc = 0.0
    for v in dry_mass_layers(rho, ruh, rvh, rmh):
        acc += float(v)
    return acc


def apmasscon(a, mass1):
    """Exner correction that restores the dry-air mass ledger (cuda_apmasscon.py:256-327)."""
    out = apmasscon_correction(a, mass1)
    apply_mass_correction(a, F(out["correction"]))
    return out


def apmasscon_correction(a, mass1):
    """The ledger (mass2, the Exner sum) and the binary32 correction apmasscon would apply, applying nothing."""
    nk, nj, ni = a["rho"].shape
    mass2 = dry_mass(a["rho"], a["ruh"], a["rvh"], a["rmh"])
    psum = 0.0
    for k in range(nk):
        t = (a["pi0"][k] + a["pp3d"][k]).astype(f64).ravel()
        acc = 0.0
        for v in t:
            acc += float(v)
        psum += acc
    corr = (np.power(f64(mass1) / f64(mass2), f64(RDDCV)) - 1.0) * psum / f64(ni * nj * nk)
    return {"mass2": mass2, "pressure_sum": psum, "correction": float(F(corr))}


def apply_mass_correction(a, corr32):
    """pp3d += correction, then prs and rho by the equation of state (cuda_apmasscon.py:300-316)."""
    a["pp3d"][...] = a["pp3d"] + F(corr32)
    pi = a["pi0"] + a["pp3d"]
    prs = P00 * _pow(pi, CPDRD)
    theta = a["th0"] + a["th3d"]
    gas = RD + np.maximum(ZERO, a["qv"]) * RV
    a["prs"][...] = prs
    a["rho"][...] = prs / ((theta * pi) * gas)
