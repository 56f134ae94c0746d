"""The synthetic initializer: the noise stream and the analytic base state.

* splitmix64: the reference's seed-0 known answer
  (``tests/test_ensemble_branch.py:58-67``) and the top-53-bit mapping of a
  word to ``amplitude*(2u-1)`` (``ensemble_branch.py:149-160``);
* the Weisman-Klemp column (``storm_initialization.rs:197-330``, CM1
  base.F:369-418): its potential temperature profile, the discrete
  hydrostatic relation its Exner function satisfies level by level (the
  first half step averaged with the surface's theta_v), 1000 hPa at the
  surface, the temperature law of the exponential theta above the
  tropopause (dT/dz = (g/cp)(T/213 - 1)), and the moisture fixed
  point qv = min(RH * rslf(p, T), 14 g/kg);
* the cos^2 bubble's amplitude and radius.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from gm1_amd import fields

G, CP, RD, RV = 9.81, 1005.7, 287.04, 461.5


def test_splitmix64_seed0_known_answer():
    words = [int(w) for w in fields.splitmix64_words(0, 4)]
    assert words == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_noise_maps_the_top_53_bits_to_a_symmetric_interval():
    first = 0xE220A8397B1DCDAF
    u = (first >> 11) * 2.0 ** -53
    got = fields.uniform_noise(0, 0.25, (4,))
    assert got[0] == np.float32(0.25 * (2.0 * u - 1.0))
    many = fields.uniform_noise(3, 0.1, (64, 64, 8))
    assert np.abs(many).max() <= np.float32(0.1) and abs(float(many.mean())) < 2e-3


def _column(cfg):
    prof = fields.profiles_for(cfg)
    return fields.base_state(cfg, prof)


@pytest.mark.parametrize("moist", [False, True])
def test_weisman_klemp_theta_profile(moist):
    cfg = fields.moist_c2(8, 8, 80) if moist else fields.dry_c1(8, 8, 40)
    base = _column(cfg)
    z = base["zs"]
    th = base["binary64"]["th"]
    below = z < 12000.0
    np.testing.assert_allclose(th[below], 300.0 + 43.0 * (z[below] / 12000.0) ** 1.25, rtol=1e-14)
    np.testing.assert_allclose(th[~below], 343.0 * np.exp(G / (213.0 * CP) * (z[~below] - 12000.0)), rtol=1e-14)
    assert base["th0"][0] == np.float32(th[0])


@pytest.mark.parametrize("moist", [False, True])
def test_base_state_is_hydrostatic_level_by_level(moist):
    """(pi_k - pi_{k-1}) = -g dz / (cp * mean theta_v), from pi = 1 (1000 hPa) at the ground."""
    cfg = fields.moist_c2(8, 8, 80) if moist else fields.dry_c1(8, 8, 40)
    base = _column(cfg)
    b = base["binary64"]
    z, pi, th, qv = base["zs"], b["pi"], b["th"], b["qv"]
    thv = th * (1.0 + qv * RV / RD) / (1.0 + qv)
    np.testing.assert_allclose(1.0 - pi[0], G * z[0] / (CP * 0.5 * (b["thv_sfc"] + thv[0])), rtol=1e-12)
    np.testing.assert_allclose(pi[:-1] - pi[1:], G * np.diff(z) / (CP * 0.5 * (thv[1:] + thv[:-1])), rtol=1e-12)
    t = th * pi
    above = np.nonzero(z > 12000.0)[0]
    if len(above) > 2:
        # theta = 343 exp(g (z - ztr)/(213 cp)) with hydrostatic pi gives dT/dz = (g/cp)(T/213 - 1):
        # isothermal at 213 K, drifting away from it at the rate the tropopause's T sets
        tm = 0.5 * (t[above][1:] + t[above][:-1])
        slope = np.diff(t[above]) / np.diff(z[above])
        np.testing.assert_allclose(slope, (G / CP) * (tm / 213.0 - 1.0), rtol=0.05)
        assert abs(t[above[0]] - 213.0) < 6.0
    # the surface theta_v average: 300 K, saturated at 1000 hPa (storm_initialization.rs:69-79)
    es = 611.2 * math.exp(17.67 * (300.0 - 273.15) / (300.0 - 29.65))
    qs = (RD / RV) * es / (1.0e5 - es)
    assert b["thv_sfc"] == pytest.approx(300.0 * (1.0 + qs * RV / RD) / (1.0 + qs), rel=1e-14)


def test_moist_column_is_a_fixed_point_of_its_humidity_profile():
    cfg = fields.moist_c2(8, 8, 80)
    base = _column(cfg)
    b = base["binary64"]
    z, pi, th, qv = base["zs"], b["pi"], b["th"], b["qv"]
    p = 1.0e5 * pi ** (CP / RD)
    t = th * pi
    es = np.minimum(611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65)), 0.5 * p)
    rslf = (RD / RV) * es / (p - es)
    rh = np.where(z <= 1200.0, 0.9, np.where(z >= 12000.0, 0.25, 0.9 - 0.65 * (z - 1200.0) / 10800.0))
    np.testing.assert_allclose(qv, np.minimum(rh * rslf, 0.014), rtol=1e-6, atol=1e-12)
    assert qv.max() == pytest.approx(0.014) and (qv[z > 12000.0] < 2e-4).all()


def test_bubble_amplitude_and_radius():
    cfg = fields.CaseConfig("bubble", 40, 40, 40, 1000.0, 1000.0, 500.0, 6.0, 6, (1, 1, 1, 1), False, "linear15",
                            ((2.0, 20000.0, 20000.0, 1250.0),), bubble_radius=(10000.0, 1400.0), noise=0.0, seed=0)
    prof = fields.profiles_for(cfg)
    base = fields.base_state(cfg, prof)
    th = fields.initial_fields(cfg, prof, base)["th3d"]
    assert th.max() == pytest.approx(2.0 * math.cos(0.5 * math.pi * math.hypot(500.0 / 10000.0, 500.0 / 10000.0)) ** 2,
                                     rel=1e-6)
    # zero outside the radius: at the domain corner and above 1250 + 1400 m
    assert th[:, 0, 0].max() == 0.0 and th[6:].max() == 0.0
