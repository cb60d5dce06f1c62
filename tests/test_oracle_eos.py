"""The oracle's equation-of-state swap point (docs/EOS.md) pinned without
the code under test:

* the gamma-law fluid p = (Gamma - 1) e at Gamma = 4/3 is the bag model: its
  Newton inversion must reproduce the closed-form root of
  3p^2 + 2Ep - (E^2 - Z^2) = 0;
* at other Gamma the recovered pressure is the root of
  p = (Gamma - 1)(E - Z^2 / (E + p)), found here by bisection;
* cons(prim(U)) = U; at rest the wave speeds are -+sqrt(Gamma - 1);
* the dof-ratio temperature interpolates a_sym -> a_broken by smoothstep.
"""
import math

import numpy as np
import pytest


def _words(bs, orc, grid=(8, 8, 8), **kw):
    return orc.eos_words(bs.sim_params_words(bs.Model(), grid), **kw)


def _states(bs, n=200, seed=7, vmax=0.97):
    rng = np.random.default_rng(seed)
    P = bs.sim_params(bs.Model(), (8, 8, 8))
    E = float(P.e_ref) * rng.uniform(0.3, 5.0, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Z = d * (E * rng.uniform(0.0, vmax, n))[:, None]
    return E, Z


def _root(E, z2, gamma):
    # bisection on f(p) = p - (gamma - 1)(E - z2 / (E + p)), increasing in p
    lo, hi = 0.0, (gamma - 1) * E
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        f = mid - (gamma - 1) * (E - z2 / (E + mid))
        lo, hi = (mid, hi) if f < 0 else (lo, mid)
    return 0.5 * (lo + hi)


def test_gamma_law_at_four_thirds_is_the_bag_model(bs, orc):
    E, Z = _states(bs)
    wb = _words(bs, orc)
    wg = _words(bs, orc, eos=1, eos_gamma=4.0 / 3.0)
    worst = 0.0
    for e, z in zip(E, Z):
        u = np.array([e, *z])
        qb = orc.prim_from_cons(wb, u)
        qg = orc.prim_from_cons(wg, u)
        worst = max(worst, float(np.max(np.abs(qg - qb) / np.maximum(np.abs(qb), 1e-300))))
    # f32 gamma (4/3 rounded) widened to f64: the two agree to its rounding
    assert worst < 1e-6


def test_gamma_law_pressure_is_the_root(bs, orc):
    for gamma in (1.2, 1.5, 1.8):
        g32 = float(np.float32(gamma))
        E, Z = _states(bs, seed=int(gamma * 10))
        w = _words(bs, orc, eos=1, eos_gamma=gamma)
        for e, z in zip(E, Z):
            q = orc.prim_from_cons(w, np.array([e, *z]))
            want = _root(e, float(z @ z), g32)
            assert q[3] == pytest.approx(want, rel=1e-12), (gamma, e, z)
            # v = Z / (E + p)
            np.testing.assert_allclose(q[:3], z / (e + want), rtol=1e-12, atol=1e-300)


def test_gamma_law_round_trip_and_sound_speed(bs, orc):
    gamma = 1.55
    g32 = float(np.float32(gamma))
    w = _words(bs, orc, eos=1, eos_gamma=gamma)
    E, Z = _states(bs, seed=3, vmax=0.9)
    for e, z in zip(E, Z):
        u = np.array([e, *z])
        back = orc.cons_from_prim(w, orc.prim_from_cons(w, u))
        np.testing.assert_allclose(back, u, rtol=1e-11, atol=1e-12 * e)
    lo, hi = orc.wave_speeds(w, np.array([0.0, 0.0, 0.0, 0.3]), 0)
    assert hi == pytest.approx(math.sqrt(g32 - 1), rel=1e-12) and lo == pytest.approx(-hi, rel=1e-12)


def test_dof_ratio_temperature(bs, orc):
    grid = (8, 8, 8)
    P = bs.sim_params(bs.Model(), grid)
    phi_b, a_rad, t_ref, e_ref = float(P.phi_b), float(P.a_rad), float(P.t_ref), float(P.e_ref)
    ratio = 0.6
    n = 512
    hs = np.linspace(-0.2, 1.2, n)
    field = np.stack([hs * phi_b, np.zeros(n)], axis=1)
    fluid = np.zeros((n, 4))
    fluid[:, 0] = e_ref
    o = orc.OracleSim(1, _words(bs, orc, grid, dof_ratio=ratio), np.zeros((0, 5), np.float32))
    o.set_state(field, fluid)
    got = o.vis_channels()[:, 2].astype(np.float64)
    h = np.clip(hs, 0, 1)
    a = a_rad * (1 + (ratio - 1) * h * h * (3 - 2 * h))
    want = (e_ref / a) ** 0.25 / t_ref - 1          # p = E / 3 at rest
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-7 * np.abs(want).max() + 1e-7)
    # the symmetric side reads the bag temperature, the broken side a_b
    assert abs(got[0] - ((e_ref / a_rad) ** 0.25 / t_ref - 1)) < 1e-6
    o.close()
