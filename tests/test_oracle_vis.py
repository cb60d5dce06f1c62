"""The oracle's visualisation channels (vis.wgsl:38-55) pinned by their
closed forms, computed here independently from the conserved state:

  p = (-E + sqrt(4 E^2 - 3 Z^2)) / 3   (root of 3p^2 + 2Ep - (E^2 - Z^2) = 0)
  v = Z / (E + p),  T = (3 p / a)^(1/4),
  channels (phi/phi_b, |v|, T/T_ref - 1, (E + p)|v|^2 / e_ref).

The bag model's reference temperature is fixed by e_ref = a T_ref^4, so a
fluid at rest at E = e_ref reads T/T_ref - 1 = 0.
"""
import numpy as np


def _oracle_with_state(bs, orc, grid, field, fluid, precision=1):
    model = bs.Model()
    params = bs.sim_params_words(model, grid)
    o = orc.OracleSim(precision, params, np.zeros((0, 5), np.float32))
    o.set_state(field, fluid)
    return o, bs.sim_params(model, grid)   # the f32 block the oracle widens


def test_vis_channels_closed_form(bs, orc):
    grid = (8, 8, 8)
    n = 512
    rng = np.random.default_rng(4)
    P = bs.sim_params(bs.Model(), grid)
    e_ref, a_rad, t_ref, phi_b = (float(P.e_ref), float(P.a_rad), float(P.t_ref), float(P.phi_b))
    E = e_ref * rng.uniform(0.5, 3.0, n)
    zdir = rng.normal(size=(n, 3))
    zdir /= np.linalg.norm(zdir, axis=1, keepdims=True)
    Z = zdir * (E * rng.uniform(0.0, 0.9, n))[:, None]   # |Z| < z_max E: no clamp
    phi = rng.uniform(-0.2, 1.2, n) * phi_b
    field = np.stack([phi, rng.normal(size=n)], axis=1).astype(np.float64)
    fluid = np.concatenate([E[:, None], Z], axis=1).astype(np.float64)
    o, _ = _oracle_with_state(bs, orc, grid, field, fluid)
    got = o.vis_channels().astype(np.float64)

    z2 = (Z ** 2).sum(axis=1)
    p = (-E + np.sqrt(4 * E ** 2 - 3 * z2)) / 3
    v = Z / (E + p)[:, None]
    v2 = (v ** 2).sum(axis=1)
    T = (3 * p / a_rad) ** 0.25
    assert abs(t_ref / (e_ref / a_rad) ** 0.25 - 1) < 1e-6   # e_ref = a T_ref^4
    want = np.stack([phi / phi_b, np.sqrt(v2), T / t_ref - 1, (E + p) * v2 / e_ref], axis=1)
    # the oracle narrows each channel to f32; the closed form is f64
    np.testing.assert_allclose(got[:, [0, 1, 3]], want[:, [0, 1, 3]], rtol=2e-7, atol=1e-30)
    np.testing.assert_allclose(got[:, 2], want[:, 2], rtol=0, atol=2e-7 * (1 + np.abs(want[:, 2])).max())
    o.close()


def test_vis_channels_at_rest_and_floored(bs, orc):
    grid = (8, 8, 8)
    P = bs.sim_params(bs.Model(), grid)
    e_ref, a_rad, t_ref = float(P.e_ref), float(P.a_rad), float(P.t_ref)
    n = 512
    field = np.zeros((n, 2))
    fluid = np.zeros((n, 4))
    fluid[:, 0] = e_ref
    fluid[0, 0] = -1.0        # below the energy floor: E -> e_floor
    o, _ = _oracle_with_state(bs, orc, grid, field, fluid)
    got = o.vis_channels()
    assert np.all(got[1:, 0] == 0) and np.all(got[:, 1] == 0) and np.all(got[:, 3] == 0)
    T = (e_ref / a_rad) ** 0.25               # p = E / 3 at rest
    assert np.abs(got[1:, 2] - (T / t_ref - 1)).max() < 1e-12
    assert abs(T / t_ref - 1) < 1e-6          # T = T_ref at E = e_ref
    # floored cell: E -> e_floor, p = max(E^2 / (E + 2E), p_floor) = p_floor
    e_floor, p_floor = float(P.e_floor), float(P.p_floor)
    p = max(e_floor / 3, p_floor)
    want = (3 * p / a_rad) ** 0.25 / t_ref - 1
    assert abs(got[0, 2] - want) < 1e-6
    o.close()
