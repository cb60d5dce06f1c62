"""Pins the oracle (oracle/oracle.cpp) against mathematics, not against itself.

Each check is chosen so that a plausible restatement mistake -- a dropped
term, a wrong sign or index, a transposed operand -- fails it:
* primitive recovery: the quadratic 3p^2 + 2Ep - (E^2 - Z^2) = 0, the
  textbook root, and the cons(prim(U)) = U round trip;
* wave speeds: relativistic velocity addition (v +- c)/(1 +- v c) for flow
  along the axis, +-c_s at rest;
* flux / HLL: the physical flux written out, HLL consistency F(U,U) = F(U),
  pure upwinding for supersonic faces;
* MC limiter: hand-worked values;
* tanh: libm within 4 ulp (f64) / 3 ulp (f32);
* the lattice: quiescent fixed point, exact periodic translation
  invariance, mirror and axis-permutation symmetry of a centred bubble to
  roundoff, energy conservation, f32 -> f64 convergence, and reduce.wgsl's
  naive f32 sums against long-double sums.
"""
import math

import numpy as np
import pytest

F = np.float32


@pytest.fixture(scope="module")
def params(bs):
    return bs.sim_params_words(bs.Model(), (32, 32, 32))


def rand_states(n, seed=0):
    rng = np.random.default_rng(seed)
    E = rng.uniform(0.05, 2.0, n)
    frac = rng.uniform(0, 0.98, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Z = d * (frac * E)[:, None]
    return np.column_stack([E, Z])


def test_prim_recovery_solves_the_quadratic(orc, params):
    for u in rand_states(300):
        q = orc.prim_from_cons(params, u)
        E, Z2 = u[0], float(np.dot(u[1:], u[1:]))
        p = q[3]
        assert abs(3 * p * p + 2 * E * p - (E * E - Z2)) <= 1e-13 * E * E
        assert abs(p - (math.sqrt(4 * E * E - 3 * Z2) - E) / 3) <= 1e-12 * E
        np.testing.assert_allclose(q[:3], u[1:] / (E + p), rtol=1e-14, atol=1e-300)
        np.testing.assert_allclose(orc.cons_from_prim(params, q), u, rtol=1e-12, atol=1e-14)


def test_prim_recovery_regularises(orc, params):
    # |Z| > z_max E is clamped on the local copy: |v| < 1, p >= p_floor
    q = orc.prim_from_cons(params, np.array([1.0, 2.0, 0.0, 0.0]))
    assert 0 < q[0] < 1 and q[3] >= F(1e-7) * F(0.125) * 0.999
    q = orc.prim_from_cons(params, np.array([-1.0, 0.0, 0.0, 0.0]))
    e_floor = float(F(F(1e-7) * F(0.125)))
    assert q[3] == pytest.approx(e_floor / 3, rel=1e-6) or q[3] == pytest.approx(e_floor, rel=1e-6)


def test_wave_speeds(orc, params):
    cs = math.sqrt(1 / 3)
    for axis in range(3):
        lo, hi = orc.wave_speeds(params, np.array([0, 0, 0, 0.1]), axis)
        assert lo == pytest.approx(-cs, abs=1e-15) and hi == pytest.approx(cs, abs=1e-15)
        for v in (0.3, -0.7, 0.95):
            q = np.zeros(4)
            q[axis] = v
            q[3] = 0.2
            lo, hi = orc.wave_speeds(params, q, axis)
            assert lo == pytest.approx((v - cs) / (1 - v * cs), rel=1e-12)
            assert hi == pytest.approx((v + cs) / (1 + v * cs), rel=1e-12)
    lo, hi = orc.wave_speeds(params, np.array([0.0, 0.999, 0.0, 0.1]), 1)
    assert hi > 0.999 and lo > 0.98  # both tend to +1 as v -> 1 along the axis


def test_physical_flux(orc, params):
    for u in rand_states(50, 3):
        q = orc.prim_from_cons(params, u)
        uu = orc.cons_from_prim(params, q)
        for axis in range(3):
            f = orc.eos_flux(q, uu, axis)
            want = np.array([uu[1 + axis], uu[1] * q[axis], uu[2] * q[axis], uu[3] * q[axis]])
            want[1 + axis] += q[3]
            np.testing.assert_allclose(f, want, rtol=1e-14, atol=1e-300)


def test_hll_consistency_and_upwinding(orc, params):
    for u in rand_states(50, 5):
        q = orc.prim_from_cons(params, u)
        uu = orc.cons_from_prim(params, q)
        for axis in range(3):
            np.testing.assert_allclose(orc.hll_flux(params, q, q, axis), orc.eos_flux(q, uu, axis),
                                       rtol=1e-12, atol=1e-15)
    # supersonic to the right: every signal speed > 0 -> the left flux
    ql = np.array([0.9, 0.0, 0.0, 0.1])
    qr = np.array([0.8, 0.05, 0.0, 0.12])
    ul = orc.cons_from_prim(params, ql)
    np.testing.assert_allclose(orc.hll_flux(params, ql, qr, 0), orc.eos_flux(ql, ul, 0), rtol=1e-12)
    ql2, qr2 = -qr * np.array([1, 1, 1, -1]), -ql * np.array([1, 1, 1, -1])
    ur2 = orc.cons_from_prim(params, qr2)
    np.testing.assert_allclose(orc.hll_flux(params, ql2, qr2, 0), orc.eos_flux(qr2, ur2, 0), rtol=1e-12)


@pytest.mark.parametrize("dm,dp,want", [(1, 1, 1), (1, 3, 2), (3, 1, 2), (-1, 2, 0), (0, 5, 0), (-2, -8, -4),
                                        (2, 2.5, 2.25), (1e-3, -1e-3, 0), (-0.5, -0.4, -0.45)])
def test_mc_limiter(orc, dm, dp, want):
    assert orc.limit_slope(dm, dp) == pytest.approx(want, abs=1e-15)
    assert orc.limit_slope_f32(dm, dp) == pytest.approx(want, abs=1e-6)


def test_tanh_against_libm(orc):
    xs = np.concatenate([np.linspace(-12, 12, 4001), np.geomspace(1e-8, 0.5, 400), -np.geomspace(1e-8, 0.5, 400)])
    for x in xs:
        t = orc.tanh_f64(float(x))
        ref = math.tanh(float(x))
        assert abs(t - ref) <= 4 * math.ulp(ref) + 1e-300
        t32 = orc.tanh_f32(float(F(x)))
        r32 = F(math.tanh(float(F(x))))
        assert abs(float(t32) - float(r32)) <= 3 * float(np.spacing(abs(r32))) + 1e-45


def test_potential_shape(orc, bs):
    p = bs.sim_params_words(bs.Model(), (8, 8, 8))
    d = bs.Model().derived()
    assert abs(orc.dpotential(p, 1.0)) < 1e-7
    assert orc.potential(p, 1.0) == pytest.approx(-d.eps, rel=1e-6)
    h = 1e-6
    for phi in (0.1, 0.4, 0.77, 1.3):
        num = (orc.potential(p, phi + h) - orc.potential(p, phi - h)) / (2 * h)
        assert orc.dpotential(p, phi) == pytest.approx(num, rel=1e-6, abs=1e-9)


# ---------------------------------------------------------------------------
#  Lattice invariants
# ---------------------------------------------------------------------------
def run(bs, orc, grid, precision, events, steps, model=None):
    model = model or bs.Model()
    o = orc.OracleSim(precision, bs.sim_params_words(model, grid), np.asarray(events, np.float32).reshape(-1, 5))
    o.advance(steps)
    return o


def test_quiescent_is_a_fixed_point(bs, orc):
    o = run(bs, orc, (16, 12, 10), 0, [], 6)
    f, u = o.state()
    assert np.all(f == 0) and np.all(u[..., 0] == F(0.125)) and np.all(u[..., 1:] == 0)
    s = o.diagnostics_sums()
    assert s[4] == 0.0 and s[2] == 0.0


def test_periodic_translation_invariance(bs, orc):
    g = (20, 16, 12)
    r = bs.Model().seed_radius()
    a = run(bs, orc, g, 1, [(0, 10, 8, 6, r)], 5).state()
    b = run(bs, orc, g, 1, [(0, 13, 3, 11, r)], 5).state()  # shifted by (3, -5, 5) mod grid
    for x, y in zip(a, b):
        assert np.array_equal(np.roll(x, shift=(5, -5, 3), axis=(0, 1, 2)), y)


def test_centred_bubble_symmetries(bs, orc):
    n = 24
    r = bs.Model().seed_radius() * 0.5
    o = run(bs, orc, (n, n, n), 1, [(0, n // 2, n // 2, n // 2, r)], 8)
    f, u = o.state()
    phi = f[..., 0]
    scale = np.abs(phi).max()
    # mirror x -> n - x (cells are at integer positions; the centre is at n/2)
    mir = lambda a: np.roll(a[:, :, ::-1], 1, axis=2)
    assert np.abs(mir(phi) - phi).max() <= 1e-12 * scale
    np.testing.assert_allclose(mir(u[..., 1]), -u[..., 1], atol=1e-12 * np.abs(u[..., 1]).max())
    # axis permutation x <-> y: phi symmetric, Zx <-> Zy
    perm = lambda a: np.swapaxes(a, 1, 2)
    assert np.abs(perm(phi) - phi).max() <= 1e-12 * scale
    np.testing.assert_allclose(perm(u[..., 1]), u[..., 2], atol=1e-12 * np.abs(u[..., 1]).max())
    assert np.abs(u[..., 1]).max() > 1e-8  # the fluid does move


def test_energy_drift_is_spatial_not_temporal(bs, orc):
    # Total energy drifts only through the spatial discretisation (the
    # diagnostic's central gradient is not the 7-point Laplacian's conserved
    # form): ~1e-4 over 40 steps here, and halving dt changes it by < 3 %, as
    # third-order time stepping must (a wrong RK3 coefficient makes it O(dt)).
    n = 32
    r = bs.Model().seed_radius() * 0.6
    drift = []
    for cfl, steps in ((0.2, 40), (0.1, 80)):
        m = bs.Model(cfl=cfl)
        o = run(bs, orc, (n, n, n), 1, [(0, 16, 16, 16, r)], 1, model=m)
        e0 = o.diagnostics_sums()[0]
        o.advance(steps)
        drift.append(o.diagnostics_sums()[0] / e0 - 1)
    assert abs(drift[0]) < 2e-4
    assert abs(drift[0] - drift[1]) < 0.03 * abs(drift[0])


def test_f32_converges_to_f64(bs, orc):
    g = (20, 20, 16)
    r = bs.Model().seed_radius() * 0.6
    ev = [(0, 10, 10, 8, r), (0.4, 2, 3, 4, r)]
    a = run(bs, orc, g, 0, ev, 12).state()
    b = run(bs, orc, g, 1, ev, 12).state()
    for x, y in zip(a, b):
        for c in range(x.shape[-1]):
            ref = y[..., c]
            if np.all(ref == 0):
                assert np.all(x[..., c] == 0)
                continue
            err = np.sqrt(np.sum((x[..., c] - ref) ** 2) / np.sum(ref ** 2))
            assert err < 1e-5, (c, err)


def test_naive_f32_reduction_vs_long_double(bs, orc):
    g = (24, 20, 16)
    r = bs.Model().seed_radius() * 0.6
    o = run(bs, orc, g, 0, [(0, 12, 10, 8, r)], 3)
    exact = o.diagnostics_sums()
    naive, bits = o.diagnostics_naive_f32()
    np.testing.assert_allclose(naive, exact[:4], rtol=1e-5)
    assert np.frombuffer(np.uint32(bits).tobytes(), dtype=np.float32)[0] == F(exact[4])
