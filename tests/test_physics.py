"""Model derivations and validation (physics.rs:44-218, config.rs:151-178).

The reference's own assertions (physics.rs:316-385, config.rs:225-242) are
ported; the closed forms of docs/PHYSICS.md:59-80 are checked in float64."""
import math

import numpy as np
import pytest

F = np.float32


def pot(m, d, p):
    # V and V' in float32, as the reference's test closures (physics.rs:320-321)
    p = F(p)
    v = F(0.5) * F(d.m2) * p * p - F(d.delta) * p * p * p / F(3.0) + F(0.25) * F(m.lambda_) * p ** 4
    dv = p * (F(d.m2) - F(d.delta) * p + F(m.lambda_) * p * p)
    return v, dv


def test_potential_has_the_requested_shape(bs):
    m = bs.Model()
    d = m.derived()
    assert d.m2 > 0.0, "symmetric minimum must be metastable"
    v, dv = pot(m, d, m.phi_b)
    assert abs(dv) < 1e-6, "phi_b must be a stationary point"
    assert abs(v + d.eps) < 1e-6
    top = max(pot(m, d, m.phi_b * i / 100.0)[0] for i in range(100))
    assert top > 0.0, "a first-order transition needs a barrier"


def test_degenerate_limit_is_the_symmetric_double_well(bs):
    d = bs.Model(eps_ratio=1e-6).derived()
    assert abs(d.m2 - 0.5 * 0.5) < 1e-4
    assert abs(d.delta - 1.5 * 0.5) < 1e-4


@pytest.mark.parametrize("eps_ratio", [0.05, 0.2, 0.3, 0.6, 0.95])
@pytest.mark.parametrize("lam", [0.1, 0.5, 1.5])
def test_critical_radius_is_never_smaller_than_the_wall(bs, eps_ratio, lam):
    d = bs.Model(eps_ratio=eps_ratio, lambda_=lam).derived()
    ratio = d.critical_radius / d.wall_width
    assert abs(ratio - d.critical_to_wall_ratio) < 1e-3 * ratio
    assert ratio > 1.0


def test_seed_radius_follows_the_chosen_mode(bs):
    m = bs.Model(seed_factor=1.15)
    d = m.derived()
    assert abs(d.seed_radius - 1.15 * d.critical_radius) < 1e-4
    assert bs.Model(eps_ratio=0.6, seed_factor=1.15).derived().seed_radius < d.seed_radius
    fixed = bs.Model(seed_cells=7.5)
    assert abs(fixed.derived().seed_radius - 7.5) < 1e-6
    assert abs(bs.Model(eps_ratio=0.6, seed_cells=7.5).derived().seed_radius - 7.5) < 1e-6
    # floor at 0.25 dx (physics.rs:159)
    assert bs.Model(seed_cells=0.1).derived().seed_radius == F(0.25)


def test_sub_critical_seeds_are_allowed_so_collapse_can_be_shown(bs):
    m = bs.Model(seed_factor=0.5)
    m.validate()
    d = m.derived()
    assert d.seed_radius < d.critical_radius


def test_defaults_match_the_documented_values(bs):
    # physics.rs:70-140 defaults: eps 0.0125, m2 0.175, delta 0.675,
    # l_w 4.0, sigma 0.0833, R_c 13.33, R_0 15.33, e_ref = a_rad 0.125, dt 0.2
    d = bs.Model().derived()
    exp = dict(eps=0.0125, m2=0.175, delta=0.675, wall_width=4.0, surface_tension=1 / 12, critical_radius=40 / 3,
               seed_radius=1.15 * 40 / 3, e_ref=0.125, a_rad=0.125, dt=0.2, t_ref=1.0, sound_speed=1 / math.sqrt(3))
    for k, v in exp.items():
        assert abs(getattr(d, k) - v) <= 2e-6 * max(1.0, abs(v)), k


@pytest.mark.parametrize("lam,er,alpha", [(0.5, 0.3, 0.1), (0.3, 0.7, 0.05), (1.2, 0.15, 0.3)])
def test_closed_forms_in_float64(bs, lam, er, alpha):
    # docs/PHYSICS.md:59-80 evaluated in double, phi_b = dx = 1
    d = bs.Model(lambda_=lam, eps_ratio=er, alpha=alpha).derived()
    lam, er, alpha = float(F(lam)), float(F(er)), float(F(alpha))
    eps = er * lam / 12
    sig = math.sqrt(lam / 2) / 6
    checks = dict(eps=eps, m2=0.5 * lam - 6 * eps, delta=1.5 * lam - 6 * eps, wall_width=2 * math.sqrt(2 / lam),
                  surface_tension=sig, critical_radius=4 / (er * math.sqrt(2 * lam)), e_ref=eps / alpha)
    for k, v in checks.items():
        assert abs(getattr(d, k) - v) <= 1e-5 * abs(v), (k, getattr(d, k), v)
    # the derived potential: V'(1) = 0 and V(1) = -eps, in double
    v1 = 0.5 * d.m2 - d.delta / 3 + 0.25 * lam
    dv1 = d.m2 - d.delta + lam
    assert abs(v1 + eps) < 1e-6 and abs(dv1) < 1e-6


def test_estimated_wall_velocity(bs):
    d = bs.Model().derived()
    x = d.eps / (0.3 * d.surface_tension)
    assert abs(d.estimated_wall_velocity - x / math.sqrt(1 + x * x)) < 1e-6
    assert 0.4 < d.estimated_wall_velocity < 0.5  # docs/NUMERICS.md:191-193 quotes 0.45


@pytest.mark.parametrize("field,val,msg", [
    ("eps_ratio", 1.0, "eps_ratio must lie in (0, 1)"),
    ("eps_ratio", 0.0, "eps_ratio must lie in (0, 1)"),
    ("lambda_", 0.0, "lambda must be positive"),
    ("alpha", -1.0, "alpha must be positive"),
    ("eta", -0.1, "eta must be non-negative"),
    ("cfl", 0.51, "cfl must lie in (0, 0.5]"),
    ("cfl", 0.0, "cfl must lie in (0, 0.5]"),
    ("seed_factor", 0.0, "seed factor must be positive"),
])
def test_validate_rejects(bs, field, val, msg):
    m = bs.Model(**{field: val})
    with pytest.raises(bs.BubbleSimError) as e:
        m.validate()
    assert msg in e.value.msg


def test_validate_accepts_limits(bs):
    bs.Model(cfl=0.5).validate()
    bs.Model(eta=0.0).validate()
    with pytest.raises(bs.BubbleSimError, match="seed radius must be positive"):
        bs.Model(seed_cells=0.0).validate()


def test_config_validate(bs):
    # config.rs:151-178 and its tests (:225-242)
    s = bs.make_spec()
    bs.validate_spec(s)
    m = bs.Model()
    d = m.derived()
    assert d.wall_width >= 3.0 and d.critical_radius > d.wall_width
    with pytest.raises(bs.BubbleSimError, match="cells thick"):
        bs.validate_spec(bs.make_spec(model=bs.Model(lambda_=50.0)))
    with pytest.raises(bs.BubbleSimError, match="--bubbles must be at least 1"):
        bs.validate_spec(bs.make_spec(bubbles=0))
    with pytest.raises(bs.BubbleSimError, match="--nucleation-duration must be positive"):
        bs.validate_spec(bs.make_spec(duration=0.0))
    with pytest.raises(bs.BubbleSimError, match="grid axis 1 is 7"):
        bs.validate_spec(bs.make_spec(grid=(8, 7, 8)))
    bs.validate_spec(bs.make_spec(grid=(8, 8, 8)))
    # an x-y plane is addressed with 32-bit byte offsets: 2^28 cells of
    # 16 B (f32) or 2^27 cells of 32 B (f64) is one too many
    with pytest.raises(bs.BubbleSimError, match="x-y plane of 268435456 cells"):
        bs.validate_spec(bs.make_spec(grid=(16384, 16384, 8)))
    bs.validate_spec(bs.make_spec(grid=(16384, 16383, 8)))
    with pytest.raises(bs.BubbleSimError, match="nx \\* ny must stay below 134217728"):
        bs.validate_spec(bs.make_spec(grid=(16384, 8192, 8), precision=bs.PRECISION_F64))
    bs.validate_spec(bs.make_spec(grid=(16384, 8191, 8), precision=bs.PRECISION_F64))


def test_grid_accepts_one_or_three_values(bs):
    # grid_accepts_one_or_three_values (config.rs:217-223)
    assert bs.parse_grid("192") == (192, 192, 192)
    assert bs.parse_grid("256,128, 64") == (256, 128, 64)
    for bad in ("1,2", "abc", "", "1,2,3,4", "-5"):
        with pytest.raises(bs.BubbleSimError):
            bs.parse_grid(bad)
