"""Nucleation schedule (physics.rs:248-308): bit-exact positions and times.

Pins, in order of strength:
* the generator core against published known-answer values of rand_pcg /
  the PCG reference suite (tests/golden/rng_published.json);
* the library's C++ schedule against an independent Python restatement
  (oracle/schedule.py: Python integers for the 128-bit generator, numpy
  float32 for every f32 operation) on committed fixtures and random cases;
* the reference's own test (schedule_respects_ordering_and_bounds,
  physics.rs:387-401) and the rejection rule (2.5 R0 periodic separation).
The f32 conversion `(next_u32 >> 8) * 2^-24` of rand 0.9's StandardUniform
is pinned by rand's own value-stability vector (three f32 draws from its
PCG32 test generator; the f64 conversion beside it fixes the order in which
PCG32 composes a u64 from two u32 words, so a shifted or swapped word fails).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import schedule as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = {"simultaneous": 0, "exponential": 1, "uniform-steps": 2}


def test_mcg128_published_vectors(bs):
    pub = json.load(open(os.path.join(GOLDEN, "rng_published.json")))
    want = [int(h, 16) for h in pub["mcg128_state42"]]
    assert [int(v) for v in bs.mcg128_from_state(42, 6)] == want
    r = S.Mcg128Xsl64(42)
    assert [r.next_u64() for _ in range(6)] == want
    seed = bytes(range(1, 17))
    st = int.from_bytes(seed, "little")
    assert int(bs.mcg128_from_state(st, 1)[0]) == int(pub["from_seed_1_to_16"])
    assert int(bs.pcg64mcg_u64(0, 1)[0]) == int(pub["seed_from_u64_0"])
    assert S.Mcg128Xsl64.seed_from_u64(0).next_u64() == int(pub["seed_from_u64_0"])


def test_streams_match_fixture(bs):
    kat = json.load(open(os.path.join(GOLDEN, "schedule_kat.json")))
    for s in kat["streams"]:
        assert [int(v) for v in bs.pcg64mcg_u64(s["seed"], 8)] == [int(v) for v in s["u64"]]
        bits = bs.pcg64mcg_f32(s["seed"], 8).view(np.uint32).tolist()
        assert bits == s["f32_bits"]


def test_f32_draws_are_24_bit_fractions(bs):
    u = bs.pcg64mcg_u64(5, 1000)
    f = bs.pcg64mcg_f32(5, 1000)
    expect = ((u & np.uint64(0xFFFFFFFF)) >> np.uint64(8)).astype(np.float64) * 2.0 ** -24
    assert np.array_equal(f.astype(np.float64), expect)
    assert f.min() >= 0.0 and f.max() < 1.0


class _Pcg32:
    """rand_pcg Lcg64Xsh32 (PCG-XSH-RR 64/32), the generator of rand's own
    float value-stability test. Pinned by the PCG demo vector below."""
    MUL = 6364136223846793005
    M64 = (1 << 64) - 1

    def __init__(self, state: int, stream: int):
        self.inc = ((stream << 1) | 1) & self.M64
        self.state = (state + self.inc) & self.M64
        self._step()

    def _step(self):
        self.state = (self.state * self.MUL + self.inc) & self.M64

    def next_u32(self) -> int:
        s = self.state
        self._step()
        rot = s >> 59
        xsh = (((s >> 18) ^ s) >> 27) & 0xFFFFFFFF
        return ((xsh >> rot) | (xsh << ((32 - rot) & 31))) & 0xFFFFFFFF

    def next_u64(self) -> int:
        lo = self.next_u32()
        return (self.next_u32() << 32) | lo


def test_f32_draw_matches_rand_value_stability_vector():
    """rand 0.9 src/distr/float.rs value_stability: the oracle's f32 draw
    (the method the schedule uses, driven by a PCG32 stream) reproduces
    rand's published samples bit for bit."""
    pub = json.load(open(os.path.join(GOLDEN, "rng_published.json")))["rand_float_value_stability"]
    r = _Pcg32(42, 54)
    assert [r.next_u32() for _ in range(6)] == [int(h, 16) for h in pub["pcg32_42_54"]]
    state, stream = int(pub["pcg32_state"], 16), int(pub["pcg32_stream"])
    r = _Pcg32(state, stream)
    got = np.array([S.Mcg128Xsl64.random_f32(r) for _ in range(3)], dtype=np.float32)
    want = np.array([np.float32(v) for v in pub["f32"]], dtype=np.float32)
    assert got.view(np.uint32).tolist() == want.view(np.uint32).tolist()
    r = _Pcg32(state, stream)
    assert [(r.next_u64() >> 11) * 2.0 ** -53 for _ in range(3)] == [float(v) for v in pub["f64"]]


def test_library_f32_draw_equals_oracle_draw(bs):
    for seed in (0, 5, 0x5EED, 2**63 + 7):
        r = S.Mcg128Xsl64.seed_from_u64(seed)
        want = np.array([r.random_f32() for _ in range(64)], dtype=np.float32)
        assert bs.pcg64mcg_f32(seed, 64).view(np.uint32).tolist() == want.view(np.uint32).tolist()


def _lib_schedule(bs, c):
    return bs.make_schedule(MODES[c["mode"]], c["count"], c["radius"], c["grid"], c["duration"], c["seed"],
                            c.get("uniform_max_step", 200), c.get("time_seed", 0x5EED), c.get("dt", 0.2))


def test_schedule_matches_fixture_bitwise(bs):
    kat = json.load(open(os.path.join(GOLDEN, "schedule_kat.json")))
    for c in kat["cases"]:
        got = _lib_schedule(bs, c).view(np.uint32).tolist()
        assert got == c["events_hex"], c["mode"]


@pytest.mark.parametrize("seed", [1, 2, 3, 17, 99, 123456789, 2**63 + 5])
@pytest.mark.parametrize("mode", ["simultaneous", "exponential", "uniform-steps"])
def test_schedule_matches_python_restatement(bs, seed, mode):
    rng = np.random.default_rng(seed % 2**32)
    grid = [int(v) for v in rng.integers(8, 200, size=3)]
    count = int(rng.integers(1, 40))
    radius = float(np.float32(rng.uniform(0.3, 30.0)))
    dur = float(np.float32(rng.uniform(1.0, 100.0)))
    c = dict(mode=mode, count=count, radius=radius, grid=grid, duration=dur, seed=seed,
             uniform_max_step=int(rng.integers(0, 300)), time_seed=int(rng.integers(0, 2**63)), dt=0.2)
    a = _lib_schedule(bs, c)
    b = S.make_schedule(mode, count, radius, grid, dur, seed, c["uniform_max_step"], c["time_seed"], 0.2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_schedule_respects_ordering_and_bounds(bs):
    # physics.rs:387-401
    grid = [64, 64, 64]
    ev = bs.make_schedule(1, 12, 8.0, grid, 40.0, 7)
    assert len(ev) == 12
    assert np.all(np.diff(ev[:, 0]) >= 0)
    assert np.all((ev[:, 0] >= 0) & (ev[:, 0] <= 40.0 + 1e-3))
    for axis, n in enumerate(grid):
        assert np.all((ev[:, 1 + axis] >= 0) & (ev[:, 1 + axis] <= n))


def test_exponential_times_closed_form(bs):
    # t_k = dur/4 ln(1 + (k + 1/2)/n (e^4 - 1)), checked in float64 (physics.rs:265-271)
    n, dur = 12, 60.0
    ev = bs.make_schedule(1, n, 15.0, 128, dur, 1)
    ref = [dur / 4 * math.log(1 + (k + 0.5) / n * (math.exp(4) - 1)) for k in range(n)]
    np.testing.assert_allclose(ev[:, 0], ref, rtol=1e-6)


def test_rejection_separates_bubbles(bs):
    # every accepted bubble is > 2.5 R0 from all earlier ones (physics.rs:276-289)
    # when the box has room; with no room the last candidate is kept
    r = 6.0
    ev = bs.make_schedule(0, 10, r, 128, 60.0, 3)
    pts = ev[:, 1:4].astype(np.float64)
    for i in range(len(pts)):
        for j in range(i):
            d = np.abs(pts[i] - pts[j])
            d = np.minimum(d, 128 - d)
            assert np.sqrt((d ** 2).sum()) > 2.5 * r - 1e-4
    crowded = bs.make_schedule(0, 50, 20.0, 16, 60.0, 3)
    assert len(crowded) == 50


def test_uniform_steps_keep_reference_positions(bs):
    # the extension draws times from a separate stream: positions equal the
    # simultaneous-mode positions of the same seed, up to the time sort
    a = bs.make_schedule(0, 12, 15.0, 128, 60.0, 1)
    b = bs.make_schedule(2, 12, 15.0, 128, 60.0, 1, 200, 11, 0.2)
    key = lambda e: tuple(e[1:4])
    assert sorted(map(key, a)) == sorted(map(key, b))
    steps = bs.event_steps(b, 0.2)
    assert np.all(steps <= 200)
    for row, s in zip(b, steps):
        assert row[0] == S.accumulated_time(int(s), 0.2)


def test_event_steps_follow_the_f32_clock(bs):
    ev = np.array([[0.0, 0, 0, 0, 1], [0.6, 0, 0, 0, 1], [0.61, 0, 0, 0, 1], [17.602346, 0, 0, 0, 1],
                   [59.37, 0, 0, 0, 1]], dtype=np.float32)
    got = bs.event_steps(ev, 0.2)
    want = [S.event_step(t, 0.2) for t in ev[:, 0]]
    assert got.tolist() == want
    assert got[0] == 0 and got[1] == 3 and got[2] == 4
