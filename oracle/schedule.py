"""Python restatement of the reference's nucleation schedule (test
infrastructure only).

Follows ``make_schedule`` / ``periodic_distance`` (src/physics.rs:248-308)
and the third-party generator stack it draws from, pinned in Cargo.lock:

* rand_core 0.9.5 ``SeedableRng::seed_from_u64``: four PCG32 (XSH-RR 64/32)
  outputs, the 64-bit LCG (multiplier 6364136223846793005, increment
  11634580027462260723) advanced before each output, laid little-endian into
  the 16-byte seed;
* rand_pcg 0.9.0 ``Mcg128Xsl64`` (= ``Pcg64Mcg``): ``from_seed`` reads the
  seed as a little-endian u128 and forces the low bit; ``next_u64``
  multiplies the state by 0x2360ED051FC65DA44385DF649FCCF645 and applies the
  XSL-RR output (rotate-right of hi^lo by state>>122); ``next_u32`` is the low
  half of ``next_u64``;
* rand 0.9.5 ``random::<f32>()``: ``(next_u32() >> 8) as f32 * 2^-24``.

Every f32 operation is carried out on numpy float32 scalars (IEEE single, no
contraction); ``exp``/``ln`` call the platform libm's ``expf``/``logf`` as
Rust's std does on Linux.  The generator is written with Python integers, an
implementation independent of the library's C++ one.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

F = np.float32
M64 = (1 << 64) - 1
M128 = (1 << 128) - 1
MCG_MUL = 0x2360ED051FC65DA44385DF649FCCF645

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]
_libm.expf.restype = ctypes.c_float
_libm.expf.argtypes = [ctypes.c_float]


def logf(x) -> np.float32:
    return F(_libm.logf(float(x)))


def expf(x) -> np.float32:
    return F(_libm.expf(float(x)))


def _rotr64(x: int, r: int) -> int:
    r &= 63
    return ((x >> r) | (x << (64 - r))) & M64 if r else x


def _rotr32(x: int, r: int) -> int:
    r &= 31
    return ((x >> r) | (x << (32 - r))) & 0xFFFFFFFF if r else x


def seed_from_u64(state: int) -> int:
    seed = bytearray()
    for _ in range(4):
        state = (state * 6364136223846793005 + 11634580027462260723) & M64
        xorshifted = (((state >> 18) ^ state) >> 27) & 0xFFFFFFFF
        rot = state >> 59
        seed += _rotr32(xorshifted, rot).to_bytes(4, "little")
    return int.from_bytes(bytes(seed), "little")


class Mcg128Xsl64:
    def __init__(self, state: int):
        self.state = (state | 1) & M128

    @classmethod
    def seed_from_u64(cls, s: int) -> "Mcg128Xsl64":
        return cls(seed_from_u64(s))

    def next_u64(self) -> int:
        self.state = (self.state * MCG_MUL) & M128
        rot = self.state >> 122
        xsl = ((self.state >> 64) ^ self.state) & M64
        return _rotr64(xsl, rot)

    def next_u32(self) -> int:
        return self.next_u64() & 0xFFFFFFFF

    def random_f32(self) -> np.float32:
        return F(1.0 / 16777216.0) * F(self.next_u32() >> 8)


def periodic_distance(a, b, n) -> np.float32:
    s = F(0.0)
    for i in range(3):
        d = F(abs(F(a[i]) - F(b[i])))
        if d > F(0.5) * n[i]:
            d = n[i] - d
        s = F(s + d * d)
    return F(np.sqrt(s))


def accumulated_time(k: int, dt) -> np.float32:
    t = F(0.0)
    for _ in range(k):
        t = F(t + F(dt))
    return t


def make_schedule(mode: str, count: int, radius, grid, duration, seed: int, uniform_max_step: int = 200,
                  time_seed: int = 0x5EED, dt=0.2) -> np.ndarray:
    """Rows (time, x, y, z, radius) as float32; mode 'simultaneous' |
    'exponential' (physics.rs:263-272) | 'uniform-steps' (extension: times at
    whole steps k in [0, uniform_max_step], k = next_u64 % (max+1) from a
    separate generator seeded with time_seed)."""
    rng = Mcg128Xsl64.seed_from_u64(seed)
    trng = Mcg128Xsl64.seed_from_u64(time_seed)
    n = [F(grid[0]), F(grid[1]), F(grid[2])]
    radius = F(radius)
    min_sep = F(2.5) * radius
    events = []
    for k in range(count):
        if mode == "simultaneous":
            t = F(0.0)
        elif mode == "exponential":
            growth = F(4.0)
            frac = F(F(k) + F(0.5)) / F(count)
            t = F(F(duration) / growth) * logf(F(1.0) + frac * F(expf(growth) - F(1.0)))
        elif mode == "uniform-steps":
            t = accumulated_time(trng.next_u64() % (uniform_max_step + 1), dt)
        else:
            raise ValueError(mode)
        pos = (F(0), F(0), F(0))
        for _attempt in range(64):
            cand = (rng.random_f32() * n[0], rng.random_f32() * n[1], rng.random_f32() * n[2])
            ok = all(periodic_distance(cand, e[1:4], n) > min_sep for e in events)
            pos = cand
            if ok:
                break
        events.append((F(t), pos[0], pos[1], pos[2], radius))
    # stable sort by time
    events = [e for _, e in sorted(enumerate(events), key=lambda ie: (ie[1][0], ie[0]))]
    return np.array(events, dtype=np.float32).reshape(-1, 5)


def event_step(t, dt) -> int:
    """First step s whose f32 clock (sum of s dt's) is >= t (sim.rs:843-879)."""
    acc = F(0.0)
    s = 0
    t = F(t)
    while not (t <= acc):
        nxt = F(acc + F(dt))
        if not (nxt > acc):
            return -1
        acc = nxt
        s += 1
    return s
