#!/usr/bin/env python3
"""Writes tests/golden/schedule_kat.json: nucleation schedules and raw
generator streams computed by the oracle's Python restatement of
make_schedule (oracle/schedule.py, physics.rs:248-308).  The library's C++
implementation is checked against these fixtures bit for bit."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import numpy as np  # noqa: E402

from oracle import schedule as S  # noqa: E402

R0 = np.float32(1.15) * np.float32(np.float32(2.0) * (np.float32(np.sqrt(np.float32(0.25))) / np.float32(6.0)) /
                                    np.float32(np.float32(np.float32(0.3) * np.float32(0.5)) / np.float32(12.0)))

CASES = [
    dict(mode="exponential", count=12, radius=float(R0), grid=[192, 192, 192], duration=60.0, seed=1),
    dict(mode="exponential", count=12, radius=8.0, grid=[64, 64, 64], duration=40.0, seed=7),
    dict(mode="exponential", count=12, radius=float(R0), grid=[128, 128, 128], duration=60.0, seed=1),
    dict(mode="simultaneous", count=1, radius=float(R0), grid=[128, 128, 128], duration=60.0, seed=1),
    dict(mode="simultaneous", count=32, radius=float(R0), grid=[512, 512, 512], duration=60.0, seed=1),
    dict(mode="simultaneous", count=40, radius=20.0, grid=[96, 80, 64], duration=60.0, seed=42),
    dict(mode="uniform-steps", count=32, radius=float(R0), grid=[512, 512, 512], duration=60.0, seed=1,
         uniform_max_step=200, time_seed=11, dt=0.2),
    dict(mode="uniform-steps", count=4, radius=6.0, grid=[48, 48, 48], duration=60.0, seed=3,
         uniform_max_step=200, time_seed=5, dt=0.2),
]


def main():
    out = {"provenance": "oracle/schedule.py (Python restatement of physics.rs:248-308 and the rand 0.9.5 / "
                         "rand_core 0.9.5 / rand_pcg 0.9.0 generator stack)",
           "cases": [], "streams": []}
    for c in CASES:
        ev = S.make_schedule(c["mode"], c["count"], c["radius"], c["grid"], c["duration"], c["seed"],
                             c.get("uniform_max_step", 200), c.get("time_seed", 0x5EED), c.get("dt", 0.2))
        out["cases"].append(dict(c, events=[[float(v) for v in row] for row in ev],
                                 events_hex=[[np.float32(v).view(np.uint32).item() for v in row] for row in ev]))
    for seed in (0, 1, 7, 42, 2**64 - 1):
        r = S.Mcg128Xsl64.seed_from_u64(seed)
        u64 = [r.next_u64() for _ in range(8)]
        r = S.Mcg128Xsl64.seed_from_u64(seed)
        f32 = [np.float32(r.random_f32()).view(np.uint32).item() for _ in range(8)]
        out["streams"].append(dict(seed=seed, u64=[str(v) for v in u64], f32_bits=f32))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedule_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
