"""Write tests/golden/configs_oracle.json: the oracle's results on the two
plain-solver workloads, for the GPU tests at those workloads.

* 64^3, one bubble at the centre cell (the --bubble-at-centre extension),
  eta 0.3, eps_ratio 0.3, seed factor 1.15, 300 steps;
* 128^3, 12 bubbles at the reference's seeded positions (seed 1), nucleation
  steps uniform in [0, 200] (time seed 0x5EED), default model, 1500 steps.

For each: the schedule, SHA-256 digests of the f32 and f64 oracle states
(field bytes then fluid bytes, x fastest; as they are, and with zero signs
canonicalised) at the listed steps, and the f64
oracle's diagnostic sums (long double, reduce.wgsl's per-cell terms) every
100 steps.  The oracle is the CPU restatement of the reference's WGSL solver
(oracle/oracle.cpp); the schedule comes from the library's host functions
(make_schedule, physics.rs:248-308), which the GPU test checks against the
Simulation's own.  Run from the repository root:

    python tests/golden/make_config_fixtures.py

About 20 minutes on 8 cores (the 128^3 runs dominate).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "bubble-simulation_amd"), ROOT]

import bubblesim as bs  # noqa: E402
from oracle import cpu as orc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs_oracle.json")


def digest(field, fluid, canonical=False):
    """SHA-256 of the field bytes then the fluid bytes; canonical: with every
    -0 written as +0 first (x + 0 maps -0 to +0 and leaves every other value
    alone), for comparisons that treat the two zeros alike."""
    h = hashlib.sha256()
    for a in (field, fluid):
        a = np.ascontiguousarray(a)
        h.update((a + a.dtype.type(0)).tobytes() if canonical else a.tobytes())
    return h.hexdigest()


def run(name, model, grid, sched, steps, digest_at):
    params = bs.sim_params_words(model, grid)
    out = {"grid": list(grid), "steps": steps,
           "model": {"eta": model.eta, "eps_ratio": model.eps_ratio, "seed_factor": model.seed_factor},
           "schedule": [[float(v) for v in row] for row in sched], "digest_at": digest_at}
    for prec, label in ((0, "f32"), (1, "f64")):
        o = orc.OracleSim(prec, params, sched)
        digests, diags = {}, []
        done = 0
        t0 = time.time()
        while done < steps:
            o.advance(100)
            done += 100
            if prec == 1:
                s = o.diagnostics_sums()
                diags.append({"step": done, "remaining": o.remaining(), "sums": [float(v) for v in s]})
            if done in digest_at:
                st = o.state()
                digests[str(done)] = {"raw": digest(*st), "canonical": digest(*st, canonical=True)}
            print(f"{name} {label}: step {done} ({time.time() - t0:.0f} s)", flush=True)
        out[label] = {"digests": digests}
        if prec == 1:
            out[label]["diagnostics"] = diags
        o.close()
    return out


def main():
    res = {"note": "oracle (oracle/oracle.cpp, the CPU restatement of the reference solver) on the plain-solver "
                   "workloads; digests are SHA-256 of field then fluid bytes; diagnostics are the f64 oracle's "
                   "long-double sums (e_total, broken, kinetic, E, max|v|) every 100 steps; written by "
                   "tests/golden/make_config_fixtures.py"}
    # 64^3 centre bubble
    m1 = bs.Model(eta=0.3, eps_ratio=0.3, seed_factor=1.15)
    s1 = np.array([(0.0, 32.0, 32.0, 32.0, m1.seed_radius())], dtype=np.float32)
    res["centre_bubble_64"] = run("64^3", m1, (64, 64, 64), s1, 300, [100, 200, 300])
    # 128^3, 12 bubbles, uniform nucleation steps
    m2 = bs.Model()
    s2 = bs.make_schedule(bs.NUC_UNIFORM_STEPS, 12, m2.seed_radius(), 128, 60.0, 1, 200, 0x5EED, m2.dt())
    res["twelve_bubbles_128"] = run("128^3", m2, (128, 128, 128), s2, 1500, [100, 500, 1500])
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
