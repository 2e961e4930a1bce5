"""Writes tests/golden/config1_prefix.npz: the reference's own Fortran (oracle/_ref,
built by oracle/Makefile.ref from /root/reference/src) on the single-fault workload
of config 1 (4000 elements, Mode III, dense kernel, tol 1e-8, injection at the
centre for 100 days) stepped with the solver loop of src/solver.f90:59-198 towards
t = 1 simulated year, for at most --max-steps accepted steps.  Saved every 250
accepted steps and at the end: the step count, the time, the state
[V | theta | sigmaN], P, and the slip of every element, the trapezoid rule over the
accepted steps (cumtrapz of V, wrapper/network.m:474-476).  The file is rewritten
at every checkpoint, so a run stopped early leaves a complete prefix.

Run where /root/reference exists; the reference is serial and built at -O1, about
0.7 s per step here once the injection has ended (steps of ~500 s), so the full
year (~50 000 steps) is out of reach and the fixture holds a prefix that runs well
past the end of the injection.

    python tests/golden/make_config1_prefix.py [--max-steps 6000]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import oracle as O  # noqa: E402
import cases  # noqa: E402

YEAR = 365.25 * 86400.0
EVERY = 250
OUT = os.path.join(HERE, "config1_prefix.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--max-steps", type=int, default=6000)
    a = ap.parse_args()
    assert O.RefSimulation.available(), "build oracle/_ref first (make -C oracle -f Makefile.ref)"
    p = cases.config1_single_fault()
    r = O.RefSimulation(p)
    N = p.N
    steps, ck = 0, dict(step=[], time=[], Y=[], P=[], slip=[])
    Y, _, _, t = r.get_state()
    slip = np.zeros(N)
    w0 = time.perf_counter()
    while t < YEAR and steps < a.max_steps:
        if r.solve(1, YEAR) == 0:
            break
        Y1, P1, _, t1 = r.get_state()
        slip += (Y[:N] + Y1[:N]) * (t1 - t) / 2.0
        Y, t = Y1, t1
        steps += 1
        if steps % EVERY == 0 or t >= YEAR or steps == a.max_steps:
            for k, v in (("step", steps), ("time", t), ("Y", Y), ("P", P1), ("slip", slip.copy())):
                ck[k].append(v)
            tmp = OUT + ".tmp.npz"
            np.savez_compressed(tmp, **{k: np.array(v) for k, v in ck.items()})
            os.replace(tmp, OUT)
            print(f"{steps} steps, t = {t:.6e} s ({time.perf_counter() - w0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
