"""Writes tests/golden/config1_year.npz: the oracle (oracle/dfn_oracle.c, the C
restatement of the reference's solver loop) on the single-fault workload of
config 1 (4000 elements, Mode III, dense kernel, tol 1e-8, injection at the centre
for 100 days) stepped with the loop of src/solver.f90:59-198 to t = 1 simulated
year exactly (the last step lands on it).  Saved every 2500 accepted steps and at
the end: the step count, the time, the state [V | theta | sigmaN], P, and the slip
of every element (the trapezoid rule over the accepted steps, cumtrapz of V,
wrapper/network.m:474-476).

The reference's own Fortran (oracle/_ref) takes about 10 hours for the year, so
tests/golden/config1_prefix.npz holds its first 2750 steps; the oracle keeps the
reference's operation order and must sit on every checkpoint of that prefix bit
for bit, which this script asserts as it passes them, before it carries the same
arithmetic on to the end of the year (6229 steps, about 10 min on 6 threads).
The file is rewritten at every checkpoint, so a run stopped early leaves a complete
prefix.

With --reverse the script adds, at the same checkpoints, the trajectory of the
oracle with the dot products of its pressure solves summed from the last element
to the first (Y_rev, P_rev, slip_rev, time_rev): the same algorithm with another
rounding, as valid as the reference's.  How far that trajectory strays from the
reference-order one is the problem's own sensitivity to the summation order of
the pressure solves, which a device's other order meets too.

    python tests/golden/make_config1_year.py [--threads 8] [--reverse]
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
EVERY = 2500
OUT = os.path.join(HERE, "config1_year.npz")
PREFIX = os.path.join(HERE, "config1_prefix.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--reverse", action="store_true")
    a = ap.parse_args()
    p = cases.config1_single_fault()
    o = O.OracleSimulation(p, threads=a.threads)
    if a.reverse:
        return add_reverse(p, o)
    ref = np.load(PREFIX, allow_pickle=False)
    ref_at = {int(s): k for k, s in enumerate(ref["step"])}
    N = p.N
    steps, ck = 0, dict(step=[], time=[], Y=[], P=[], slip=[])
    Y, _, _, t = o.get_state()
    slip = np.zeros(N)
    w0 = time.perf_counter()
    while t < YEAR:
        o.solve(1, YEAR)
        Y1, P1, _, t1 = o.get_state()
        if t1 <= t:  # no step taken (the loop stopped on an error)
            break
        slip += (Y[:N] + Y1[:N]) * (t1 - t) / 2.0
        Y, t = Y1, t1
        steps += 1
        if steps in ref_at:  # the reference's own checkpoint: bit for bit
            k = ref_at[steps]
            assert t == float(ref["time"][k]), (steps, t, float(ref["time"][k]))
            np.testing.assert_array_equal(Y, ref["Y"][k])
            np.testing.assert_array_equal(P1, ref["P"][k])
            np.testing.assert_array_equal(slip, ref["slip"][k])
        if steps % EVERY == 0 or t >= YEAR:
            for key, v in (("step", steps), ("time", t), ("Y", Y), ("P", P1), ("slip", slip.copy())):
                ck[key].append(v)
            tmp = OUT + ".tmp.npz"
            np.savez_compressed(tmp, **{key: np.array(v) for key, v in ck.items()})
            os.replace(tmp, OUT)
            print(f"{steps} steps, t = {t:.6e} s ({time.perf_counter() - w0:.0f} s)", flush=True)


def add_reverse(p, o):
    z = dict(np.load(OUT, allow_pickle=False))
    o.set_cg_variant(0, True)
    N = p.N
    steps, t = 0, 0.0
    Y, _, _, t = o.get_state()
    slip = np.zeros(N)
    out = dict(time_rev=[], Y_rev=[], P_rev=[], slip_rev=[])
    w0 = time.perf_counter()
    for n in z["step"]:
        while steps < int(n):
            o.solve(1, YEAR)
            Y1, P1, _, t1 = o.get_state()
            slip += (Y[:N] + Y1[:N]) * (t1 - t) / 2.0
            Y, t = Y1, t1
            steps += 1
        for key, v in (("time_rev", t), ("Y_rev", Y), ("P_rev", P1), ("slip_rev", slip.copy())):
            out[key].append(v)
        print(f"reverse: {steps} steps, t = {t:.6e} s ({time.perf_counter() - w0:.0f} s)", flush=True)
    z.update({key: np.array(v) for key, v in out.items()})
    tmp = OUT + ".tmp.npz"
    np.savez_compressed(tmp, **z)
    os.replace(tmp, OUT)


if __name__ == "__main__":
    main()
