"""Config 1 over a long horizon against the reference's own Fortran.

tests/golden/config1_prefix.npz holds checkpoints, every 250 accepted steps, of the
reference's solver loop (src/solver.f90:59-198, built from its sources into
oracle/_ref) on the config-1 fault (4000 elements, Mode III, dense kernel, tol
1e-8, injection for the first 100 days), well past the end of the injection.  The
device runs the same problem through dfn_solve from the same initial state and
must sit on every checkpoint with the same accepted-step count.

Its clock drifts from the reference's by about 1e-9 relative per thousand steps:
the embedded error estimate is a difference of nearly equal stage sums, so the
device's other summation order moves each proposed step in its ninth digit (no
accept/reject decision differs).  So at every checkpoint: the accepted-step
counts are equal, the clocks agree to 1e-8 relative, sigmaN to 1e-12, and V,
theta and the slip (trapezoid rule over the accepted steps,
wrapper/network.m:474-476) agree to 1e-8 relative beyond what the device state
moves over twice the clock difference (|dY/dt| 2 |t_dev - t_ref|, from the
device's own right-hand side), and P, beyond |dP/dt| 2 |t_dev - t_ref|, to 1e-8
of its largest value (the level the pressure solves stop at, tol_solver = 1e-8).
The largest raw differences and the clock offsets are reported.
"""
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config1_prefix.npz")


def _rel_max(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_config1_long_prefix_against_reference_fortran():
    from demystification_dfn_amd.device import DeviceSimulation
    z = np.load(FIX, allow_pickle=False)
    p = cases.config1_single_fault()
    N = p.N
    print(f"[gpu] config 1 prefix: N={N}, {len(z['step'])} checkpoints to step {int(z['step'][-1])}, "
          f"t = {float(z['time'][-1]):.6e} s", flush=True)
    d = DeviceSimulation(p)
    d.monitor(True)
    done, first_miss = 0, None
    worst = dict(time=0.0, V=0.0, theta=0.0, sigmaN=0.0, slip=0.0, P=0.0, V_beyond=0.0)
    for k, n in enumerate(z["step"]):
        st = d.solve(int(n) - done)
        done = int(n)
        assert st["accepted_steps"] == done
        d.monitor_take()  # the per-step maxima are not compared; keep the record short
        Y, P, Pdot, t = d.get_state()
        _, slip = d.moment()
        Yr, Pr, tr, sr = z["Y"][k], z["P"][k], float(z["time"][k]), z["slip"][k]
        band = 2.0 * abs(t - tr)
        f = d.friction_law(1.0, Y)  # dY/dt at the device state (pressure rate over the next second)
        e = dict(time=abs(t - tr) / tr, V=_rel_max(Y[:N], Yr[:N]), theta=_rel_max(Y[N:2 * N], Yr[N:2 * N]),
                 sigmaN=_rel_max(Y[2 * N:], Yr[2 * N:]), slip=_rel_max(slip, sr),
                 P=float(np.max(np.abs(P - Pr))) / max(float(np.abs(Pr).max()), 1.0))
        # differences beyond the clock band, relative (P: to its largest value)
        beyond = np.abs(Y - Yr) - band * np.abs(f)
        e["V_beyond"] = float(np.max(beyond[:N] / np.abs(Yr[:N])))
        e["theta_beyond"] = float(np.max(beyond[N:2 * N] / np.abs(Yr[N:2 * N])))
        e["slip_beyond"] = float(np.max((np.abs(slip - sr) - band * np.abs(Y[:N])) / np.abs(sr)))
        e["P_beyond"] = float(np.max(np.abs(P - Pr) - band * np.abs(Pdot))) / max(float(np.abs(Pr).max()), 1.0)
        for key, v in e.items():
            worst[key] = max(worst.get(key, 0.0), v)
        ok = (e["time"] <= 1e-8 and e["sigmaN"] <= 1e-12 and e["V_beyond"] <= 1e-8 and e["theta_beyond"] <= 1e-8
              and e["slip_beyond"] <= 1e-8 and e["P_beyond"] <= 1e-8)
        if not ok and first_miss is None:
            first_miss = (int(n), tr, e)
    print(f"[gpu] config 1 prefix over {done} steps, t = {t:.6e} s: largest relative differences {worst}; "
          f"first miss {first_miss}", flush=True)
    assert first_miss is None, first_miss
