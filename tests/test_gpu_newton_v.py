"""The Newton-V form (dfn_set_formulation), an extension with no reference
counterpart: slip and log state are integrated and every evaluation solves the
stress balance for the slip rate.  It integrates a different ODE for the same
physics, so it is compared with the reference form at trajectory tolerance:
switching forms round-trips the state, and from a common start both forms are
run to a common final time (the last step landing on it) at the same tolerance
together with the reference form at a 1000x tighter tolerance as the accurate
solution; the Newton-V solution must be as close to it as the reference form at
the same tolerance is (within a factor 10 plus a floor)."""
import dataclasses

import numpy as np
import pytest

import cases
from demystification_dfn_amd import _abi, problems

pytestmark = pytest.mark.gpu


def _real(p):
    real = np.ones(p.N, bool)
    real[p.fake_indices()] = False
    return real


def _errors(p, a, b):
    """slip rate relative to max |V|, log state absolute, sigma_N relative."""
    N, real = p.N, _real(p)
    ev = np.max(np.abs(a[:N][real] - b[:N][real])) / np.max(np.abs(b[:N][real]))
    ep = np.max(np.abs(a[N:2 * N][real] - b[N:2 * N][real]))
    es = np.max(np.abs(a[2 * N:][real] - b[2 * N:][real]) / np.abs(b[2 * N:][real]))
    return np.array([ev, ep, es])


@pytest.mark.parametrize("make", [cases.single_fault_small, cases.crossing_network])
def test_switching_forms_round_trips_the_state(make):
    from demystification_dfn_amd.device import DeviceSimulation
    p = make()
    print(f"[gpu] newton-v switch {p.name} N={p.N}", flush=True)
    d = DeviceSimulation(p)
    d.solve(5)
    Y0, P0, _, t0 = d.get_state()
    d.set_formulation(_abi.FORM_NEWTON_V)
    Y1, P1, _, t1 = d.get_state()
    assert t1 == t0
    np.testing.assert_array_equal(P1, P0)
    real = _real(p)
    N = p.N
    assert np.all(np.abs(Y1[:N][real] - Y0[:N][real]) <= 1e-12 * np.abs(Y0[:N][real]))  # V from the balance
    np.testing.assert_array_equal(Y1[N:], Y0[N:])  # psi, sigma_N
    with pytest.raises(_abi.DfnError):
        d.monitor(True)  # the event monitor reads V from the state: reference form only
    d.set_formulation(_abi.FORM_REFERENCE)
    Y2 = d.get_state()[0]
    np.testing.assert_array_equal(Y2, Y1)


def _strengthening_crossing_network():
    """The crossing network with b = a / 2 everywhere (rate strengthening), so the
    pulse and the injection drive slow slip and no event within the window."""
    p = cases.crossing_network()
    p.arrays["b"] = 0.5 * p.arrays["a"]
    return p


def _solve_to(d, t_end, max_steps=60_000, chunk=5_000):
    """Steps to t_end in chunks (a progress line per chunk); fails instead of
    running on if t_end is not reached within max_steps."""
    s = None
    for _ in range(max_steps // chunk):
        s = d.solve(chunk, t_final=t_end)
        if s["time"] >= t_end:
            return s
        print(f"[gpu]   ... t = {s['time']:.4e} after {s['accepted_steps']} steps", flush=True)
    raise AssertionError(f"t_end {t_end} not reached in {max_steps} steps")


@pytest.mark.parametrize("make,t_end", [(cases.single_fault_small, 2.0e6), (_strengthening_crossing_network, 5.0e3)])
def test_newton_v_trajectory_tracks_the_reference_form(make, t_end):
    from demystification_dfn_amd.device import DeviceSimulation
    base = make()
    tol = 1e-7
    runs = {}
    for key, form, tl in (("fine", _abi.FORM_REFERENCE, tol * 1e-3), ("ref", _abi.FORM_REFERENCE, tol),
                          ("nv", _abi.FORM_NEWTON_V, tol)):
        p = dataclasses.replace(base, tol_solver=tl)
        print(f"[gpu] {base.name} {key} (tol {tl:g})", flush=True)
        d = DeviceSimulation(p)
        d.set_formulation(form)
        s = _solve_to(d, t_end)
        Y, P, _, t = d.get_state()
        assert t == t_end, (key, t)
        runs[key] = Y
        print(f"[gpu] {base.name} {key}: {s['accepted_steps']} steps, {s['rejected_attempts']} rejected, "
              f"max V {s['max_V']:.3e}", flush=True)
    e_ref = _errors(base, runs["ref"], runs["fine"])
    e_nv = _errors(base, runs["nv"], runs["fine"])
    print(f"[gpu] {base.name} errors vs 1000x tighter reference form: reference form {e_ref}, Newton-V {e_nv}",
          flush=True)
    assert np.all(e_nv <= 10 * e_ref + 1e-7), (e_nv, e_ref)


def test_newton_v_through_an_earthquake():
    """Two velocity-weakening faults under loading through their first events
    (the event-monitor case): both forms at the same tolerance, stepped one
    accepted step at a time to a common final time; the first event (onset and
    end at max V = 1e-3 m/s per fault) agrees to 1e-3 relative."""
    from demystification_dfn_amd.catalog import event_catalog
    from demystification_dfn_amd.device import DeviceSimulation
    p = problems.loaded_faults()
    ranges, k = [], 0
    for n in p.fault_nb_element:
        ranges.append((k, k + int(n)))
        k += int(n) + 1
    cats, stats, t_end = {}, {}, None
    for key, form in (("ref", _abi.FORM_REFERENCE), ("nv", _abi.FORM_NEWTON_V)):
        d = DeviceSimulation(p)
        d.set_formulation(form)
        times, maxv = [], []
        print(f"[gpu] loaded_faults {key}", flush=True)
        for i in range(20_000):
            if i % 500 == 499:
                print(f"[gpu]   ... {i + 1} steps, t = {times[-1]:.4e}", flush=True)
            if key == "ref":
                if len(times) == 2500:
                    break
                d.solve(1)
            else:
                if times and times[-1] >= t_end:
                    break
                d.solve(1, t_final=t_end)
            Y, _, _, t = d.get_state()
            times.append(t)
            maxv.append([np.max(np.abs(Y[a:b])) for a, b in ranges])
        if key == "ref":
            t_end = times[-1]
        stats[key] = d.stats()
        cats[key] = event_catalog(np.array(times), np.array(maxv), 1e-3, first_fault=1)
        print(f"[gpu] loaded_faults {key}: {stats[key]['accepted_steps']} steps, "
              f"{stats[key]['rejected_attempts']} rejected, events {cats[key]}", flush=True)
    assert times[-1] == t_end
    # the first event (a later one may straddle t_end in one form and not the other)
    er, en = cats["ref"][0], cats["nv"][0]
    assert er[0] == en[0]
    assert abs(er[3] - en[3]) <= 1e-3 * er[3] and abs(er[4] - en[4]) <= 1e-3 * er[4], (er, en)
