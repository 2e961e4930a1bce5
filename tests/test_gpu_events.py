"""The device event monitor (dfn_monitor, dfn_monitor_take, dfn_moment) against
the oracle's trajectory post-processed as the reference does: per-fault max |V|
at every accepted step, slip = cumtrapz(t, V), M0 = sum(mu * slip * ds), and the
event catalog rule.  Two velocity-weakening faults under shear loading go
through their first earthquakes.  Bars: the recorded series agree to 1e-8 over
the interseismic steps; event onset and end times and the moments to 1e-4."""
import numpy as np
import pytest

from demystification_dfn_amd import problems
from demystification_dfn_amd.catalog import event_catalog, moment
from oracle.oracle import OracleSimulation, geometry

pytestmark = pytest.mark.gpu

STEPS = 2500


def _fault_ranges(p):
    out, k = [], 0
    for n in p.fault_nb_element:
        out.append((k, k + int(n)))
        k += int(n) + 1
    return out


def test_event_catalog_and_moment_match_oracle():
    from demystification_dfn_amd.device import DeviceSimulation
    p = problems.loaded_faults()
    print(f"[gpu] loaded_faults N={p.N}", flush=True)
    d = DeviceSimulation(p)
    d.monitor(True)
    d.solve(STEPS)
    t_d, v_d = d.monitor_take()
    M0_d, slip_d = d.moment()
    assert t_d.shape[0] == STEPS and v_d.shape == (STEPS, p.nb_fault)

    o = OracleSimulation(p, threads=8)
    ds = geometry(o)["ds"]
    N = p.N
    ranges = _fault_ranges(p)
    real = np.zeros(N, bool)
    for a, b in ranges:
        real[a:b] = True
    Vold, t_old = o.get_state()[0][:N].copy(), 0.0
    slip = np.zeros(N)
    t_o, v_o = [], []
    for _ in range(STEPS):
        o.solve(1)
        Y, _, _, t = o.get_state()
        V = Y[:N]
        slip[real] += 0.5 * (Vold[real] + V[real]) * (t - t_old)
        t_o.append(t)
        v_o.append([np.max(np.abs(V[a:b])) for a, b in ranges])
        Vold, t_old = V.copy(), t
    t_o, v_o = np.array(t_o), np.array(v_o)
    M0_o = np.array([moment(p.mu, slip[a:b], ds[a:b]) for a, b in ranges])

    early = 100
    assert np.max(np.abs(t_d[:early] - t_o[:early]) / t_o[:early]) <= 1e-10
    assert np.max(np.abs(v_d[:early] - v_o[:early]) / v_o[:early]) <= 1e-8
    cat_d = event_catalog(t_d, v_d, 1e-3, first_fault=1)
    cat_o = event_catalog(t_o, v_o, 1e-3, first_fault=1)
    print(f"[gpu] events device {cat_d}\n[gpu] events oracle {cat_o}", flush=True)
    assert len(cat_o) >= 1 and len(cat_d) == len(cat_o)
    for ed, eo in zip(cat_d, cat_o):
        assert ed[0] == eo[0]
        assert abs(ed[3] - eo[3]) <= 1e-4 * eo[3] and abs(ed[4] - eo[4]) <= 1e-4 * eo[4]
    print(f"[gpu] M0 device {M0_d} oracle {M0_o}", flush=True)
    assert np.all(np.abs(M0_d - M0_o) <= 1e-4 * np.abs(M0_o))
