"""Error conventions of the C ABI against the oracle, which implements the same
ABI: every entry point returns a status and never aborts, where the reference
STOPs (src/solver.f90:412-420 for an extensional normal stress) or would read
unset module state (a kernel call before initiate_hmat,
src/kernel_calculation/hmatrixModeII.f90:51-62).  Both libraries must give the
same status, and after a stop the same state."""
import ctypes as C

import numpy as np
import pytest

import cases
from demystification_dfn_amd import _abi, problems
from oracle.oracle import OracleSimulation

pytestmark = pytest.mark.gpu


def _both(p):
    from demystification_dfn_amd.device import DeviceSimulation
    print(f"[gpu] errors {p.name} N={p.N}", flush=True)
    return DeviceSimulation(p), OracleSimulation(p)


def _status(fn, *args):
    try:
        fn(*args)
    except _abi.DfnError as e:
        return e.code
    return _abi.DFN_OK


def test_extensional_normal_stress_stops_both_after_the_accepted_step():
    """sigma_N > 0 on one element: the solver stops with DFN_ERR_TENSILE after the
    step that the reference checks (check_errors on the accepted state), and both
    libraries hold the same state at the stop."""
    p = cases.crossing_network()
    d, o = _both(p)
    Y, P, _, t = o.get_state()
    N = p.N
    real = np.flatnonzero(~np.isin(np.arange(N), p.fake_indices()))
    Y = Y.copy()
    Y[2 * N + real[len(real) // 2]] = 1.0e3  # extensional by 1 kPa
    for s in (d, o):
        s.set_state(Y, P, t)
    codes = [_status(s.solve, 5) for s in (d, o)]
    assert codes == [_abi.DFN_ERR_TENSILE, _abi.DFN_ERR_TENSILE], codes
    sd, so = d.stats(), o.stats()
    assert sd["accepted_steps"] == so["accepted_steps"] == 1
    Yd, Pd, _, td = d.get_state()
    Yo, Po, _, to = o.get_state()
    assert abs(td - to) <= 1e-12 * to
    np.testing.assert_allclose(Yd[real], Yo[real], rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose(Yd[2 * N:][real], Yo[2 * N:][real], rtol=1e-8)


def test_hmatrix_kernel_before_initiate_hmat_is_a_state_error():
    p = problems.network_small(n_faults=6, seed=3, mode=2, eta=2.0, eps_aca=1e-6, nleaf=16,
                               static_kernel=_abi.KERNEL_HMATRIX)
    d, o = _both(p)
    x = np.ones(p.N)
    for s in (d, o):
        assert _status(s.kernel, x) == _abi.DFN_ERR_STATE
        assert _status(s.Hgemv_right, x) == _abi.DFN_ERR_STATE
    assert _status(d.tuning) == _abi.DFN_ERR_STATE
    for s in (d, o):
        s.initiate_hmat()
        assert _status(s.kernel, x) == _abi.DFN_OK
        assert _status(s.Hgemv_right, x, 2) == _abi.DFN_ERR_ARG  # operator index out of range


def test_null_pointers_and_bad_options_are_argument_errors():
    p = cases.single_fault_small()
    d, o = _both(p)
    for s in (d, o):
        f, h = s._fns, s._h
        null = C.POINTER(C.c_double)()
        assert f["kernel"](h, null, null, null) == _abi.DFN_ERR_ARG
        assert f["friction_law"](h, 0.0, null, null) == _abi.DFN_ERR_ARG
        assert f["rkqs"](h, null, null, None, 1.0, 1e-6, null, None, None) == _abi.DFN_ERR_ARG
        assert f["kernel"](None, null, null, null) == _abi.DFN_ERR_ARG
    assert d._fns["set_precision"](d._h, 7) == _abi.DFN_ERR_ARG  # the oracle stores FP64 only
    # the handle is still usable after the refused calls, and both agree on it
    sd, so = d.solve(3), o.solve(3)
    assert sd["accepted_steps"] == so["accepted_steps"] == 3
    np.testing.assert_allclose(d.get_state()[0], o.get_state()[0], rtol=1e-8, atol=1e-300)


@pytest.mark.parametrize("late_part", [1, 2, 3])
def test_cooperative_compute_P_reports_any_workgroup_that_lost_contact(late_part, monkeypatch):
    """The multi-workgroup compute_P of a 4855-element network spans three
    workgroups per system.  A host switch (DFN_CG_LATE_PART) makes one of them
    report a message wait that gave up, while every kernel still runs to its end:
    whichever workgroup it is, the solve reports it (DFN_ERR_CG_TIMEOUT), instead
    of a converged pressure or a step halving; the same network without the
    switch solves."""
    from demystification_dfn_amd.device import DeviceSimulation
    monkeypatch.setenv("DFN_CG_PATH", "coop")
    p = cases.connected_network_spans()
    print(f"[gpu] errors coop timeout part {late_part} N={p.N}", flush=True)
    monkeypatch.setenv("DFN_CG_LATE_PART", str(late_part))
    d = DeviceSimulation(p)
    assert _status(d.compute_P, 5000.0) == _abi.DFN_ERR_CG_TIMEOUT
    assert _status(d.solve, 1) == _abi.DFN_ERR_CG_TIMEOUT
    assert d.stats()["accepted_steps"] == 0 and d.stats()["cg_halvings"] == 0
    d.close()
    monkeypatch.setenv("DFN_CG_LATE_PART", "0")
    d = DeviceSimulation(p)
    Pt, _, ier = d.compute_P(5000.0)
    assert ier == 0 and np.isfinite(Pt).all()
    assert d.solve(1)["accepted_steps"] == 1
