"""The CPU oracle against fixtures written by the reference's own Fortran modules
(tests/golden/make_golden.py).  The oracle restates the reference arithmetic in
the same operation order, so these comparisons are exact."""
import os

import numpy as np
import pytest

import cases
from oracle.oracle import OracleSimulation, RefSimulation, eval_matrix, geometry, hmat_tree


def _load(golden_dir, name):
    return dict(np.load(os.path.join(golden_dir, name)))


@pytest.mark.parametrize("fixture,make", [("single_fault_300.npz", cases.single_fault_small),
                                          ("crossing_network.npz", cases.crossing_network),
                                          ("network_rate_strengthening.npz", cases.reference_network),
                                          ("coupling_network.npz", cases.coupling_network)])
def test_oracle_matches_reference_fixtures(golden_dir, fixture, make):
    g = _load(golden_dir, fixture)
    prob = make()
    o = OracleSimulation(prob)
    sT, sN = o.kernel(g["kernel_in"])
    np.testing.assert_array_equal(sT, g["kernel_T"])
    if prob.fracture_mode == 2:
        np.testing.assert_array_equal(sN, g["kernel_N"])
    Y0 = prob.initial_Y()
    np.testing.assert_array_equal(Y0, g["Y0"])
    for x, ref in zip(g["rhs_x"], g["rhs"]):
        np.testing.assert_array_equal(o.friction_law(x, Y0), ref)
    for dt, pt, pd, ier in zip(g["P_dt"], g["P_temp"], g["P_dot"], g["P_ier"]):
        a, b, e = o.compute_P(dt)
        assert e == ier
        if not ier:
            np.testing.assert_array_equal(a, pt)
            np.testing.assert_array_equal(b, pd)
    y1, x1, hdid, hnext = o.rkqs(Y0, g["rkqs_dydx"], 0.0, 200.0, prob.tol_solver, g["rkqs_yscal"])
    np.testing.assert_array_equal(y1, g["rkqs_y"])
    assert (x1, hdid, hnext) == (float(g["rkqs_x"]), float(g["rkqs_hdid"]), float(g["rkqs_hnext"]))
    o2 = OracleSimulation(prob)
    o2.solve(int(g["solve_steps"]))
    Y, P, _, t = o2.get_state()
    np.testing.assert_array_equal(Y, g["solve_Y"])
    np.testing.assert_array_equal(P, g["solve_P"])
    assert t == float(g["solve_time"])


def test_oracle_trees_match_reference_fixture(golden_dir):
    g = _load(golden_dir, "trees_network20.npz")
    q = cases.tree_network()
    o = OracleSimulation(q)
    np.testing.assert_array_equal(geometry(o)["elem"], g["elem"])
    o.initiate_hmat()
    t = hmat_tree(o, 0)
    np.testing.assert_array_equal(t["mat_corr"], g["mat_corr"])
    np.testing.assert_array_equal(t["ind_clusters"], g["ind_clusters"])
    np.testing.assert_array_equal(t["bltree"], g["bltree"])
    split = t["bltree"][:, 0] == 2
    np.testing.assert_array_equal(t["info"][split, :4], g["info"][split, :4])


@pytest.mark.skipif(not RefSimulation.available(), reason="reference build oracle/_ref not present")
def test_oracle_matches_live_reference_network():
    """A seeded 12-fault network (no crossings), Mode II direct: kernels, RHS and 25 steps."""
    from demystification_dfn_amd import problems
    prob = problems.network_small(n_faults=12, seed=7, static_kernel=0)
    o, r = OracleSimulation(prob), RefSimulation(prob)
    V = np.random.default_rng(5).random(prob.N) * 1e-8
    for a, b in zip(o.kernel(V), r.kernel(V)):
        np.testing.assert_array_equal(a, b)
    Y0 = prob.initial_Y()
    np.testing.assert_array_equal(o.friction_law(500.0, Y0), r.friction_law(500.0, Y0))
    o.solve(25)
    r.solve(25)
    for a, b in zip(o.get_state(), r.get_state()):
        np.testing.assert_array_equal(a, b)


ENTRY_CASES = [("single_fault_300", cases.single_fault_small), ("crossing_network", cases.crossing_network),
               ("network20", cases.tree_network), ("network20_mode3", cases.mode3_network),
               ("network_rate_strengthening", cases.reference_network)]


@pytest.mark.parametrize("name,make", ENTRY_CASES)
def test_oracle_kernel_entries_match_reference_fixture(golden_dir, name, make):
    """The H-matrix entry functions eval_kernel_ModeIII / _ModeII_T / _ModeII_N
    (src/hmatrix/eval_kernel_hmatrix.f90:7-199) against the reference build's, bit
    for bit, at the diagonal, the neighbour diagonals, the fake elements' rows and
    columns and seeded pairs (faults of every strike, fake elements between them)."""
    z = np.load(os.path.join(golden_dir, "kernel_entries.npz"))
    prob = make()
    o = OracleSimulation(prob)
    k, l = z[f"{name}__k"].astype(np.int64), z[f"{name}__l"].astype(np.int64)
    for op in ((0,) if prob.fracture_mode == 3 else (0, 1)):
        np.testing.assert_array_equal(eval_matrix(o, op)[k, l], z[f"{name}__op{op}"])


@pytest.mark.skipif(not RefSimulation.available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("make", [cases.tree_network, cases.mode3_network])
def test_oracle_kernel_matrices_match_live_reference(make):
    """Whole kernel matrices of a 20-fault network against the reference build."""
    prob = make()
    o, r = OracleSimulation(prob), RefSimulation(prob)
    for op in ((0,) if prob.fracture_mode == 3 else (0, 1)):
        np.testing.assert_array_equal(eval_matrix(o, op), r.eval_matrix(op))
