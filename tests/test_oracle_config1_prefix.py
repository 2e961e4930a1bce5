"""The oracle on config 1 against the reference's own Fortran over its first
checkpoint (tests/golden/config1_prefix.npz, written by make_config1_prefix.py
from oracle/_ref: 250 accepted steps of src/solver.f90:59-198 on the 4000-element
fault with injection, t = 3.03e6 s).  The restatement keeps the reference's
operation order, so the state, the pressure and the clock are bit-identical."""
import os

import numpy as np

import cases
from oracle.oracle import OracleSimulation

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config1_prefix.npz")


def test_oracle_equals_reference_over_first_config1_checkpoint():
    z = np.load(FIX, allow_pickle=False)
    p = cases.config1_single_fault()
    o = OracleSimulation(p, threads=min(8, os.cpu_count() or 1))
    o.solve(int(z["step"][0]))
    Y, P, _, t = o.get_state()
    assert t == float(z["time"][0])
    np.testing.assert_array_equal(Y, z["Y"][0])
    np.testing.assert_array_equal(P, z["P"][0])
