"""The Fortran host (fortran/drive_solver.f90) runs the reference's solver-loop call
sequence against the device library through the ISO_C_BINDING layer
(fortran/dfn_hip.f90); its trajectory must match the CPU oracle's."""
import os
import subprocess

import numpy as np
import pytest

import cases
from demystification_dfn_amd import fortran_io
from oracle.oracle import OracleSimulation

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "demystification_dfn_amd", "fortran", "build", "drive_solver")


@pytest.mark.parametrize("make,nsteps", [(cases.single_fault_small, 15), (cases.crossing_network, 15)])
def test_fortran_host_trajectory(tmp_path, make, nsteps):
    assert os.path.exists(DRIVER), "Fortran driver not built (__graft_entry__.build())"
    p = make()
    fin, fout = str(tmp_path / "prob.bin"), str(tmp_path / "out.bin")
    fortran_io.write_problem(p, fin, nsteps)
    print(f"[gpu] fortran host {p.name} N={p.N} steps={nsteps}", flush=True)
    subprocess.run([DRIVER, fin, fout], check=True, timeout=120)
    t, Y, P = fortran_io.read_result(fout, p.N)
    o = OracleSimulation(p)
    o.solve(nsteps)
    Yo, Po, _, to = o.get_state()
    assert abs(t - to) <= 1e-8 * to
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for c in range(3):
        a, b = Y[c * N:(c + 1) * N][real], Yo[c * N:(c + 1) * N][real]
        assert np.all(np.abs(a - b) <= 1e-8 * np.abs(b)), c
