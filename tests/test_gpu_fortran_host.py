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


def test_fortran_host_output_files(tmp_path):
    """The loop in one dfn_solve call from Fortran with the library writing the
    output files (dfn_set_output / dfn_flush_output through dfn_hip): the files
    equal those of the same run driven from Python."""
    from demystification_dfn_amd.config_io import OUTPUT_FIELDS, read_output
    from demystification_dfn_amd.device import DeviceSimulation
    assert os.path.exists(DRIVER), "Fortran driver not built (__graft_entry__.build())"
    p = cases.crossing_network()
    nsteps, stride, freq = 9, 2, 2
    fin, fout = str(tmp_path / "prob.bin"), str(tmp_path / "out.bin")
    fdir, pdir = tmp_path / "fortran", tmp_path / "python"
    fdir.mkdir()
    pdir.mkdir()
    fortran_io.write_problem(p, fin, nsteps)
    print(f"[gpu] fortran host output files {p.name} N={p.N}", flush=True)
    subprocess.run([DRIVER, fin, fout, str(fdir), str(stride), str(freq)], check=True, timeout=120)
    d = DeviceSimulation(p)
    d.set_output(str(pdir), stride, freq)
    d.solve(nsteps)
    d.flush_output()
    names = sorted(os.listdir(pdir))
    assert names == ["output0000003", "output0000007", "output0000009"]
    assert sorted(os.listdir(fdir)) == names
    for nm in names:
        a, b = read_output(str(fdir / nm)), read_output(str(pdir / nm))
        for k in OUTPUT_FIELDS:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{nm}/{k}")
