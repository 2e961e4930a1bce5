"""Solver output files (dfn_set_output / dfn_flush_output) against the oracle.

The device run saves every stride_time-th accepted step and writes a file each
freq_writing_file saves, named by the step count, plus the final partial file
(src/solver.f90:180-202, 211-317).  The oracle is stepped one accepted step at a
time and its state read at the same steps; the files are read back with the
library's HDF5 reader.  Bars as for trajectories: V, theta, sigma_N and the shear
stress rate 1e-8 relative, P to the CG tolerance, time 1e-8 relative."""
import os

import numpy as np
import pytest

import cases
from demystification_dfn_amd import _abi
from oracle.oracle import OracleSimulation, stress_rate

pytestmark = pytest.mark.gpu


def close(a, b, rtol, floor):
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), floor))


@pytest.mark.parametrize("make", [cases.single_fault_small, cases.crossing_network])
def test_output_files_follow_the_oracle(make, tmp_path):
    from demystification_dfn_amd.config_io import read_output
    from demystification_dfn_amd.device import DeviceSimulation
    p = make()
    print(f"[gpu] output files, {p.name} N={p.N}", flush=True)
    d, o = DeviceSimulation(p), OracleSimulation(p)
    stride, freq, nsteps = 2, 3, 10
    d.set_output(str(tmp_path), stride, freq)
    d.solve(4)
    d.solve(nsteps - 4)  # the saves continue across solve calls
    d.flush_output()
    d.flush_output()  # nothing held: no second file
    assert sorted(os.listdir(tmp_path)) == ["output0000005", "output0000010"]
    f1, f2 = read_output(str(tmp_path / "output0000005")), read_output(str(tmp_path / "output0000010"))
    N = p.N
    assert f1["V"].shape == (N, freq) and f1["time"].shape == (freq,)
    assert f2["V"].shape == (N, 2) and f2["time"].shape == (2,)
    got = {k: np.concatenate([f1[k], f2[k]], axis=-1) for k in f1}
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    A = p.arrays
    col = 0
    for it in range(1, nsteps + 1):
        o.solve(1)
        if (it - 1) % stride:
            continue
        Y, P, _, t = o.get_state()
        sT = stress_rate(o)
        theta = A["Dc"] / A["V0"] * np.exp(Y[N:2 * N])
        assert abs(got["time"][col] - t) <= 1e-8 * t, it
        for key, ref in (("V", Y[:N]), ("theta", theta), ("tractionNel", Y[2 * N:])):
            assert close(got[key][real, col], ref[real], 1e-8, 1e-300), (it, key)
        scale = np.abs(sT[real]).max()
        assert np.max(np.abs(got["tractionTel"][real, col] - sT[real])) <= 1e-8 * scale, it
        assert np.max(np.abs(got["P"][:, col] - P)) <= 1e-9 * max(np.abs(P).max(), 1.0) + 10 * p.tol_solver, it
        np.testing.assert_array_equal(got["permeability"][real, col], A["permeability"][real])
        col += 1
    assert col == 5


def test_output_argument_checks(tmp_path):
    from demystification_dfn_amd.device import DeviceSimulation
    d = DeviceSimulation(cases.single_fault_small())
    with pytest.raises(_abi.DfnError):
        d.set_output(str(tmp_path), 0, 3)
    with pytest.raises(_abi.DfnError):
        d.set_output(str(tmp_path), 1, 0)
    d.set_output(None, 0, 0)  # off
    d.solve(2)
    d.flush_output()
    assert os.listdir(tmp_path) == []
