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


ENS_DRIVER = os.path.join(ROOT, "demystification_dfn_amd", "fortran", "build", "drive_ensemble")


def _ens_nets():
    from demystification_dfn_amd import _abi, problems
    return [problems.network_small(n_faults=nf, seed=100 + i, mode=2, eta=2.0, eps_aca=1e-6, nleaf=16,
                                   static_kernel=_abi.KERNEL_HMATRIX, tol=tol)
            for i, (nf, tol) in enumerate(((9, 1e-6), (14, 1e-9), (11, 1e-7)))]


@pytest.mark.parametrize("learned", [False, True])
def test_fortran_host_ensemble(tmp_path, learned):
    """An ensemble driven from Fortran through the dfn_ensemble_* / dfn_mlp_*
    interfaces of dfn_hip (fortran/drive_ensemble.f90) equals the same calls made
    from Python, bit for bit: every network's state and step counts, the recorded
    (features, target) pairs and, with the learned controller (FP32 operators and
    the trained proposal network), the network's outputs on those pairs and one
    training epoch on them."""
    from demystification_dfn_amd import _abi
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceMlp
    from demystification_dfn_amd.train_step_network import load
    assert os.path.exists(ENS_DRIVER), "Fortran ensemble driver not built (__graft_entry__.build())"
    probs = _ens_nets()
    files = []
    for i, p in enumerate(probs):
        files.append(str(tmp_path / f"net{i}.bin"))
        fortran_io.write_problem(p, files[-1], 0)
    prec = _abi.PREC_FP32 if learned else _abi.PREC_FP64
    ctrl = _abi.STEP_LEARNED if learned else _abi.STEP_RKQS
    wfile = "-"
    if learned:
        params, norm, _ = load()
        wfile = str(tmp_path / "weights.bin")
        fortran_io.write_weights(wfile, params, norm)
    nsteps = 40
    fout = str(tmp_path / "ens.bin")
    print(f"[gpu] fortran ensemble N={[p.N for p in probs]} learned={learned}", flush=True)
    subprocess.run([ENS_DRIVER, fout, str(nsteps), str(prec), str(ctrl), wfile] + files, check=True, timeout=300)
    got = fortran_io.read_ensemble_result(fout, [p.N for p in probs], _abi.MLP_PARAMS)

    ens = DeviceEnsemble(probs)
    ens.initiate_hmat()
    ens.set_precision(prec)
    mlp = None
    if learned:
        mlp = DeviceMlp()
        mlp.set_params(params, norm)
    ens.set_controller(ctrl, mlp)
    ens.set_recording(True)
    ens.solve(nsteps)
    for i, g in enumerate(got["nets"]):
        Y, P, _, t = ens.get_state(i)
        s = ens.stats(i)
        assert g["time"] == t and g["accepted"] == s["accepted_steps"] == nsteps
        assert g["rejected"] == s["rejected_attempts"]
        np.testing.assert_array_equal(g["Y"], Y)
        np.testing.assert_array_equal(g["P"], P)
    X, T = ens.take_records()
    assert len(T) > 0
    np.testing.assert_array_equal(got["X"], X)
    np.testing.assert_array_equal(got["T"], T)
    if learned:
        np.testing.assert_array_equal(got["Y"], mlp.forward(X))
        loss = mlp.train(X, T, epochs=1, batch=64, lr=1e-3, seed=5)
        assert got["loss"] == loss
        np.testing.assert_array_equal(got["params"], mlp.get_params()[0])
    ens.close()
