"""Parity of the gfx950 library (through its C ABI) with the CPU oracle on the same
seeded inputs.

Bars: stress-transfer matvec relative error <= 1e-12 in
plain FP64; slip rate / log state within 1e-8 relative over interseismic steps.
compute_P is a CG iteration whose sums are reduced in a different order on the
device, so pressures are compared to the CG stopping tolerance (max residual of
the row-scaled system below tol_solver Pa)."""
import sys

import numpy as np
import pytest

import cases
from demystification_dfn_amd import _abi, problems
from oracle.oracle import OracleSimulation

pytestmark = pytest.mark.gpu


def dev(problem):
    from demystification_dfn_amd.device import DeviceSimulation
    print(f"[gpu] {problem.name} N={problem.N} mode={problem.fracture_mode} kernel={problem.static_kernel}",
          flush=True)
    sys.stdout.flush()
    return DeviceSimulation(problem)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def close_elementwise(a, b, rtol, floor):
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), floor))


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("make", [cases.single_fault_small, cases.crossing_network])
def test_direct_kernel_and_rhs(make):
    p = make()
    d, o = dev(p), OracleSimulation(p)
    V = np.random.default_rng(0).random(p.N) * 1e-9
    for a, b in zip(d.kernel(V), o.kernel(V)):
        if np.linalg.norm(b) > 0:
            assert rel(a, b) <= 1e-12
    Y0 = p.initial_Y()
    for x in (1.0, 1000.0, 20000.0):
        yd, yo = d.friction_law(x, Y0), o.friction_law(x, Y0)
        assert close_elementwise(yd, yo, 1e-9, 1e-300 + np.abs(yo).max() * 1e-12), x


# compute_P has three kernels (one workgroup per system with the direction in
# LDS, cooperative multi-workgroup, plain one-workgroup); the library picks by
# size and DFN_CG_PATH forces a choice so each is checked here.
CG_PATHS = ["auto", "coop", "plain"]


@pytest.mark.parametrize("cg_path", CG_PATHS)
@pytest.mark.parametrize("make", [cases.single_fault_small, cases.crossing_network, cases.grid_network,
                                  cases.connected_network, cases.connected_network_spans])
def test_compute_P(make, cg_path, monkeypatch):
    monkeypatch.setenv("DFN_CG_PATH", cg_path)
    p = make()
    d, o = dev(p), OracleSimulation(p)
    for dt in (10.0, 5000.0, 86400.0):
        a, ad, ea = d.compute_P(dt)
        b, bd, eb = o.compute_P(dt)
        assert ea == eb, dt
        if ea == 0:
            scale = max(np.abs(b).max(), 1.0)
            assert np.max(np.abs(a - b)) <= 1e-9 * scale + 10 * p.tol_solver, dt
            assert np.max(np.abs(ad - bd)) <= (1e-9 * scale + 10 * p.tol_solver) / dt, dt


@pytest.mark.parametrize("cg_path", CG_PATHS)
@pytest.mark.parametrize("make,nsteps", [(cases.single_fault_small, 20), (cases.crossing_network, 30),
                                          (cases.grid_network, 4), (cases.connected_network, 12)])
def test_solver_trajectory(make, nsteps, cg_path, monkeypatch):
    monkeypatch.setenv("DFN_CG_PATH", cg_path)
    p = make()
    d, o = dev(p), OracleSimulation(p)
    sd, so = d.solve(nsteps), o.solve(nsteps)
    Yd, Pd, _, td = d.get_state()
    Yo, Po, _, to = o.get_state()
    assert sd["accepted_steps"] == so["accepted_steps"] == nsteps
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    assert abs(td - to) <= 1e-8 * to
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    assert close_elementwise(Yd[:N][real], Yo[:N][real], 1e-8, 1e-300)          # slip rate
    assert close_elementwise(Yd[N:2 * N][real], Yo[N:2 * N][real], 1e-8, 1e-300)  # log state
    assert close_elementwise(Yd[2 * N:][real], Yo[2 * N:][real], 1e-8, 1e-300)   # normal stress


def _permeable_crossing_network():
    """The crossing network with 100 times the permeability: compute_P needs more
    than 200 iterations at some trial steps (13 halvings over 30 steps)."""
    p = cases.crossing_network()
    p.arrays["permeability"] = p.arrays["permeability"] * 100.0
    return p


@pytest.mark.parametrize("make", [_permeable_crossing_network])
def test_step_halvings_match_the_oracle(make):
    """Trial steps whose compute_P fails: the device solves compute_P(t_try / 2)
    beside compute_P(t_try) and uses it when t_try fails; the halving sequence,
    step counts and states must be those of the oracle, which halves and solves
    again as the reference does (src/solver.f90:81-92)."""
    p = make()
    d, o = dev(p), OracleSimulation(p)
    sd, so = d.solve(30), o.solve(30)
    print(f"[gpu] halvings device {sd['cg_halvings']} oracle {so['cg_halvings']}", flush=True)
    assert sd["cg_halvings"] == so["cg_halvings"] > 0
    assert sd["accepted_steps"] == so["accepted_steps"] and sd["rejected_attempts"] == so["rejected_attempts"]
    Yd, Yo = d.get_state()[0], o.get_state()[0]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Yd[sl][real], Yo[sl][real], 1e-8, 1e-300)


def test_rkqs_single_call():
    p = cases.single_fault_small()
    d, o = dev(p), OracleSimulation(p)
    Y0 = p.initial_Y()
    yd = o.friction_law(200.0, Y0)
    yscal = np.abs(Y0) + np.abs(yd * 200.0) + p.tol_solver
    y1, x1, h1, n1 = d.rkqs(Y0, yd, 0.0, 200.0, p.tol_solver, yscal)
    y2, x2, h2, n2 = o.rkqs(Y0, yd, 0.0, 200.0, p.tol_solver, yscal)
    assert h1 == pytest.approx(h2, rel=1e-12) and n1 == pytest.approx(n2, rel=1e-9)
    assert close_elementwise(y1, y2, 1e-10, 1e-300)


# ---------------------------------------------------------------------------
def _hnet(mode, n_faults=14, seed=5):
    return problems.network_small(n_faults=n_faults, seed=seed, mode=mode, eta=2.0, eps_aca=1e-6, nleaf=16,
                                  static_kernel=_abi.KERNEL_HMATRIX)


@pytest.mark.parametrize("mode", [2, 3])
def test_hmatvec_on_identical_block_lists(mode):
    """Device H-matvec vs the oracle applying the very same block lists."""
    p = _hnet(mode)
    d, o = dev(p), OracleSimulation(p)
    d.initiate_hmat()
    ops = (0, 1) if mode == 2 else (0,)
    for op in ops:
        o.hmat_import(op, d.hmat_export(op))
    rng = np.random.default_rng(1)
    for trial in range(3):
        V = 10.0 ** rng.uniform(-12, 0, p.N)
        (aT, aN), (bT, bN) = d.kernel(V), o.kernel(V)
        assert rel(aT, bT) <= 1e-12
        if mode == 2:
            assert rel(aN, bN) <= 1e-12
    x = rng.standard_normal(p.N)
    for op in ops:
        assert rel(d.Hgemv_right(x, op=op), o.Hgemv_right(x, op=op)) <= 1e-12


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_FP32, _abi.PREC_MIXED, _abi.PREC_FP32_STORE,
                                  _abi.PREC_MIXED_SPLIT])
def test_autotuned_launch_shapes_leave_results_unchanged(prec, monkeypatch):
    """The autotuner only picks launch shapes that keep every sum in the same order:
    the stress rates of the tuned layout, of the default shapes and of every
    alternative shape (DFN_AUTOTUNE=alt) are equal bit for bit, and so are those with
    short phase-1 tasks two to a wave or one (DFN_VTX_PAIRS) and, with FP64 factors,
    those read with non-temporal loads or ordinary ones (DFN_VTX_NT)."""
    p = _hnet(2, n_faults=40, seed=8)
    keys = ("vtx_waves", "vtx_order", "panel_rows", "panel_order", "panel_split")
    out = []
    # (autotune, single launch of phase 1 and the near-field panels, phase-1 wave items,
    # non-temporal FP64 factor loads)
    runs = [("1", None, None, None), ("0", None, None, None), ("alt", None, None, None), ("0", None, "1", None),
            ("1", None, "0", None), ("alt", None, "0", None)]
    if prec == _abi.PREC_MIXED:
        runs += [("1", "1", None, None), ("1", "0", None, None), ("0", "1", None, None)]
    if prec == _abi.PREC_FP64:
        runs += [("0", None, None, "1"), ("1", None, "1", "1"), ("alt", None, None, "0"), ("1", None, None, "0")]
    for tune, fuse, pairs, nt in runs:
        monkeypatch.setenv("DFN_AUTOTUNE", tune)
        if nt is None:
            monkeypatch.delenv("DFN_VTX_NT", raising=False)
        else:
            monkeypatch.setenv("DFN_VTX_NT", nt)
        if fuse is None:
            monkeypatch.delenv("DFN_FUSE_NEAR", raising=False)
        else:
            monkeypatch.setenv("DFN_FUSE_NEAR", fuse)
        if pairs is None:
            monkeypatch.delenv("DFN_VTX_PAIRS", raising=False)
        else:
            monkeypatch.setenv("DFN_VTX_PAIRS", pairs)
        d = dev(p)
        d.initiate_hmat()
        if prec != _abi.PREC_FP64:
            d.set_precision(prec)
        t = d.tuning()
        print(f"[gpu] autotune={tune} fuse={fuse} pairs={pairs} nt={nt} {t}", flush=True)
        assert t["vtx_waves"] in (2, 4, 8) and t["vtx_order"] in (0, 1, 2) and t["panel_rows"] in (2, 4)
        assert t["panel_order"] in (0, 1) and t["panel_split"] in (0, 32, 16)
        if tune == "0":
            assert [t[k] for k in keys] == [4, 0, 2, 0, 0]
        if tune == "alt":
            assert [t[k] for k in keys] == [8, 2, 4, 1, 16]
        assert t["rank_ceiling"] in (16, 20, 32)
        if tune == "alt":
            assert t["rank_ceiling"] == 32
        if pairs is not None:
            assert t["vtx_pairs"] == int(pairs)
        elif tune == "alt":
            assert t["vtx_pairs"] == 1
        elif tune == "0":
            assert t["vtx_pairs"] == 0
        if fuse is not None:
            assert t["fused_near"] == int(fuse)
        elif prec != _abi.PREC_MIXED:
            assert t["fused_near"] == 0
        V = 10.0 ** np.random.default_rng(4).uniform(-12, 0, p.N)
        out.append(d.kernel(V))
    for other in out[1:]:
        for a, b in zip(out[0], other):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("mode", [2, 3])
def test_kernel_timing_covers_every_stress_evaluation(mode):
    """dfn_kernel_timing times one launch per stress evaluation of the solver (the
    H-matrix panel kernel) and nothing once stopped."""
    p = _hnet(mode, n_faults=20, seed=3)
    d = dev(p)
    d.initiate_hmat()
    s0 = d.stats()
    d.kernel_timing(True)
    d.solve(6)
    ms, n = d.kernel_timing_read()
    s1 = d.stats()
    rhs = s1["rhs_evaluations"] - s0["rhs_evaluations"]
    # a step whose first attempt ran before its compute_P(t_try) failed evaluates
    # the stress of its state once more
    assert n >= rhs >= 36
    if s1["cg_halvings"] == s0["cg_halvings"]:
        assert n == rhs
    assert 0.0 < ms < 100.0
    d.kernel_timing(False)
    d.solve(2)
    assert d.kernel_timing_read() == (0.0, 0)


@pytest.mark.parametrize("builder", ["device", "host"])
@pytest.mark.parametrize("mode", [2, 3])
def test_device_built_operator_matches_oracle_construction(mode, builder, monkeypatch):
    """Trees, ranks and factors of the library's construction (GPU assembly and
    ACA, or the all-host path) vs the oracle's restatement of src/hmatrix
    (cluster tree, block tree, full-pivot ACA, QR+SVD recompression)."""
    monkeypatch.setenv("DFN_HBUILD", builder)
    p = _hnet(mode)
    d, o = dev(p), OracleSimulation(p)
    d.initiate_hmat()
    o.initiate_hmat()
    for op in ((0, 1) if mode == 2 else (0,)):
        a, b = d.hmat_export(op), o.hmat_export(op)
        for key in ("perm", "dense_row0", "dense_col0", "dense_m", "dense_n", "lr_row0", "lr_col0", "lr_m", "lr_n",
                    "lr_k"):
            np.testing.assert_array_equal(a[key], b[key], err_msg=key)
        np.testing.assert_array_equal(a["dense_data"], b["dense_data"])
        x = np.random.default_rng(op).standard_normal(p.N)
        assert rel(d.Hgemv_right(x, op=op), o.Hgemv_right(x, op=op)) <= 1e-12


OP_KEYS = ("perm", "dense_row0", "dense_col0", "dense_m", "dense_n", "dense_data", "lr_row0", "lr_col0", "lr_m",
           "lr_n", "lr_k", "lr_u", "lr_v")


@pytest.mark.parametrize("mode", [2, 3])
def test_gpu_construction_equals_host_construction_larger(mode, monkeypatch):
    """A 60-fault network (a few thousand elements, admissible leaves spanning
    many 64 x 256 tiles): the GPU assembly + ACA gives the very same operator,
    factor for factor, as the library's host construction."""
    p = problems.network_small(n_faults=60, seed=21, mode=mode, eta=2.0, eps_aca=1e-6, nleaf=32,
                               static_kernel=_abi.KERNEL_HMATRIX)
    monkeypatch.setenv("DFN_HBUILD", "host")
    a = dev(p)
    a.initiate_hmat()
    monkeypatch.setenv("DFN_HBUILD", "device")
    b = dev(p)
    b.initiate_hmat()
    for op in ((0, 1) if mode == 2 else (0,)):
        ea, eb = a.hmat_export(op), b.hmat_export(op)
        assert a.hmat_stats(op)["nb_lowrank"] > 50
        for key in OP_KEYS:
            np.testing.assert_array_equal(ea[key], eb[key], err_msg=key)


def test_hmatrix_solver_trajectory_mode2():
    p = _hnet(2)
    d, o = dev(p), OracleSimulation(p)
    d.initiate_hmat()
    for op in (0, 1):
        o.hmat_import(op, d.hmat_export(op))
    sd, so = d.solve(25), o.solve(25)
    Yd = d.get_state()[0]
    Yo = o.get_state()[0]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    assert close_elementwise(Yd[:N][real], Yo[:N][real], 1e-8, 1e-300)
    assert close_elementwise(Yd[N:2 * N][real], Yo[N:2 * N][real], 1e-8, 1e-300)


@pytest.mark.parametrize("prec", [_abi.PREC_FP32, _abi.PREC_FP32_STORE])
@pytest.mark.parametrize("mode", [2, 3])
def test_fp32_operator_storage_error(mode, prec):
    """Precision form: FP32 factors with per-block scaling (FP32 sums, or FP64 sums
    over the same storage) vs the FP64 operator, on log-uniform slip rates in
    [1e-12, 1] m/s; the bar is 1e-5 relative."""
    p = _hnet(mode, n_faults=20, seed=9)
    d = dev(p)
    d.initiate_hmat()
    rng = np.random.default_rng(3)
    for _ in range(20):
        V = 10.0 ** rng.uniform(-12, 0, p.N)
        d.set_precision(_abi.PREC_FP64)
        a = d.kernel(V)
        d.set_precision(prec)
        b = d.kernel(V)
        for y64, y32 in zip(a, b):
            if np.linalg.norm(y64) > 0:
                err = rel(y32, y64)
                assert err <= 1e-5, err
                assert err > 0  # really a different storage


@pytest.mark.parametrize("mode", [2, 3])
def test_mixed_operator_storage_error(mode):
    """FP64 near-field blocks beside FP32 low-rank factors read into FP64 products
    and sums (DFN_PREC_MIXED) vs the FP64 operator on log-uniform slip rates in
    [1e-12, 1] m/s: norm-wise within 1e-6, and on every row above 1e-3 of the
    largest stress rate within 1e-5 relative -- the rows where all-FP32 storage
    loses the most, because an entry stored in FP32 keeps its rounding times the
    row's cancellation ratio.  The split near-field storage (an FP32 value and an
    8-bit residual per entry, DFN_PREC_MIXED_SPLIT) keeps about 31 bits of a task's
    largest near-field entries: its stress rates sit within 1e-8 of the FP64
    near-field layout's, far closer than the 1e-6 an FP32 near-field alone would
    allow."""
    p = _hnet(mode, n_faults=20, seed=9)
    d = dev(p)
    d.initiate_hmat()
    rng = np.random.default_rng(3)
    worst = {_abi.PREC_FP32: 0.0, _abi.PREC_FP32_STORE: 0.0, _abi.PREC_MIXED: 0.0, _abi.PREC_MIXED_SPLIT: 0.0}
    split_vs_mixed = 0.0
    for _ in range(20):
        V = 10.0 ** rng.uniform(-12, 0, p.N)
        d.set_precision(_abi.PREC_FP64)
        a = d.kernel(V)
        outs = {}
        for prec in (_abi.PREC_FP32, _abi.PREC_FP32_STORE, _abi.PREC_MIXED, _abi.PREC_MIXED_SPLIT):
            d.set_precision(prec)
            b = outs[prec] = d.kernel(V)
            for y64, y in zip(a, b):
                if np.linalg.norm(y64) == 0:
                    continue
                if prec in (_abi.PREC_MIXED, _abi.PREC_MIXED_SPLIT):
                    assert 0 < rel(y, y64) <= 1e-6
                big = np.abs(y64) > 1e-3 * np.abs(y64).max()
                worst[prec] = max(worst[prec], float(np.max(np.abs(y[big] - y64[big]) / np.abs(y64[big]))))
        for ym, ys in zip(outs[_abi.PREC_MIXED], outs[_abi.PREC_MIXED_SPLIT]):
            if np.linalg.norm(ym) > 0:
                split_vs_mixed = max(split_vs_mixed, rel(ys, ym))
    print(f"[gpu] significant-row max error: FP32 {worst[_abi.PREC_FP32]:.2e}, FP32 storage with FP64 sums "
          f"{worst[_abi.PREC_FP32_STORE]:.2e}, mixed {worst[_abi.PREC_MIXED]:.2e}, split near-field "
          f"{worst[_abi.PREC_MIXED_SPLIT]:.2e}; split vs mixed {split_vs_mixed:.2e}", flush=True)
    assert worst[_abi.PREC_MIXED] <= 1e-5 and worst[_abi.PREC_MIXED_SPLIT] <= 1e-5
    assert worst[_abi.PREC_MIXED] < worst[_abi.PREC_FP32]
    assert 0 < split_vs_mixed <= 1e-8


@pytest.mark.parametrize("prec", [_abi.PREC_FP32, _abi.PREC_MIXED, _abi.PREC_FP32_STORE, _abi.PREC_MIXED_SPLIT])
def test_reduced_operator_trajectory_tracks_fp64(prec):
    p = _hnet(2)
    a, b = dev(p), dev(p)
    a.initiate_hmat()
    b.initiate_hmat()
    b.set_precision(prec)
    a.solve(15)
    b.solve(15)
    Ya, Yb = a.get_state()[0], b.get_state()[0]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    assert close_elementwise(Yb[:N][real], Ya[:N][real], 1e-4, 1e-300)


# ---------------------------------------------------------------------------
def _single_fault_4000():
    return problems.single_fault(n=4000, length=40e3, static_kernel=_abi.KERNEL_DIRECT, tol=1e-8)


def test_single_fault_4000_direct_kernel():
    """The 4000-element Mode III fault with the dense direct kernel: tiled device
    kernel vs the oracle's element loop, on smooth, random and log-uniform rates."""
    p = _single_fault_4000()
    d, o = dev(p), OracleSimulation(p, threads=8)
    rng = np.random.default_rng(11)
    x = np.linspace(-1, 1, p.N)
    for V in (1e-9 * np.exp(-x * x / 0.01), rng.random(p.N) * 1e-9, 10.0 ** rng.uniform(-12, 0, p.N)):
        assert rel(d.kernel(V)[0], o.kernel(V)[0]) <= 1e-12


def test_single_fault_4000_trajectory():
    """Injection into the 4000-element fault: 40 accepted solver steps (friction
    law, compute_P with the source, rkqs) track the oracle to 1e-8."""
    p = _single_fault_4000()
    d, o = dev(p), OracleSimulation(p, threads=8)
    sd, so = d.solve(40), o.solve(40)
    Yd, Pd, _, td = d.get_state()
    Yo, Po, _, to = o.get_state()
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    assert abs(td - to) <= 1e-8 * to
    N = p.N
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Yd[sl], Yo[sl], 1e-8, 1e-300)
    assert np.max(np.abs(Pd - Po)) <= 1e-9 * max(np.abs(Po).max(), 1.0) + 10 * p.tol_solver


# ---------------------------------------------------------------------------
def test_reference_input_file_network_against_reference_outputs():
    """The reference's own problems/network_rate_strengthening/config (10 Mode II
    faults, 1029 elements, read by the library's HDF5 reader into the fixture) with
    the dense Mode II kernel: the device's kernel, friction law and 20 solver
    steps against the outputs of the reference's compiled Fortran on it."""
    import os
    p = cases.reference_network()
    g = dict(np.load(os.path.join(os.path.dirname(cases.__file__), "network_rate_strengthening.npz")))
    d = dev(p)
    aT, aN = d.kernel(g["kernel_in"])
    assert rel(aT, g["kernel_T"]) <= 1e-12 and rel(aN, g["kernel_N"]) <= 1e-12
    Y0 = p.initial_Y()
    for x, ref in zip(g["rhs_x"], g["rhs"]):
        assert close_elementwise(d.friction_law(x, Y0), ref, 1e-9, 1e-300 + np.abs(ref).max() * 1e-12)
    s = d.solve(int(g["solve_steps"]))
    Y, P, _, t = d.get_state()
    assert abs(t - float(g["solve_time"])) <= 1e-8 * float(g["solve_time"])
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Y[sl][real], g["solve_Y"][sl][real], 1e-8, 1e-300)


def test_reference_input_file_network_hmatrix_on_identical_block_lists():
    """The same reference network with the H-matrix kernel at the reference's
    default eta / ACA / leaf (3, 1e-8, 100): device vs oracle on the same block
    lists, matvec and 15 solver steps."""
    p = cases.reference_network()
    p.static_kernel = _abi.KERNEL_HMATRIX
    d, o = dev(p), OracleSimulation(p)
    d.initiate_hmat()
    for op in (0, 1):
        o.hmat_import(op, d.hmat_export(op))
    x = np.random.default_rng(5).standard_normal(p.N)
    for op in (0, 1):
        assert rel(d.Hgemv_right(x, op=op), o.Hgemv_right(x, op=op)) <= 1e-12
    sd, so = d.solve(15), o.solve(15)
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    Yd, Yo = d.get_state()[0], o.get_state()[0]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Yd[sl][real], Yo[sl][real], 1e-8, 1e-300)


def test_connected_network_hmatrix_trajectory():
    """The snapped network (52 crossings) with the H-matrix kernel on identical
    block lists: 12 steps with rejections and crossing-coupled diffusion."""
    p = cases.connected_network(static_kernel=_abi.KERNEL_HMATRIX)
    d, o = dev(p), OracleSimulation(p, threads=8)
    d.initiate_hmat()
    for op in (0, 1):
        o.hmat_import(op, d.hmat_export(op))
    sd, so = d.solve(12), o.solve(12)
    assert sd["rejected_attempts"] == so["rejected_attempts"] > 0
    Yd, Pd = d.get_state()[:2]
    Yo, Po = o.get_state()[:2]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Yd[sl][real], Yo[sl][real], 1e-8, 1e-300)
    assert np.max(np.abs(Pd - Po)) <= 1e-9 * np.abs(Po).max() + 10 * p.tol_solver


def test_connected_network_spans_trajectory():
    """4855 elements, 143 crossings: the cooperative compute_P (three workgroup
    spans per system, crossing operands in other spans read from global memory)
    inside 6 H-matrix steps on identical block lists."""
    p = cases.connected_network_spans(static_kernel=_abi.KERNEL_HMATRIX)
    d, o = dev(p), OracleSimulation(p, threads=8)
    d.initiate_hmat()
    for op in (0, 1):
        o.hmat_import(op, d.hmat_export(op))
    sd, so = d.solve(6), o.solve(6)
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    assert sd["cg_iterations"] > 0
    Yd, Pd = d.get_state()[:2]
    Yo, Po = o.get_state()[:2]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Yd[sl][real], Yo[sl][real], 1e-8, 1e-300)
    assert np.max(np.abs(Pd - Po)) <= 1e-9 * np.abs(Po).max() + 10 * p.tol_solver


# ---------------------------------------------------------------------------
# The device library's own setup (host code of dfn_create) against the oracle,
# whose geometry, loading projections and crossings the CPU tests pin
# (test_oracle_setup_pins.py).
@pytest.mark.parametrize("mode", [2, 3])
def test_setup_loading_projections_match_the_oracle(mode):
    """Faults of every strike under all five stress rates: the uploaded state
    (Mode III adds the projected stress rates to sigma_N) and the friction law,
    which reads the projected loading rates, equal the oracle's."""
    p = cases.strike_fan(mode)
    d, o = dev(p), OracleSimulation(p)
    Yd, Yo = d.get_state()[0], o.get_state()[0]
    np.testing.assert_allclose(Yd, Yo, rtol=1e-15, atol=0)
    for x in (1.0, 1000.0, 20000.0):
        yd, yo = d.friction_law(x, Yo), o.friction_law(x, Yo)
        assert close_elementwise(yd, yo, 1e-9, 1e-300 + np.abs(yo).max() * 1e-12), x


@pytest.mark.parametrize("cg_path", CG_PATHS)
def test_junction_crossings_pressure_and_trajectory(cg_path, monkeypatch):
    """A T junction, an X crossing and a near miss on each side of the crossing
    threshold: compute_P and 10 solver steps, which carry the crossing terms of
    the pressure stencil, against the oracle."""
    monkeypatch.setenv("DFN_CG_PATH", cg_path)
    p = cases.junction_network()
    d, o = dev(p), OracleSimulation(p)
    for dt in (10.0, 5000.0, 86400.0):
        a, ad, ea = d.compute_P(dt)
        b, bd, eb = o.compute_P(dt)
        assert ea == eb, dt
        if ea == 0:
            assert np.max(np.abs(a - b)) <= 1e-9 * np.abs(b).max() + 10 * p.tol_solver, dt
    sd, so = d.solve(10), o.solve(10)
    assert sd["accepted_steps"] == so["accepted_steps"] and sd["rejected_attempts"] == so["rejected_attempts"]
    Yd, Yo = d.get_state()[0], o.get_state()[0]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Yd[sl][real], Yo[sl][real], 1e-8, 1e-300)


@pytest.mark.parametrize("cg_path", CG_PATHS)
def test_permeability_coupling_pressure_and_trajectory(cg_path, monkeypatch):
    """permeability_coupling = 1 (src/fluid_diffusion.f90:68-70): every compute_P
    first sets the permeability from the effective normal stress of the step-start
    state; the pressures, the permeability the steps leave and 30 solver steps
    against the oracle, which tests/golden/coupling_network.npz pins to the
    reference build."""
    monkeypatch.setenv("DFN_CG_PATH", cg_path)
    p = cases.coupling_network()
    d, o = dev(p), OracleSimulation(p)
    for dt in (10.0, 5000.0, 86400.0):
        a, ad, ea = d.compute_P(dt)
        b, bd, eb = o.compute_P(dt)
        assert ea == eb, dt
        if ea == 0:
            assert np.max(np.abs(a - b)) <= 1e-9 * np.abs(b).max() + 10 * p.tol_solver, dt
    sd, so = d.solve(30), o.solve(30)
    assert sd["accepted_steps"] == so["accepted_steps"] == 30
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    assert sd["cg_halvings"] == so["cg_halvings"]
    Yd, Pd, _, td = d.get_state()
    Yo, Po, _, to = o.get_state()
    assert abs(td - to) <= 1e-8 * to
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for sl in (slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)):
        assert close_elementwise(Yd[sl][real], Yo[sl][real], 1e-8, 1e-300)
    assert np.max(np.abs(Pd - Po)) <= 1e-8 * np.abs(Po).max() + 10 * p.tol_solver
