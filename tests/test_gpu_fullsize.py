"""The bench workload at its full size: 800 seeded faults, 79 378 elements, Mode II,
H-matrix operators (eta 2, ACA 1e-6, leaf 64).  The oracle applies the very same
block lists (exported from the device build, imported into the oracle), so the
stress-transfer operators must agree to 1e-12 relative (norm-wise, FP64) and the
first 30 solver steps from the common initial state to 1e-8 elementwise with the
same step counts, as the small-network parity tests require.  FP32 operator
storage keeps the stress rates within 1e-6 norm-wise of FP64 at this size.  The
same network with every fault intersection snapped to a shared node (3403
hydraulic crossings, the bench's connected variant) runs its pressure solves over
the crossing terms and halves steps whose solve fails: its first steps must take
the oracle's halvings and steps, with the same state."""
import numpy as np
import pytest

from demystification_dfn_amd import _abi, problems
from oracle.oracle import OracleSimulation

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    from demystification_dfn_amd.device import DeviceSimulation
    p = problems.network_80k()
    print(f"[gpu] full-size network N={p.N}: device build", flush=True)
    d = DeviceSimulation(p)
    d.initiate_hmat()
    print("[gpu] full-size network: oracle on the exported block lists", flush=True)
    o = OracleSimulation(p, threads=16)
    for op in (0, 1):
        o.hmat_import(op, d.hmat_export(op))
    return p, d, o


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_full_size_operators_match_on_identical_block_lists(pair):
    p, d, o = pair
    rng = np.random.default_rng(11)
    for op in (0, 1):
        for x in (rng.standard_normal(p.N), 10.0 ** rng.uniform(-12.0, 0.0, p.N)):
            err = _rel(d.Hgemv_right(x, op=op), o.Hgemv_right(x, op=op))
            assert err <= 1e-12, (op, err)
    V = 10.0 ** rng.uniform(-12.0, 0.0, p.N)
    for a, b in zip(d.kernel(V), o.kernel(V)):  # through the permutation and the Mode II scale
        assert _rel(a, b) <= 1e-12


def test_full_size_first_steps_match_the_oracle(pair):
    p, d, o = pair
    nsteps = 30
    print(f"[gpu] full-size network: {nsteps} steps on both", flush=True)
    sd, so = d.solve(nsteps), o.solve(nsteps)
    assert sd["accepted_steps"] == so["accepted_steps"] == nsteps
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    Yd, Pd, _, td = d.get_state()
    Yo, Po, _, to = o.get_state()
    assert abs(td - to) <= 1e-8 * to
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for c in range(3):  # slip rate, log state, normal stress
        a, b = Yd[c * N:(c + 1) * N][real], Yo[c * N:(c + 1) * N][real]
        assert np.all(np.abs(a - b) <= 1e-8 * np.abs(b)), c
    scale = max(np.abs(Po).max(), 1.0)
    assert np.max(np.abs(Pd - Po)) <= 1e-9 * scale + 10 * p.tol_solver


def test_full_size_fp32_operator_error(pair):
    p, d, _ = pair
    rng = np.random.default_rng(12)
    worst = 0.0
    try:
        for _ in range(20):
            V = 10.0 ** rng.uniform(-12.0, 0.0, p.N)
            d.set_precision(_abi.PREC_FP64)
            a = d.kernel(V)
            d.set_precision(_abi.PREC_FP32)
            b = d.kernel(V)
            worst = max(worst, max(_rel(y32, y64) for y32, y64 in zip(b, a)))
    finally:
        d.set_precision(_abi.PREC_FP64)
    print(f"[gpu] full-size FP32 operators: worst norm-wise error {worst:.2e}", flush=True)
    assert worst <= 1e-6


def test_full_size_split_trajectory_tracks_fp64(pair):
    """The 80k network stepped 20 times from its initial state with FP64 operators and
    with the split near-field storage (DFN_PREC_MIXED_SPLIT, the bench's headline
    form), each on a fresh handle that builds its own operators (so both start from
    the same trial step): the same accepted and rejected steps, the clocks within 1e-9
    relative and the slip rate, log state and normal stress within 1e-8 relative on
    every real element (measured on an MI355X: 4e-11 and 5e-10 at most)."""
    from demystification_dfn_amd.device import DeviceSimulation
    p = pair[0]
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    runs = {}
    for prec in (_abi.PREC_FP64, _abi.PREC_MIXED_SPLIT):
        e = DeviceSimulation(p)
        try:
            e.initiate_hmat()
            e.set_precision(prec)
            s = e.solve(20)
            Y, _, _, t = e.get_state()
            runs[prec] = (Y, t, s["accepted_steps"], s["rejected_attempts"])
        finally:
            e.close()
    (Ya, ta, acc_a, rej_a), (Yb, tb, acc_b, rej_b) = runs[_abi.PREC_FP64], runs[_abi.PREC_MIXED_SPLIT]
    diffs = [float(np.max(np.abs(Yb[c * N:(c + 1) * N][real] - Ya[c * N:(c + 1) * N][real]) /
                          np.abs(Ya[c * N:(c + 1) * N][real]))) for c in range(3)]
    print(f"[gpu] full-size split vs FP64 over 20 steps: steps {acc_a}/{rej_a} and {acc_b}/{rej_b}, "
          f"clocks {abs(tb - ta) / ta:.2e}, V / theta / sigmaN {diffs}", flush=True)
    assert acc_a == acc_b == 20 and rej_a == rej_b
    assert abs(tb - ta) <= 1e-9 * ta
    assert max(diffs) <= 1e-8


def test_full_size_mixed_operator_error(pair):
    """The 80k network with FP64 near-field blocks and FP32 low-rank factors read
    into FP64 sums: the stress-rate error over 20 log-uniform slip-rate vectors,
    norm-wise and on the rows above 1e-3 of the largest stress rate (bar 1e-5),
    beside all-FP32 storage, and with the near-field entries split into an FP32
    value and an 8-bit residual (DFN_PREC_MIXED_SPLIT).  The FP64 references come
    from one multi-RHS call."""
    p, d, _ = pair
    rng = np.random.default_rng(13)
    X = 10.0 ** rng.uniform(-12.0, 0.0, (20, p.N))
    Y64 = d.kernel_multi(X)
    worst = {_abi.PREC_FP32: [0.0, 0.0], _abi.PREC_MIXED: [0.0, 0.0], _abi.PREC_MIXED_SPLIT: [0.0, 0.0]}
    try:
        for prec in (_abi.PREC_FP32, _abi.PREC_MIXED, _abi.PREC_MIXED_SPLIT):
            d.set_precision(prec)
            for r in range(len(X)):
                for y, y64 in zip(d.kernel(X[r]), (Y64[0][r], Y64[1][r])):
                    big = np.abs(y64) > 1e-3 * np.abs(y64).max()
                    w = worst[prec]
                    w[0] = max(w[0], _rel(y, y64))
                    w[1] = max(w[1], float(np.max(np.abs(y[big] - y64[big]) / np.abs(y64[big]))))
    finally:
        d.set_precision(_abi.PREC_FP64)
    print(f"[gpu] full-size norm-wise / significant-row errors: FP32 {worst[_abi.PREC_FP32]}, "
          f"mixed {worst[_abi.PREC_MIXED]}, split near-field {worst[_abi.PREC_MIXED_SPLIT]}", flush=True)
    for prec in (_abi.PREC_MIXED, _abi.PREC_MIXED_SPLIT):
        assert worst[prec][0] <= 1e-6
        assert worst[prec][1] <= 1e-5


@pytest.fixture(scope="module")
def connected_pair():
    from demystification_dfn_amd.device import DeviceSimulation
    p = problems.network_80k(snap=True, name="network800_80k_connected_seed2024")
    print(f"[gpu] full-size connected network N={p.N}, {p.meta['snapped_crossings']} crossings: device build",
          flush=True)
    d = DeviceSimulation(p)
    d.initiate_hmat()
    o = OracleSimulation(p, threads=16)
    for op in (0, 1):
        o.hmat_import(op, d.hmat_export(op))
    return p, d, o


def test_full_size_connected_steps_match_the_oracle(connected_pair):
    p, d, o = connected_pair
    assert p.meta["snapped_crossings"] > 3000
    nsteps = 30  # the first steps ramp up from t_try = 100 s; the halvings come once t_try outgrows the solve
    sd, so = d.solve(nsteps), o.solve(nsteps)
    print(f"[gpu] connected: device {sd}, oracle {so}", flush=True)
    assert sd["accepted_steps"] == so["accepted_steps"] == nsteps
    assert sd["rejected_attempts"] == so["rejected_attempts"]
    assert sd["cg_halvings"] == so["cg_halvings"] > 0
    Yd, Pd, _, td = d.get_state()
    Yo, Po, _, to = o.get_state()
    assert abs(td - to) <= 1e-8 * to
    N = p.N
    real = np.ones(N, bool)
    real[p.fake_indices()] = False
    for c in range(3):
        a, b = Yd[c * N:(c + 1) * N][real], Yo[c * N:(c + 1) * N][real]
        assert np.all(np.abs(a - b) <= 1e-8 * np.abs(b)), (c, float(np.max(np.abs(a - b) / np.abs(b))))
    scale = max(np.abs(Po).max(), 1.0)
    assert np.max(np.abs(Pd - Po)) <= 1e-9 * scale + 10 * p.tol_solver
