"""Multi-RHS stress transfer (dfn_kernel_multi): the FP64 H-matrix applied to many
slip-rate vectors in one call, both phases as dense products on the matrix cores.

Every output vector must agree with the one-vector path (dfn_kernel, itself pinned
to the oracle) and with the oracle applying the very same block lists, to 1e-12
relative norm-wise: the bar of the plain FP64 matvec.  Vector counts that are not
multiples of 64 leave padding vectors in the last tile group, counts above 1024
split into batches, and blocks with more than 512 columns take the chunked phase 1
with its reduction."""
import sys

import numpy as np
import pytest

from demystification_dfn_amd import _abi, problems
from oracle.oracle import OracleSimulation

pytestmark = pytest.mark.gpu


def _net(mode, big=False):
    if big:  # 4 562 elements; 17 low-rank blocks wider than one 512-column phase-1 chunk
        return problems.network_small(n_faults=60, seed=3, box=2000.0, lmin=300.0, lmax=2000.0, mode=mode, eta=2.0,
                                      eps_aca=1e-6, nleaf=64, static_kernel=_abi.KERNEL_HMATRIX)
    return problems.network_small(n_faults=14, seed=5, mode=mode, eta=2.0, eps_aca=1e-6, nleaf=16,
                                  static_kernel=_abi.KERNEL_HMATRIX)


def _dev(p):
    from demystification_dfn_amd.device import DeviceSimulation
    print(f"[gpu] multi-RHS {p.name} N={p.N} mode={p.fracture_mode}", flush=True)
    sys.stdout.flush()
    d = DeviceSimulation(p)
    d.initiate_hmat()
    return d


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _vectors(n, N, seed):
    return 10.0 ** np.random.default_rng(seed).uniform(-12.0, 0.0, (n, N))


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("nrhs", [1, 37, 64, 130])
def test_multi_rhs_matches_one_vector_path_and_oracle(mode, nrhs):
    p = _net(mode)
    d = _dev(p)
    o = OracleSimulation(p)
    for op in ((0, 1) if mode == 2 else (0,)):
        o.hmat_import(op, d.hmat_export(op))
    X = _vectors(nrhs, p.N, 100 + nrhs)
    sT, sN = d.kernel_multi(X)
    assert sT.shape == X.shape and (sN is None) == (mode == 3)
    worst = 0.0
    for r in range(nrhs):
        aT, aN = d.kernel(X[r])
        worst = max(worst, _rel(sT[r], aT))
        if mode == 2:
            worst = max(worst, _rel(sN[r], aN))
    for r in sorted({0, nrhs // 2, nrhs - 1}):
        bT, bN = o.kernel(X[r])
        worst = max(worst, _rel(sT[r], bT))
        if mode == 2:
            worst = max(worst, _rel(sN[r], bN))
    print(f"[gpu] multi-RHS mode {mode} nrhs {nrhs}: worst relative error {worst:.2e}", flush=True)
    assert worst <= 1e-12


@pytest.mark.parametrize("mode", [2, 3])
def test_multi_rhs_chunked_blocks_and_batches(mode):
    """1100 vectors (a 1024-vector batch and a 76-vector one) on a network whose
    widest low-rank blocks span several phase-1 chunks."""
    p = _net(mode, big=True)
    d = _dev(p)
    X = _vectors(1100, p.N, 7)
    X[5] = 0.0  # an all-zero vector gives exactly zero stress rates
    sT, sN = d.kernel_multi(X)
    assert not np.any(sT[5]) and (sN is None or not np.any(sN[5]))
    for r in (0, 511, 1023, 1024, 1099):
        aT, aN = d.kernel(X[r])
        assert _rel(sT[r], aT) <= 1e-12, r
        if mode == 2:
            assert _rel(sN[r], aN) <= 1e-12, r
    # linearity over the vectors: H(a x + b y) = a H x + b H y
    Z = np.stack([X[1], X[2], 2.0 * X[1] - 3.0 * X[2]])
    zT, zN = d.kernel_multi(Z)
    assert _rel(zT[2], 2.0 * zT[0] - 3.0 * zT[1]) <= 1e-12
    if mode == 2:
        assert _rel(zN[2], 2.0 * zN[0] - 3.0 * zN[1]) <= 1e-12


def test_multi_rhs_edge_cases():
    p = _net(2)
    d = _dev(p)
    sT, sN = d.kernel_multi(np.zeros((0, p.N)))
    assert sT.shape == (0, p.N)
    with pytest.raises(ValueError):
        d.kernel_multi(np.zeros((3, p.N + 1)))
    # dense direct handles have no H-matrix to apply
    from demystification_dfn_amd.device import DeviceSimulation
    q = problems.network_small(n_faults=6, seed=2, mode=2, static_kernel=_abi.KERNEL_DIRECT)
    dd = DeviceSimulation(q)
    with pytest.raises(_abi.DfnError) as e:
        dd.kernel_multi(np.ones((2, q.N)))
    assert e.value.code == _abi.DFN_ERR_STATE
