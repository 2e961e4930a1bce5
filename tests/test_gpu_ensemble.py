"""The ensemble engine (dfn_ensemble_*) against independent single-network runs.

Every network of an ensemble must follow exactly the trajectory it follows in
its own dfn_handle: same accepted and rejected steps, same clock, and the same
state, bit for bit (the ensemble's block-diagonal operator layout keeps each
network's panels, and its stage and compute_P kernels keep each network's
arithmetic).  Networks run with different tolerances so that some reject
attempts while others accept, which exercises the masked retries."""
import numpy as np
import pytest

from demystification_dfn_amd import _abi, problems

pytestmark = pytest.mark.gpu


def _nets(mode=2, n=5, faults=(8, 16), tols=(1e-6, 1e-9), seed=11, **kw):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nf = int(rng.integers(faults[0], faults[1] + 1))
        out.append(problems.network_small(n_faults=nf, seed=100 + i, mode=mode, eta=2.0, eps_aca=1e-6, nleaf=16,
                                          static_kernel=_abi.KERNEL_HMATRIX, tol=tols[i % len(tols)], **kw))
    return out


def _compare(probs, nsteps, prec=_abi.PREC_FP64, graph=True, env=None, monkeypatch=None):
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceSimulation
    print(f"[gpu] ensemble of {len(probs)} networks, N={[p.N for p in probs]} prec={prec} graph={graph}", flush=True)
    ens = DeviceEnsemble(probs)
    ens.set_graph(graph)
    ens.initiate_hmat()
    if prec != _abi.PREC_FP64:
        ens.set_precision(prec)
    # several solve calls: the captured graphs are replayed across calls
    for _ in range(3):
        ens.solve(nsteps)
    if env and monkeypatch:
        for k, v in env.items():
            monkeypatch.setenv(k, v)
    rejected = 0
    for i, p in enumerate(probs):
        d = DeviceSimulation(p)
        d.initiate_hmat()
        if prec != _abi.PREC_FP64:
            d.set_precision(prec)
        sd = d.solve(3 * nsteps)
        Yd, Pd, Pdd, td = d.get_state()
        Ye, Pe, Pde, te = ens.get_state(i)
        se = ens.stats(i)
        assert se["accepted_steps"] == sd["accepted_steps"] == 3 * nsteps
        assert se["rejected_attempts"] == sd["rejected_attempts"], (i, se, sd)
        assert te == td
        np.testing.assert_array_equal(Ye, Yd)
        np.testing.assert_array_equal(Pe, Pd)
        np.testing.assert_array_equal(Pde, Pdd)
        rejected += sd["rejected_attempts"]
    return rejected


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_mode2_equals_independent_runs(graph):
    rejected = _compare(_nets(2), 6, graph=graph)
    assert rejected > 0  # some networks retried attempts while others had accepted


def test_ensemble_mode3_equals_independent_runs():
    _compare(_nets(3, n=4, seed=5), 6)


def test_ensemble_fp32_operators_equal_independent_fp32_runs():
    _compare(_nets(2, n=4, seed=9), 5, prec=_abi.PREC_FP32)


def test_ensemble_with_a_large_network_uses_the_plain_compute_P(monkeypatch):
    """A network above the one-workgroup compute_P size: the ensemble solves every
    system with the plain one-workgroup kernel, as a single run does under
    DFN_CG_PATH=plain."""
    big = problems.fault_network(n_faults=120, box=2000.0, lmin=100.0, lmax=800.0, alpha=0.5, seed=3, eta=2.0,
                                 eps_aca=1e-6, nleaf=32, tol=1e-8)
    assert big.N > 4096
    probs = [big] + _nets(2, n=2, seed=13)
    _compare(probs, 3, env={"DFN_CG_PATH": "plain"}, monkeypatch=monkeypatch)
