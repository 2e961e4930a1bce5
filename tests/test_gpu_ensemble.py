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


@pytest.mark.parametrize("prec", [_abi.PREC_FP32, _abi.PREC_MIXED, _abi.PREC_FP32_STORE, _abi.PREC_MIXED_SPLIT])
def test_ensemble_reduced_operators_equal_independent_runs(prec):
    _compare(_nets(2, n=4, seed=9), 5, prec=prec)


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_ensemble_mixed_fused_phase1_equals_unfused_independent_runs(fuse, monkeypatch):
    """Mixed precision: the ensemble with phase 1 and the near-field panels in one
    launch (or two), its networks' own runs the other way; every network is
    bit-identical (the tasks of networks sitting out a retry return at once in
    either kind of workgroup)."""
    monkeypatch.setenv("DFN_FUSE_NEAR", fuse)
    _compare(_nets(2, n=4, seed=9), 5, prec=_abi.PREC_MIXED, env={"DFN_FUSE_NEAR": "0" if fuse == "1" else "1"},
             monkeypatch=monkeypatch)


@pytest.mark.parametrize("cg_path", ["auto", "plain"])
def test_ensemble_with_a_large_network(cg_path, monkeypatch):
    """A network above the one-workgroup compute_P size beside two small ones: the
    ensemble solves the large network's systems with the cooperative kernel (by
    default; each system over its own span of workgroups, the launches mapping
    workgroups to systems) or the plain one-workgroup kernel (DFN_CG_PATH=plain),
    the small ones with the one-workgroup LDS kernel, as single runs do."""
    monkeypatch.setenv("DFN_CG_PATH", cg_path)
    big = problems.fault_network(n_faults=120, box=2000.0, lmin=100.0, lmax=800.0, alpha=0.5, seed=3, eta=2.0,
                                 eps_aca=1e-6, nleaf=32, tol=1e-8)
    assert big.N > 4096
    probs = [big] + _nets(2, n=2, seed=13)
    _compare(probs, 3)


@pytest.mark.timeout(900)
def test_full_size_ensemble_networks_equal_independent_runs():
    """The 64-network ensemble at its stated size (50-400 faults per network in
    7 km boxes, ~1.4M elements): after 10 FP64 steps, 10 more with FP32 operators
    and 10 with the split near-field storage of the bench's headline, the smallest,
    the median and the largest network are bit-identical to their own runs of the
    same 30 steps."""
    import time
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceSimulation
    probs = problems.network_ensemble(64, faults=(50, 400), box=7000.0, seed=2024)
    Ns = np.array([p.N for p in probs])
    picks = [int(np.argmin(Ns)), int(np.argsort(Ns)[len(Ns) // 2]), int(np.argmax(Ns))]
    print(f"[gpu] full-size ensemble: {len(probs)} networks, {Ns.sum()} elements, checking {picks} "
          f"(N={Ns[picks].tolist()})", flush=True)
    t0 = time.perf_counter()
    ens = DeviceEnsemble(probs)
    ens.initiate_hmat()
    print(f"[gpu] ensemble built in {time.perf_counter() - t0:.1f} s", flush=True)
    ens.solve(10)
    ens.set_precision(_abi.PREC_FP32)
    print(f"[gpu] 10 FP64 steps done, FP32 layout at {time.perf_counter() - t0:.1f} s", flush=True)
    ens.solve(10)
    ens.set_precision(_abi.PREC_MIXED_SPLIT)
    print(f"[gpu] 10 FP32 steps done, split layout at {time.perf_counter() - t0:.1f} s", flush=True)
    ens.solve(10)
    print(f"[gpu] 10 split steps done at {time.perf_counter() - t0:.1f} s", flush=True)
    for i in picks:
        d = DeviceSimulation(probs[i])
        d.initiate_hmat()
        d.solve(10)
        d.set_precision(_abi.PREC_FP32)
        d.solve(10)
        d.set_precision(_abi.PREC_MIXED_SPLIT)
        sd = d.solve(10)
        Yd, Pd, Pdd, td = d.get_state()
        Ye, Pe, Pde, te = ens.get_state(i)
        se = ens.stats(i)
        print(f"[gpu] network {i} N={probs[i].N}: accepted {se['accepted_steps']} rejected "
              f"{se['rejected_attempts']} t={te:.6e}", flush=True)
        assert se["accepted_steps"] == sd["accepted_steps"] == 30
        assert se["rejected_attempts"] == sd["rejected_attempts"]
        assert te == td
        np.testing.assert_array_equal(Ye, Yd)
        np.testing.assert_array_equal(Pe, Pd)
        np.testing.assert_array_equal(Pde, Pdd)
        d.close()


def test_pi_and_learned_controllers_track_the_reference_controller():
    """Recorded (features, accepted-step) pairs train the proposal network; the
    rkqs, PI and learned controllers then step the same networks to a common
    final time (the last step clipped to land on it), and so does rkqs at a
    1000x tighter tolerance as the accurate solution.  Every step is still
    accepted by the Cash-Karp test, so the PI and learned solutions must be as
    close to the accurate one as the rkqs solution at the same tolerance is
    (within a factor 10 plus 1e-6): slip rate in max norm relative to max |V|,
    log state and sigma_N in max norm."""
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceMlp
    tol = 1e-6
    probs = _nets(2, n=3, seed=21, tols=(tol,))
    fine = _nets(2, n=3, seed=21, tols=(tol * 1e-3,))
    mlp = DeviceMlp()
    rec = DeviceEnsemble(probs)
    rec.initiate_hmat()
    rec.set_recording(True)
    rec.solve(80)
    X, T = rec.take_records()
    # steps whose successor the one-day cap sets are left out, so only the steps
    # before the networks settle at the cap give pairs
    print(f"[gpu] {T.shape[0]} recorded pairs", flush=True)
    assert X.shape == (T.shape[0], 8) and T.shape[0] >= 8
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(T))
    norm = np.concatenate([X.mean(0), X.std(0) + 1e-3, [T.mean()], [T.std() + 1e-3]])
    p0, _ = mlp.get_params()
    mlp.set_params(p0, norm)
    l0, _ = mlp.loss_grad(X, T)
    lf = mlp.train(X, T, epochs=20, batch=8, lr=1e-3, seed=1)
    assert lf <= l0
    t_end = min(rec.stats(i)["time"] for i in range(len(probs)))
    states = {}
    for key, ps, mode in (("fine", fine, _abi.STEP_RKQS), ("rkqs", probs, _abi.STEP_RKQS),
                          ("pi", probs, _abi.STEP_PI), ("learned", probs, _abi.STEP_LEARNED)):
        ens = DeviceEnsemble(ps)
        ens.initiate_hmat()
        ens.set_controller(mode, mlp if mode == _abi.STEP_LEARNED else None)
        ens.solve_until(t_end, max_steps=20000)
        states[key] = [ens.get_state(i) for i in range(len(ps))]
        for i in range(len(ps)):
            assert states[key][i][3] == t_end
            print(f"[gpu] {key} net {i}: {ens.stats(i)['accepted_steps']} steps, "
                  f"{ens.stats(i)['rejected_attempts']} rejected", flush=True)

    def err(key, i, p):
        N = p.N
        real = np.ones(N, bool)
        real[p.fake_indices()] = False
        a, b = states["fine"][i][0], states[key][i][0]
        ev = np.max(np.abs(a[:N][real] - b[:N][real])) / np.max(np.abs(a[:N][real]))
        ep = np.max(np.abs(a[N:2 * N][real] - b[N:2 * N][real]))
        es = np.max(np.abs(a[2 * N:][real] - b[2 * N:][real]) / np.abs(a[2 * N:][real]))
        return np.array([ev, ep, es])

    for i, p in enumerate(probs):
        e_ref = err("rkqs", i, p)
        for key in ("pi", "learned"):
            e = err(key, i, p)
            print(f"[gpu] net {i} {key}: errors {e} (rkqs {e_ref})", flush=True)
            assert np.all(e <= 10 * e_ref + 1e-6), (key, i, e, e_ref)


def test_bench_ensemble_sampled_networks_equal_independent_runs():
    """The bench's ensemble workload at its size (64 seeded networks, 123k elements,
    FP64 with graphs): after 20 ensemble steps, eight networks spread over the size
    range are bit-identical to their own single-network runs."""
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceSimulation
    probs = problems.network_ensemble(64, faults=(5, 36), box=2500.0, seed=2024)
    print(f"[gpu] bench ensemble: {len(probs)} networks, {sum(p.N for p in probs)} elements", flush=True)
    ens = DeviceEnsemble(probs)
    ens.initiate_hmat()
    ens.solve(20)
    order = np.argsort([p.N for p in probs])
    for i in order[::9]:
        p = probs[i]
        d = DeviceSimulation(p)
        d.initiate_hmat()
        sd = d.solve(20)
        Yd, Pd, _, td = d.get_state()
        Ye, Pe, _, te = ens.get_state(int(i))
        assert ens.stats(int(i))["accepted_steps"] == sd["accepted_steps"] == 20
        assert te == td
        np.testing.assert_array_equal(Ye, Yd)
        np.testing.assert_array_equal(Pe, Pd)


def test_trained_proposal_network_against_rkqs():
    """The proposal network trained on the GPU on 1e6 recorded pairs
    (demystification_dfn_amd/data/step_network.npz, written by
    train_step_network) on held-out networks: rkqs (src/ode_solver/ode_bs.f90:
    156-170) and the learned controller step the same 40 seeded networks
    (tolerances 1e-6 to 1e-10) from t = 0 to a common final time.  Reported: the
    rejected-attempt fraction, the stress evaluations (6 per first attempt, 5 per
    retried one) and the aggregate simulated time per wall hour of each.  The
    learned controller may not cost more: at most 1 % more stress evaluations and
    no more rejected attempts than rkqs."""
    import time
    from demystification_dfn_amd import train_step_network as tsn
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceMlp
    params, norm, meta = tsn.load()
    assert meta["recording"]["pairs"] >= 1_000_000
    assert meta["training"]["loss_final"] < 0.5 * meta["training"]["loss_initial"]
    mlp = DeviceMlp()
    mlp.set_params(params, norm)
    probs = tsn.recording_ensemble(40, seed=99)  # seeds 11.. were recorded for training
    t_end = 2.0e6
    out = {}
    for key, mode in (("rkqs", _abi.STEP_RKQS), ("learned", _abi.STEP_LEARNED)):
        ens = DeviceEnsemble(probs)
        ens.initiate_hmat()
        ens.set_controller(mode, mlp if mode == _abi.STEP_LEARNED else None)
        w0 = time.perf_counter()
        ens.solve_until(t_end, max_steps=200000)
        wall = time.perf_counter() - w0
        st = [ens.stats(i) for i in range(len(probs))]
        acc = sum(s["accepted_steps"] for s in st)
        rej = sum(s["rejected_attempts"] for s in st)
        out[key] = dict(accepted=acc, rejected=rej, rejected_fraction=rej / (acc + rej), evaluations=6 * acc + 5 * rej,
                        network_yr_per_hr=len(probs) * t_end / 3.15576e7 / (wall / 3600.0))
        print(f"[gpu] {key}: {out[key]}", flush=True)
        ens.close()
    r, l = out["rkqs"], out["learned"]
    print(f"[gpu] learned / rkqs: evaluations {l['evaluations'] / r['evaluations']:.4f}, rejected "
          f"{l['rejected']} vs {r['rejected']}, rate {l['network_yr_per_hr'] / r['network_yr_per_hr']:.3f}", flush=True)
    assert l["rejected"] <= r["rejected"]
    assert l["evaluations"] <= 1.01 * r["evaluations"]
