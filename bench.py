"""Benchmark of the per-RK-stage hot path on one MI355X.

A step is one accepted step of the solver loop (src/solver.f90:59-198) of every
network: CG pressure solves, the stress-transfer H-matvec and the rate-state RHS
at every Cash-Karp stage, the embedded error norm and the step controller, all
on the device.

The default workload is the headline one: an ensemble of 64 seeded fault
networks of 50-400 faults each (~1.37M elements, H-matrix eta 2 / ACA 1e-6 /
leaf 64) advanced together, with reduced-precision operator storage read into
FP64 products and sums (near-field entries as an FP32 value plus an 8-bit
residual, FP32 low-rank factors; --headline-precision mixed keeps the near-field
in FP64; FP64 state, RHS, CG and error norm) and the learned step-size proposals (data/step_network.npz), against
the plain FP64 path (one dfn_handle per network, FP64 operators, rkqs) over the
same simulated interval of every network.

Prints ONE JSON line: simulated network-years per wall-clock hour of the whole
ensemble, the ratio to the plain FP64 path, the roofline of the dominant kernel
(the phase-2 panel kernels, HIP events on the ensemble's stream), a CPU
baseline (the oracle on this host's cores, from the state at the start of the
timed window), and the 80k-element single network (FP64 operators; FP32 and
mixed operators with their stress-rate errors over 1000 vectors) and
its hydraulically connected variant as sub-objects.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simulated yrs/wall-hr + H-matvec GB/s (%HBM peak), mixed+learned vs plain FP64"
# HBM bytes per launch of the dominant kernel, from two rocprofv3 PMC runs
# (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a 1 GiB copy kernel
# (FETCH x 2.000, WRITE x 1.000); summaries and commands in
# profiles/pmc_traffic_network80k_fp64.md, profiles/pmc_traffic_ensemble64_split.md and
# profiles/pmc_traffic_ensemble64_mixed.md.
# Keyed by (workload, seed, phase 1 fused with the near-field panels).
MEASURED_TRAFFIC = {
    ("network800_80k", 2024, False): 2.4846e9,
    ("ensemble64_1374k_split_learned", 2024, False): 17.550e9,
    ("ensemble64_1374k_mixed_learned", 2024, False): 21.246e9,
    ("ensemble64_1374k_mixed_learned", 2024, True): 20.702e9,
}
YEAR = 365.25 * 86400.0
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 peak (vector and matrix cores alike)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, seed):
    from demystification_dfn_amd import _abi, problems
    if name == "network80k":
        p = problems.network_80k(seed=seed)
        desc = dict(workload="network800_80k", n_faults=800, box_m=7000, length_law="p(L)~L^-0.1 on [50 m, 2 km]",
                    element_m=10, fracture_mode="II", kernel="hmatrix", eta=2.0, eps_aca=1e-6, leaf=64,
                    friction="aging, a~U[0.008,0.012], b=0.01 (70%) / 0.006, Dc=1 mm, sigma_n=-30 MPa",
                    pulse="10 MPa on one seeded element", seed=seed)
    elif name == "network80k_connected":
        p = problems.network_80k(seed=seed, snap=True, name=f"network800_80k_connected_seed{seed}")
        desc = dict(workload="network800_80k_connected", n_faults=800, box_m=7000,
                    length_law="p(L)~L^-0.1 on [50 m, 2 km]", element_m=10, fracture_mode="II", kernel="hmatrix",
                    eta=2.0, eps_aca=1e-6, leaf=64, crossings=p.meta["snapped_crossings"],
                    note="the 80k network with every fault intersection snapped to a shared node (hydraulic "
                         "crossings)", seed=seed)
    elif name == "single_fault":
        p = problems.single_fault(n=4000, length=40e3, static_kernel=_abi.KERNEL_DIRECT, tol=1e-8)
        desc = dict(workload="single_fault_4000", fracture_mode="III", kernel="direct", length_m=40e3,
                    injection="Q=1.25e-6 for 100 days", tol=1e-8)
    else:
        raise SystemExit(f"unknown workload {name}")
    desc["N"] = p.N
    return p, desc


def cpu_info():
    """The host cores this job may use: OMP_NUM_THREADS when the host sets it
    (the cores allotted per GPU), else the cores of this process's affinity mask;
    with the CPU model and the logical CPU count for the record."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(omp) if omp.isdigit() and int(omp) > 0 else affinity
    return dict(model=model, logical_cpus=os.cpu_count(), affinity_cpus=affinity, omp_num_threads=omp or None,
                threads=threads)


def cpu_baseline_reference(state, p, seconds):
    """The reference's own Fortran modules (oracle/_ref, built with the reference's
    -O1 serial flags) from `state` = (Y, P, time, t_try), timed for ~`seconds`."""
    from oracle.oracle import RefSimulation
    r = RefSimulation(p)
    Y, P, t, t_try = state
    r.set_state(Y, P, t)
    r.set_t_try(t_try)
    steps = 0
    w0 = time.perf_counter()
    while time.perf_counter() - w0 < seconds:
        steps += r.solve(1)
    wall = time.perf_counter() - w0
    sim_dt = r.get_state()[3] - t
    return dict(value=(sim_dt / YEAR) / (wall / 3600.0), unit="simulated yr / wall hr", cores=1, kind="reference",
                steps=steps, wall_s=wall, sim_s=sim_dt, ms_per_step=1e3 * wall / max(steps, 1),
                mean_dt_s=sim_dt / max(steps, 1),
                sample=f"{steps} accepted steps of the reference Fortran (directIII, friction law, compute_P, rkqs; "
                       f"-O1, one thread) from the GPU's state at the start of its timed window ({wall:.1f} s wall, "
                       f"{sim_dt:.3e} s simulated)")


def cpu_baseline(sim, state, p, seconds, threads, t_stop=None):
    """The oracle (plain-C restatement) from `state` = (Y, P, time, t_try) on the
    same operator block lists as the device handle `sim`, timed for ~`seconds` of
    wall time (or until t_stop)."""
    from oracle.oracle import OracleSimulation, RefSimulation
    if p.static_kernel == 0 and RefSimulation.available():
        return cpu_baseline_reference(state, p, seconds)
    o = OracleSimulation(p, threads=threads)
    if p.static_kernel == 1:
        for op in ((0, 1) if p.fracture_mode == 2 else (0,)):
            o.hmat_import(op, sim.hmat_export(op))
    Y, P, t, t_try = state
    o.set_state(Y, P, t)
    o.set_t_try(t_try)
    t_start = o.stats()["time"]
    steps = 0
    w0 = time.perf_counter()
    while time.perf_counter() - w0 < seconds and (t_stop is None or o.stats()["time"] < t_stop):
        o.solve(1)
        steps += 1
    wall = time.perf_counter() - w0
    sim_dt = o.stats()["time"] - t_start
    o.close()
    return dict(value=(sim_dt / YEAR) / (wall / 3600.0), unit="simulated yr / wall hr", cores=threads,
                kind="port", steps=steps, wall_s=wall, sim_s=sim_dt, ms_per_step=1e3 * wall / max(steps, 1),
                mean_dt_s=sim_dt / max(steps, 1),
                sample=f"{steps} accepted steps from the GPU's state at the start of its timed window "
                       f"({wall:.1f} s wall, {sim_dt:.3e} s simulated, OpenMP matvec on {threads} threads)")


def device_state(sim):
    Y, P, _, t = sim.get_state()
    return (Y, P, t, sim.stats()["t_try"])


def mixed_precision(sim, p, args, kb64):
    """Reduced operator storage on a single network: FP32 (per-block power-of-two
    scales, FP32 accumulation, compensated near-field sums), the same storage read
    into FP64 products and sums, and FP64 near-field blocks beside FP32 low-rank
    factors read into FP64 sums (the mixed mode), and the mixed mode with each
    near-field entry split into an FP32 value and an 8-bit residual (5 bytes), each
    against the FP64 operator over
    random log-uniform slip rates in [1e-12, 1] m/s (norm-wise and on the rows above
    1e-3 of the largest stress rate); then FP32 solver steps continued from the
    FP64 run's state.  State, RHS, CG and error norm stay FP64 throughout."""
    from demystification_dfn_amd import _abi
    rng = np.random.default_rng(7)
    errs = {_abi.PREC_FP32: ([], []), _abi.PREC_MIXED: ([], []), _abi.PREC_FP32_STORE: ([], []),
            _abi.PREC_MIXED_SPLIT: ([], [])}
    # the FP64 stress rates of every vector in one multi-RHS application
    X = 10.0 ** rng.uniform(-12.0, 0.0, (args.mixed_vectors, p.N))
    w0 = time.perf_counter()
    Y64 = sim.kernel_multi(X)
    multi_wall = time.perf_counter() - w0
    multi = multi_rhs_bench(sim, X, kb64)
    multi["wall_s_with_host_copies"] = multi_wall
    log(f"[bench] FP64 multi-RHS over {multi['vectors']} vectors: {multi['ms']:.1f} ms, phase 2 at "
        f"{multi['panel_TFLOP_s']:.1f} TFLOP/s")
    # one precision switch each (a switch rebuilds the reduced-precision layout)
    for prec in (_abi.PREC_FP32, _abi.PREC_FP32_STORE, _abi.PREC_MIXED, _abi.PREC_MIXED_SPLIT):
        sim.set_precision(prec)
        log(f"[bench] stress-rate error over {args.mixed_vectors} vectors, precision {prec}")
        for r in range(args.mixed_vectors):
            a = tuple(y[r] for y in Y64 if y is not None)
            b = sim.kernel(X[r])
            for y64, y in zip(a, b):
                errs[prec][0].append(np.linalg.norm(y - y64) / np.linalg.norm(y64))
                big = np.abs(y64) > 1e-3 * np.abs(y64).max()
                errs[prec][1].append(np.max(np.abs(y[big] - y64[big]) / np.abs(y64[big])))

    def stats(prec):
        e, g = errs[prec]
        return dict(vectors=args.mixed_vectors, norm_max=float(np.max(e)), norm_median=float(np.median(e)),
                    significant_rows_max=float(np.max(g)), significant_rows_median=float(np.median(g)))
    sim.set_precision(_abi.PREC_MIXED)
    kbm = sim.bench_kernel(reps=20)
    sim.set_precision(_abi.PREC_MIXED_SPLIT)
    kbs = sim.bench_kernel(reps=20)
    sim.set_precision(_abi.PREC_FP32)
    tune32 = sim.tuning()
    kb = sim.bench_kernel(reps=20)
    s0 = sim.stats()
    w0 = time.perf_counter()
    sim.solve(args.mixed_steps)
    wall = time.perf_counter() - w0
    s1 = sim.stats()
    sim.set_precision(_abi.PREC_FP64)
    sim_dt = s1["time"] - s0["time"]
    return dict(value=(sim_dt / YEAR) / (wall / 3600.0), unit="simulated yr / wall hr", steps=args.mixed_steps,
                ms_per_step=1e3 * wall / args.mixed_steps, mean_accepted_dt_s=sim_dt / args.mixed_steps,
                hmatvec_ms=kb["ms_kernel"], hmatvec_GB_s=kb["alg_bytes"] / (kb["ms_kernel"] * 1e-3) / 1e9,
                hmatvec_speedup_vs_fp64=kb64["ms_kernel"] / kb["ms_kernel"],
                panel_GB_s=kb["dominant_bytes"] / (kb["ms_dominant"] * 1e-3) / 1e9,
                stress_rate_rel_err=stats(_abi.PREC_FP32),
                fp64_near_field=dict(stress_rate_rel_err=stats(_abi.PREC_MIXED), hmatvec_ms=kbm["ms_kernel"],
                                     alg_bytes=kbm["alg_bytes"],
                                     hmatvec_speedup_vs_fp64=kb64["ms_kernel"] / kbm["ms_kernel"],
                                     note="FP64 near-field blocks beside FP32 low-rank factors read into FP64 "
                                          "products and sums (DFN_PREC_MIXED)"),
                split_near_field=dict(stress_rate_rel_err=stats(_abi.PREC_MIXED_SPLIT), hmatvec_ms=kbs["ms_kernel"],
                                      alg_bytes=kbs["alg_bytes"],
                                      hmatvec_speedup_vs_fp64=kb64["ms_kernel"] / kbs["ms_kernel"],
                                      note="DFN_PREC_MIXED with each near-field entry stored as its FP32 value and "
                                           "an 8-bit residual times a scale per panel task "
                                           "(DFN_PREC_MIXED_SPLIT)"),
                fp32_storage_fp64_sums=dict(stress_rate_rel_err=stats(_abi.PREC_FP32_STORE),
                                            note="the FP32 storage read into FP64 products and sums "
                                                 "(DFN_PREC_FP32_STORE): the error of the storage alone"),
                autotune=tune32, fp64_multi_rhs=multi,
                note="continues from the FP64 run's final state; operators FP32, state FP64")


def multi_rhs_bench(sim, X, kb64):
    """The FP64 H-matrix applied to the rows of X in one multi-RHS call
    (dfn_kernel_multi, phase 2 on the matrix cores), timed on the device with the
    vectors resident, against as many one-vector applications."""
    b = sim.bench_kernel_multi(X[:1024], reps=3)
    n = min(len(X), 1024)
    tflops = b["panel_flops"] / (b["ms_panel"] * 1e-3) / 1e12
    return dict(vectors=n, ms=b["ms_apply"], ms_per_vector=b["ms_apply"] / n,
                one_vector_ms=kb64["ms_kernel"], speedup_vs_one_vector_calls=kb64["ms_kernel"] * n / b["ms_apply"],
                panel_ms=b["ms_panel"], panel_TFLOP_s=tflops, panel_frac_fp64_peak=tflops / FP64_PEAK_TFLOPS,
                alg_GB_s=b["alg_bytes"] / (b["ms_apply"] * 1e-3) / 1e9,
                note="phase-2 flops = 2 x panel entries (dense and U) x vectors rounded up to 64; peak = FP64 matrix peak "
                     "(equal to the FP64 vector peak on MI355X)")


def learned_step(args, probs):
    """Step-size control on the ensemble: pairs recorded from an rkqs-controlled
    run train the proposal network on the device; then the same networks, from
    the same initial state, are stepped to a common final time under rkqs, the
    PI controller and the learned controller.  Reports steps, rejected-attempt
    fraction and aggregate simulated time per wall hour for each."""
    from demystification_dfn_amd import _abi
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceMlp
    rec = DeviceEnsemble(probs)
    rec.initiate_hmat()
    rec.set_recording(True)
    rec.solve(args.learned_steps)
    X, T = rec.take_records()
    t_end = min(rec.stats(i)["time"] for i in range(len(probs)))
    rec.close()
    mlp = DeviceMlp()
    out = dict(pairs=int(T.shape[0]), t_end_s=t_end)
    if T.shape[0] >= 16:
        norm = np.concatenate([X.mean(0), X.std(0) + 1e-3, [T.mean()], [T.std() + 1e-3]])
        p0, _ = mlp.get_params()
        mlp.set_params(p0, norm)
        l0, _ = mlp.loss_grad(X, T)
        w0 = time.perf_counter()
        lf = mlp.train(X, T, epochs=args.learned_epochs, batch=min(256, T.shape[0]), lr=1e-3, seed=1)
        out["training"] = dict(epochs=args.learned_epochs, loss_before=l0, loss_after=lf,
                               seconds=time.perf_counter() - w0)
    for key, mode in (("rkqs", _abi.STEP_RKQS), ("pi", _abi.STEP_PI), ("learned", _abi.STEP_LEARNED)):
        if mode == _abi.STEP_LEARNED and "training" not in out:
            continue
        ens = DeviceEnsemble(probs)
        ens.initiate_hmat()
        ens.set_controller(mode, mlp if mode == _abi.STEP_LEARNED else None)
        w0 = time.perf_counter()
        ens.solve_until(t_end, max_steps=100000)
        wall = time.perf_counter() - w0
        acc = sum(ens.stats(i)["accepted_steps"] for i in range(len(probs)))
        rej = sum(ens.stats(i)["rejected_attempts"] for i in range(len(probs)))
        out[key] = dict(value=(len(probs) * t_end / YEAR) / (wall / 3600.0), unit="simulated network-yr / wall hr",
                        accepted_steps=acc, rejected_attempts=rej, rejected_fraction=rej / max(acc + rej, 1),
                        wall_s=wall)
        ens.close()
    out["note"] = ("pairs exclude steps whose successor the one-day cap set; all controllers start from t=0 and "
                   "land on t_end; every step is accepted or rejected by the Cash-Karp test")
    return out


def run_ensemble(args):
    """Ensemble workload: independent seeded networks advanced together by the
    ensemble engine (shared stage launches, per-network clocks and step sizes,
    HIP-graph replay).  Throughput is the aggregate simulated time over all
    networks per wall hour; the same networks run one handle at a time, and the
    ensemble with graphs off, are timed beside it."""
    from demystification_dfn_amd import _abi, problems
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceSimulation
    lo, hi = (int(v) for v in args.ensemble_faults.split(","))
    probs = problems.network_ensemble(args.networks, faults=(lo, hi), box=args.ensemble_box, seed=args.seed)
    Ns = [p.N for p in probs]
    desc = dict(workload=f"ensemble{len(probs)}_{sum(Ns) // 1000}k", networks=len(probs),
                faults_per_network=f"{lo}-{hi}", box_m=args.ensemble_box,
                length_law="p(L)~L^-0.1 on [50 m, 2 km]", element_m=10, fracture_mode="II",
                kernel="hmatrix", eta=2.0, eps_aca=1e-6, leaf=64, elements=sum(Ns), min_elements=min(Ns),
                max_elements=max(Ns), seed=args.seed,
                friction="per network: a~U[0.008,0.012] per fault, b=0.01 (70%) / 0.006, Dc=1 mm, sigma_n=-30 MPa")
    log(f"[bench] ensemble of {len(probs)} networks, {sum(Ns)} elements (max {max(Ns)})")
    ens = DeviceEnsemble(probs)
    if args.no_graph:
        ens.set_graph(False)
    t0 = time.perf_counter()
    ens.initiate_hmat()
    setup_s = time.perf_counter() - t0
    info = ens.info()
    log(f"[bench] ensemble setup {setup_s:.1f} s {info}")

    def timed(steps):
        t_before = [ens.stats(i)["time"] for i in range(len(probs))]
        w0 = time.perf_counter()
        ens.solve(steps)
        wall = time.perf_counter() - w0
        dt = sum(ens.stats(i)["time"] - t_before[i] for i in range(len(probs)))
        return dt, wall

    ens.solve(args.warmup)
    dt, wall = timed(args.steps)
    value = (dt / YEAR) / (wall / 3600.0)
    log(f"[bench] ensemble {args.steps} steps in {wall:.2f} s")
    ens.set_graph(False)
    dt_e, wall_e = timed(max(args.steps // 3, 1))
    ens.set_graph(not args.no_graph)
    ens.set_precision(_abi.PREC_FP32)
    dt32, wall32 = timed(args.steps)
    ens.set_precision(_abi.PREC_FP64)
    rej = sum(ens.stats(i)["rejected_attempts"] for i in range(len(probs)))
    acc = sum(ens.stats(i)["accepted_steps"] for i in range(len(probs)))
    # the same networks one handle at a time (a sample of them, same step count)
    nsamp = min(len(probs), args.ensemble_single_sample)
    dt_s, wall_s = 0.0, 0.0
    for p in probs[:nsamp]:
        d = DeviceSimulation(p)
        d.initiate_hmat()
        d.solve(args.warmup)
        s0 = d.stats()["time"]
        w0 = time.perf_counter()
        d.solve(args.steps)
        wall_s += time.perf_counter() - w0
        dt_s += d.stats()["time"] - s0
        d.close()
    single = (dt_s / YEAR) / (wall_s / 3600.0) if wall_s > 0 else None
    out = dict(
        metric="simulated network-yr / wall hr, " + str(len(probs)) + "-network ensemble", value=value,
        unit="simulated network-yr / wall hr", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
        data="synthetic (seeded generator)", config=desc,
        ensemble=dict(setup_s=setup_s, S_entries=info["S_entries"], alg_bytes_per_stress_eval=info["alg_bytes"],
                      accepted_steps=acc, rejected_attempts=rej,
                      eager_no_graph=dict(value=(dt_e / YEAR) / (wall_e / 3600.0),
                                          ms_per_step=1e3 * wall_e / max(args.steps // 3, 1)),
                      fp32_operators=dict(value=(dt32 / YEAR) / (wall32 / 3600.0), ms_per_step=1e3 * wall32 / args.steps),
                      one_handle_at_a_time=dict(value=single, networks_sampled=nsamp,
                                                note="same networks, one dfn_handle each, run in turn"),
                      speedup_vs_one_at_a_time=value / single if single else None),
        roofline=None, cpu_baseline=None)
    if args.learned_steps > 0:
        log("[bench] learned step-size control")
        out["step_control"] = learned_step(args, probs)
    print(json.dumps(out), flush=True)


def run_network(workload, args, mixed=True, cpu_seconds=None):
    """One network on one dfn_handle: warmup, `steps` timed solver steps, the
    dominant kernel timed live over further steps, the isolated matvec, FP32
    operator storage and the CPU baseline from the state at the start of the
    timed window.  Returns the result dict."""
    from demystification_dfn_amd.device import DeviceSimulation
    p, desc = make_workload(workload, args.seed)
    log(f"[bench] workload {desc['workload']} N={p.N}")
    sim = DeviceSimulation(p)
    t0 = time.perf_counter()
    sim.initiate_hmat()
    setup_s = time.perf_counter() - t0
    hm = {}
    if p.static_kernel == 1:
        ops = (0, 1) if p.fracture_mode == 2 else (0,)
        st = [sim.hmat_stats(op) for op in ops]
        hm = dict(S_entries=int(sum(s["S_dense"] + s["S_lowrank_U"] + s["S_lowrank_V"] for s in st)),
                  max_rank=int(max(s["max_rank"] for s in st)),
                  lowrank_blocks=int(sum(s["nb_lowrank"] for s in st)),
                  dense_blocks=int(sum(s["nb_dense"] for s in st)), autotune=sim.tuning())
    log(f"[bench] H-matrix setup {setup_s:.1f} s {hm}")
    sim.solve(args.warmup)
    s0 = sim.stats()
    state0 = device_state(sim)  # the CPU baseline starts where the timed window starts
    w0 = time.perf_counter()
    sim.solve(args.steps)
    wall = time.perf_counter() - w0
    s1 = sim.stats()
    sim_dt = s1["time"] - s0["time"]
    value = (sim_dt / YEAR) / (wall / 3600.0)
    log(f"[bench] {args.steps} steps in {wall:.2f} s, simulated {sim_dt:.3e} s")
    # the dominant kernel of every stress evaluation, timed with HIP events on the
    # library's stream (dfn_kernel_timing) over further steps of the same run: the
    # event records cost host time per launch, so they stay out of the timed steps
    n_extra = min(args.steps, 50)
    sim.kernel_timing(True)
    sim.solve(n_extra)
    ms_live, n_live = sim.kernel_timing_read()
    sim.kernel_timing(False)
    log(f"[bench] dominant kernel {ms_live * 1e3:.1f} us mean over {n_live} launches in {n_extra} more steps")
    kb = sim.bench_kernel(reps=20)
    dominant = "panel_kernel_f64 (H-matvec phase 2)" if p.static_kernel == 1 else (
        "direct3_tile_kernel + reduce (dense dislocation kernel evaluated on the fly, 11 flops per "
        "element-node pair)" if p.fracture_mode == 3 else
        "direct2_tile_kernel + reduces (dense dislocation kernel evaluated on the fly, 40 flops per pair)")
    ms_dom = ms_live if n_live > 0 else kb["ms_dominant"]
    timing = dict(launches_timed=n_live, ms_per_launch_isolated=kb["ms_dominant"],
                  timing=f"mean over the launches of {n_extra} solver steps after the timed ones, HIP events "
                         f"on the library's stream")
    achieved = kb["dominant_bytes"] / (ms_dom * 1e-3) / 1e9
    whole = kb["alg_bytes"] / (kb["ms_kernel"] * 1e-3) / 1e9
    if p.static_kernel == 1:
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=MEASURED_TRAFFIC.get((desc["workload"], args.seed, False)), kernel=dominant,
                    bytes_per_launch=kb["dominant_bytes"], ms_per_launch=ms_dom, **timing)
        matvec = dict(ms=kb["ms_kernel"], alg_bytes=kb["alg_bytes"], GB_s=whole, frac_hbm=whole / HBM_PEAK_GBS, **hm)
    else:
        tflops = kb["dominant_bytes"] / (ms_dom * 1e-3) / 1e12
        roof = dict(bound="valu-fp64", achieved=tflops, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=tflops / FP64_PEAK_TFLOPS, traffic=None, kernel=dominant,
                    flops_per_launch=kb["dominant_bytes"], ms_per_launch=ms_dom, **timing)
        matvec = dict(ms=kb["ms_kernel"], flops=kb["alg_bytes"], TFLOP_s=tflops)
    out = dict(
        metric=METRIC, value=value, unit="simulated yr / wall hr", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
        data="synthetic (seeded generator)", config=desc,
        roofline=roof,
        hmatvec=matvec,
        run=dict(accepted=s1["accepted_steps"] - s0["accepted_steps"],
                 rejected=s1["rejected_attempts"] - s0["rejected_attempts"],
                 cg_iterations=s1["cg_iterations"] - s0["cg_iterations"], window_start_s=s0["time"],
                 sim_time_s=s1["time"], mean_accepted_dt_s=sim_dt / args.steps, max_V=s1["max_V"],
                 setup_s=setup_s),
    )
    if mixed and args.mixed_steps > 0 and p.static_kernel == 1:
        out["mixed"] = mixed_precision(sim, p, args, kb)
    cpu_seconds = args.cpu_seconds if cpu_seconds is None else cpu_seconds
    if not args.no_cpu and cpu_seconds > 0:
        ci = cpu_info()
        log(f"[bench] cpu baseline on {ci['threads']} threads")
        cb = cpu_baseline(sim, state0, p, cpu_seconds, ci["threads"])
        cb.update(cpu=ci, gpu_ms_per_step=out["ms_per_step"], gpu_mean_dt_s=sim_dt / args.steps,
                  gpu_vs_cpu_sim_rate=value / cb["value"] if cb["value"] > 0 else None,
                  gpu_vs_cpu_per_step=cb["ms_per_step"] / out["ms_per_step"])
        out["cpu_baseline"] = cb
        if cb["kind"] == "port" and ci["threads"] > 1:
            log("[bench] cpu baseline on 1 thread")  # the reference as built runs on one thread
            out["cpu_baseline_1thread"] = cpu_baseline(sim, state0, p, cpu_seconds / 2, 1)
    else:
        out["cpu_baseline"] = None
    sim.close()
    return out


def run_headline(args):
    """Configs 4 and 5: the 64-network ensemble with reduced-precision operator
    storage read into FP64 products and sums (FP32 low-rank factors; near-field
    entries as an FP32 value and an 8-bit residual, DFN_PREC_MIXED_SPLIT, or in FP64,
    DFN_PREC_MIXED) and the learned step-size proposals, against the plain FP64 path
    (one dfn_handle per network, FP64 operators, rkqs) over the same simulated
    interval of every network; then, continuing, the ensemble with rkqs in the same
    precision and in FP64."""
    from demystification_dfn_amd import _abi, problems
    from demystification_dfn_amd import train_step_network as tsn
    from demystification_dfn_amd.device import DeviceEnsemble, DeviceMlp, DeviceSimulation
    lo, hi = (int(v) for v in args.ensemble_faults.split(","))
    probs = problems.network_ensemble(args.networks, faults=(lo, hi), box=args.ensemble_box, seed=args.seed)
    B = len(probs)
    Ns = np.array([p.N for p in probs])
    split = args.headline_precision == "split"
    prec = _abi.PREC_MIXED_SPLIT if split else _abi.PREC_MIXED
    desc = dict(workload=f"ensemble{B}_{Ns.sum() // 1000}k_{args.headline_precision}_learned", networks=B,
                faults_per_network=f"{lo}-{hi}", box_m=args.ensemble_box,
                length_law="p(L)~L^-0.1 on [50 m, 2 km]", element_m=10, fracture_mode="II",
                kernel="hmatrix", eta=2.0, eps_aca=1e-6, leaf=64, elements=int(Ns.sum()), min_elements=int(Ns.min()),
                max_elements=int(Ns.max()), seed=args.seed,
                friction="per network: a~U[0.008,0.012] per fault, b=0.01 (70%) / 0.006, Dc=1 mm, sigma_n=-30 MPa",
                pulse="10 MPa on one seeded element per network (fluid diffusion on)",
                precision=("DFN_PREC_MIXED_SPLIT: near-field entries as their FP32 value plus an 8-bit residual "
                           "times a scale per panel task (5 bytes; at least 24, for a task's largest entries "
                           "about 31 significant bits); " if split else
                           "DFN_PREC_MIXED: FP64 near-field blocks; ") +
                          "FP32 low-rank factors (power-of-two block scales); all read into FP64 products and sums; "
                          "FP64 state, RHS, CG and error norm",
                controller="learned proposals (data/step_network.npz) with the PI controller as fallback; the "
                           "Cash-Karp error test accepts or rejects every step")
    log(f"[bench] headline: ensemble of {B} networks, {Ns.sum()} elements (max {Ns.max()})")
    params, norm, meta = tsn.load()
    mlp = DeviceMlp()
    mlp.set_params(params, norm)
    ens = DeviceEnsemble(probs)
    t0 = time.perf_counter()
    ens.initiate_hmat()
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens.set_precision(prec)
    layout32_s = time.perf_counter() - t0
    log(f"[bench] ensemble setup {setup_s:.1f} s, {args.headline_precision} layout {layout32_s:.1f} s {ens.info()}")
    ens.set_controller(_abi.STEP_LEARNED, mlp)

    def window(steps):
        ta = np.array([ens.stats(i)["time"] for i in range(B)])
        acc0 = sum(ens.stats(i)["accepted_steps"] for i in range(B))
        rej0 = sum(ens.stats(i)["rejected_attempts"] for i in range(B))
        w0 = time.perf_counter()
        ens.solve(steps)
        wall = time.perf_counter() - w0
        tb = np.array([ens.stats(i)["time"] for i in range(B)])
        acc = sum(ens.stats(i)["accepted_steps"] for i in range(B)) - acc0
        rej = sum(ens.stats(i)["rejected_attempts"] for i in range(B)) - rej0
        sim = float((tb - ta).sum())
        return ta, tb, wall, dict(value=(sim / YEAR) / (wall / 3600.0), unit="simulated network-yr / wall hr",
                                  steps=steps, ms_per_step=1e3 * wall / steps, mean_accepted_dt_s=sim / max(acc, 1),
                                  accepted=acc, rejected=rej, rejected_fraction=rej / max(acc + rej, 1),
                                  window_start_s_median=float(np.median(ta)))

    head_dtype = ("f32 + 8-bit residual near-field / f32 low-rank storage, f64 arithmetic and state" if split else
                  "f64 near-field / f32 low-rank storage, f64 arithmetic and state")
    ens.solve(args.warmup)
    tA, tB, wall, head = window(args.steps)
    value = head["value"]
    log(f"[bench] {args.headline_precision}+learned: {args.steps} steps in {wall:.2f} s "
        f"({head['ms_per_step']:.2f} ms/step)")
    # the dominant kernel, HIP events on the ensemble's stream over further steps in
    # which every network takes part
    n_extra = min(args.steps, 10)
    ens.kernel_timing(True)
    ens.solve(n_extra)
    ms_dom, n_dom, bytes_dom = ens.kernel_timing_read()
    ens.kernel_timing(False)
    achieved = bytes_dom / (ms_dom * 1e-3) / 1e9 if n_dom else None
    fused = bool(ens.info()["fused_near"])
    dom_name = ("vtx_near_kernel: phase 1 over the FP32 low-rank V factors (FP64 sums) and the FP64 near-field "
                "panels in one launch; the whole ensemble's tasks" if fused else
                "phase-2 panels (panel_kernel_split on the near-field set stored as FP32 + 8-bit residual, then "
                "panel_kernel_f32w on the FP32 low-rank set; the whole ensemble's panels)" if split else
                "phase-2 panels (panel_kernel_f64 on the FP64 near-field set, then panel_kernel_f32w on the FP32 "
                "low-rank set; the whole ensemble's panels)")
    log(f"[bench] dominant kernel ({'phase 1 + near-field' if fused else 'phase-2 panels'}): {ms_dom:.3f} ms mean "
        f"over {n_dom} launches")
    roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=achieved / HBM_PEAK_GBS if achieved else None,
                traffic=MEASURED_TRAFFIC.get((desc["workload"], args.seed, fused)),
                kernel=dom_name, bytes_per_launch=bytes_dom, ms_per_launch=ms_dom, launches_timed=n_dom,
                timing=f"mean over the dominant launches of {n_extra} ensemble steps after the timed ones in which "
                       f"every network took part, HIP events on the ensemble's stream")
    info32 = ens.info()
    # the same ensemble, continuing: rkqs with the same operators, then with FP64 operators
    ens.set_controller(_abi.STEP_RKQS)
    _, _, _, rkqs32 = window(args.steps)
    ens.set_precision(_abi.PREC_FP64)
    info64 = ens.info()
    _, _, _, rkqs64 = window(args.steps)
    ens.close()
    log(f"[bench] continuing: {args.headline_precision} rkqs {rkqs32['ms_per_step']:.2f} ms/step, FP64 rkqs "
        f"{rkqs64['ms_per_step']:.2f}")
    # the plain FP64 path: every network on its own dfn_handle, from t = 0 to the start
    # of the ensemble's timed window (untimed), then timed over the same interval
    mlp.close()
    order = np.argsort(Ns)
    sample = [int(order[B // 4]), int(order[B // 2])] if not args.no_cpu and args.cpu_seconds > 0 else []
    if args.no_plain:  # profiling runs: the ensemble's kernels only
        return dict(metric=METRIC, value=value, unit="simulated network-yr / wall hr", n_gpus=1, steps=args.steps,
                    warmup=args.warmup, ms_per_step=head["ms_per_step"], higher_is_better=True, scaling="weak",
                    vs_baseline=None, dtype=head_dtype,
                    data="synthetic (seeded generator)",
                    config=desc, roofline=roof, window=head, cpu_baseline=None,
                    note="--no-plain: the plain FP64 path and the CPU baseline were not run")
    plain_wall, plain_steps, plain_rej, states, per_net = 0.0, 0, 0, {}, {}
    for i, p in enumerate(probs):
        d = DeviceSimulation(p)
        d.initiate_hmat()
        d.solve(10 ** 9, t_final=float(tA[i]))
        s0 = d.stats()
        if i in sample:
            states[i] = device_state(d)
        w0 = time.perf_counter()
        d.solve(10 ** 9, t_final=float(tB[i]))
        w = time.perf_counter() - w0
        s1 = d.stats()
        plain_wall += w
        plain_steps += s1["accepted_steps"] - s0["accepted_steps"]
        plain_rej += s1["rejected_attempts"] - s0["rejected_attempts"]
        per_net[i] = (w, s1["accepted_steps"] - s0["accepted_steps"])
        if i in sample:
            states[i] = (states[i], d)  # keep the handle for the CPU baseline's block lists
        else:
            d.close()
        if i % 16 == 15:
            log(f"[bench] plain FP64 path: {i + 1}/{B} networks")
    sim_total = float((tB - tA).sum())
    plain_value = (sim_total / YEAR) / (plain_wall / 3600.0)
    ratio = value / plain_value
    log(f"[bench] plain FP64 path {plain_value:.1f} network-yr/hr; {args.headline_precision}+learned ensemble "
        f"{ratio:.2f}x")
    out = dict(
        metric=METRIC, value=value, unit="simulated network-yr / wall hr", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=head["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype=head_dtype, data="synthetic (seeded generator)",
        config=desc,
        roofline=roof,
        vs_plain_fp64=dict(plain_value=plain_value, unit="simulated network-yr / wall hr", ratio=ratio,
                           target_ratio=20.0, meets_target=bool(ratio >= 20.0), plain_wall_s=plain_wall,
                           plain_accepted=plain_steps, plain_rejected=plain_rej,
                           note="the same 64 networks, one dfn_handle each (FP64 operators, rkqs), each stepped "
                                "from t=0 to the start of the ensemble's timed window (untimed) and then timed to "
                                "the window's end: the same simulated interval as the ensemble's, per network"),
        window=head,
        ensemble=dict(setup_s=setup_s, reduced_layout_s=layout32_s, S_entries=info64["S_entries"],
                      alg_bytes_per_stress_eval_fp64=info64["alg_bytes"],
                      alg_bytes_per_stress_eval_reduced=info32["alg_bytes"],
                      continuing_reduced_rkqs=rkqs32, continuing_fp64_rkqs=rkqs64,
                      note="continuing windows follow the timed one in simulated time; rates over different "
                           "windows compare per step (ms_per_step) rather than per simulated year"),
        proposal_network=dict(pairs=meta["recording"]["pairs"], loss_initial=meta["training"]["loss_initial"],
                              loss_final=meta["training"]["loss_final"], features=meta["features"]),
    )
    # CPU baseline: the oracle on two sampled networks (25th and 50th percentile of
    # size) from the plain path's state at the start of the timed window
    if sample:
        ci = cpu_info()
        runs, gpu_w, gpu_sim = [], 0.0, 0.0
        for i in sample:
            st, d = states[i]
            log(f"[bench] cpu baseline: network {i} (N={probs[i].N}) on {ci['threads']} threads")
            r = cpu_baseline(d, st, probs[i], args.cpu_seconds / len(sample), ci["threads"], t_stop=float(tB[i]))
            d.close()
            r["network"], r["N"] = i, int(probs[i].N)
            # the plain GPU path over the same network's window, at the same start
            w, n = per_net[i]
            r["gpu_plain_ms_per_step"] = 1e3 * w / max(n, 1)
            runs.append(r)
            gpu_w += w
            gpu_sim += float(tB[i] - tA[i])
        cpu_sim = sum(r["sim_s"] for r in runs)
        cpu_wall = sum(r["wall_s"] for r in runs)
        cpu_rate = (cpu_sim / YEAR) / (cpu_wall / 3600.0)
        gpu_rate = (gpu_sim / YEAR) / (gpu_w / 3600.0)
        # over the whole ensemble, as the sampled networks' CPU / plain-GPU ratio
        est = cpu_rate * plain_value / gpu_rate
        out["cpu_baseline"] = dict(
            value=est, unit="simulated network-yr / wall hr", cores=ci["threads"], kind="port",
            sample=(f"the oracle (OpenMP matvec on {ci['threads']} threads) on networks {sample} (N = "
                    f"{[int(probs[i].N) for i in sample]}, the 25th and 50th size percentiles) from the GPU's state "
                    f"at the start of the timed window, {cpu_wall:.1f} s wall in all; value = their CPU rate scaled "
                    f"to the 64 networks by the plain GPU path's rate on them against its rate on all 64"),
            measured_on_sample=dict(value=cpu_rate, unit="simulated network-yr / wall hr", runs=runs),
            plain_gpu_on_sample=gpu_rate, plain_gpu_vs_cpu=gpu_rate / cpu_rate if cpu_rate > 0 else None, cpu=ci)
    else:
        out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="headline",
                    help="headline (the 64-network ensemble, mixed operators + learned step vs the plain FP64 path), "
                         "network80k, network80k_connected, single_fault, ensemble")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sub", action="store_true", help="headline: skip the 80k-network sub-runs")
    ap.add_argument("--no-plain", action="store_true",
                    help="headline: skip the plain FP64 path and the CPU baseline (profiling runs)")
    ap.add_argument("--mixed-steps", type=int, default=20,
                    help="steps of the FP32-operator (mixed precision) run on a single network; 0 skips it")
    ap.add_argument("--mixed-vectors", type=int, default=1000,
                    help="random log-uniform slip-rate vectors for the FP32-vs-FP64 stress-rate error")
    ap.add_argument("--networks", type=int, default=64, help="networks of the ensemble workloads")
    ap.add_argument("--ensemble-faults", default="50,400", help="fault-count range per network of the ensembles")
    ap.add_argument("--ensemble-box", type=float, default=7000.0, help="box size (m) of the ensemble's networks")
    ap.add_argument("--headline-precision", choices=("split", "mixed"), default="split",
                    help="operator storage of the headline ensemble: split (DFN_PREC_MIXED_SPLIT, near-field entries "
                         "as FP32 + 8-bit residual) or mixed (DFN_PREC_MIXED, FP64 near-field)")
    ap.add_argument("--ensemble-single-sample", type=int, default=16,
                    help="--workload ensemble: networks also run one handle at a time")
    ap.add_argument("--no-graph", action="store_true",
                    help="--workload ensemble: issue the launches one by one instead of replaying HIP graphs")
    ap.add_argument("--learned-steps", type=int, default=0,
                    help="--workload ensemble: steps of a recording run that trains the step-size network; 0 skips")
    ap.add_argument("--learned-epochs", type=int, default=30)
    ap.add_argument("--copy-probe-gb", type=float, default=2.0,
                    help="size of the copy kernel whose rate is reported as the achievable HBM bandwidth")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every workload here runs on a single GPU", file=sys.stderr)
        sys.exit(2)

    from demystification_dfn_amd.device import bench_stream, device_info
    log("[bench]", device_info())
    if args.workload == "ensemble":
        return run_ensemble(args)
    if args.workload != "headline":
        out = run_network(args.workload, args)
        out["copy_kernel_GB_s"] = bench_stream(int(args.copy_probe_gb * (1 << 30)), 20) if args.copy_probe_gb else None
        print(json.dumps(out), flush=True)
        return
    out = run_headline(args)
    if not args.no_sub:
        # the 800-fault 80k-element network (FP64, with its reduced-storage runs) and its connected variant
        out["network80k"] = run_network("network80k", args, mixed=True, cpu_seconds=args.cpu_seconds / 2)
        out["network80k_connected"] = run_network("network80k_connected", args, mixed=False, cpu_seconds=0)
    out["copy_kernel_GB_s"] = bench_stream(int(args.copy_probe_gb * (1 << 30)), 20) if args.copy_probe_gb else None
    print(json.dumps(out), flush=True)
    maps = os.environ.get("DFN_BENCH_MAPS")
    if maps:  # the process's library map at its end (attributes a fault in the exit handlers)
        with open("/proc/self/maps") as src, open(maps, "w") as dst:
            dst.write(src.read())


if __name__ == "__main__":
    main()
