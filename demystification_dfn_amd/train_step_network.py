"""Trains the step-size proposal network on the GPU and writes its weights.

Pairs (features, log10 of the accepted next step over rkqs's proposal) are
recorded from ensembles stepped by the reference controller (rkqs,
src/ode_solver/ode_bs.f90:165-169) with per-network tolerances between 1e-6 and
1e-10, so that most steps are set by the error test rather than by the one-day
cap (capped steps carry no information about the controller and are not
recorded).  The network (dfn_mlp_*: hand-written forward, backward and Adam
kernels) is trained on the recorded pairs and written, with its input and output
normalisation and the recording statistics, to data/step_network.npz.

    python -m demystification_dfn_amd.train_step_network [--pairs 1000000] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "step_network.npz")
TOLERANCES = (1e-6, 1e-7, 1e-8, 1e-9, 1e-10)


def recording_ensemble(n_networks, seed):
    """Seeded small networks (5-36 faults in 2.5 km boxes), tolerances cycling
    through TOLERANCES."""
    from . import problems
    probs = problems.network_ensemble(n_networks, faults=(5, 36), box=2500.0, seed=seed)
    for i, p in enumerate(probs):
        p.tol_solver = TOLERANCES[i % len(TOLERANCES)]
    return probs


def record(pairs, seed, networks, max_seconds, log):
    """At least `pairs` pairs (or what max_seconds of stepping gives) from fresh
    ensembles of `networks` networks."""
    from . import _abi
    from .device import DeviceEnsemble
    Xs, Ts, steps, t0 = [], [], 0, time.perf_counter()
    have, ensembles, stops = 0, 0, 0
    while have < pairs and time.perf_counter() - t0 < max_seconds:
        ens = DeviceEnsemble(recording_ensemble(networks, seed + ensembles))
        ensembles += 1
        ens.initiate_hmat()
        ens.set_recording(True)
        for _ in range(200):
            try:
                ens.solve(25)
                steps += 25
            except _abi.DfnError as e:
                # a network turned extensional (the reference stops there): keep the
                # pairs recorded so far and go on with the next ensemble
                if e.code != _abi.DFN_ERR_TENSILE:
                    raise
                stops += 1
                done = True
            else:
                done = False
            X, T = ens.take_records()
            Xs.append(X)
            Ts.append(T)
            have += len(T)
            if done or have >= pairs or time.perf_counter() - t0 >= max_seconds:
                break
        log(f"[train] {have} pairs after {steps} ensemble steps of {ensembles} ensembles "
            f"({time.perf_counter() - t0:.0f} s)")
        ens.close()
    X = np.concatenate(Xs)[:pairs]
    T = np.concatenate(Ts)[:pairs]
    return X, T, dict(pairs=int(len(T)), ensemble_steps=steps, networks=networks, ensembles=ensembles,
                      stopped_extensional=stops, record_seconds=time.perf_counter() - t0)


def train(X, T, epochs, batch, lr, seed):
    """Normalisation from the data, then Adam on the device; returns (mlp, info)."""
    from .device import DeviceMlp
    mlp = DeviceMlp()
    norm = np.concatenate([X.mean(0), X.std(0) + 1e-3, [T.mean()], [T.std() + 1e-3]]).astype(np.float32)
    p0, _ = mlp.get_params()
    mlp.set_params(p0, norm)
    loss0, _ = mlp.loss_grad(X[:65536], T[:65536])
    w0 = time.perf_counter()
    loss = mlp.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, dict(epochs=epochs, batch=batch, lr=lr, loss_initial=float(loss0), loss_final=float(loss),
                     train_seconds=time.perf_counter() - w0, target_mean=float(T.mean()), target_std=float(T.std()),
                     rejected_pairs_fraction=float(np.mean(T < -1e-6)))


def save(path, mlp, meta):
    params, norm = mlp.get_params()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=params.astype(np.float32), norm=norm.astype(np.float32), meta=json.dumps(meta))


def load(path=DEFAULT_OUT):
    """(params, norm, meta) of a trained network written by save()."""
    z = np.load(path, allow_pickle=False)
    return z["params"], z["norm"], json.loads(str(z["meta"]))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--networks", type=int, default=256)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--max-seconds", type=float, default=300.0, help="cap on the recording time")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, file=sys.stderr, flush=True)

    X, T, rec = record(a.pairs, a.seed, a.networks, a.max_seconds, log)
    log(f"[train] training on {len(T)} pairs")
    mlp, info = train(X, T, a.epochs, a.batch, a.lr, a.seed)
    meta = dict(recording=rec, training=info, tolerances=list(TOLERANCES),
                features=["max log10 V", "mean log10 V", "d(max log10 V)/dt", "log10 err n", "log10 err n-1",
                          "log10 err n-2", "log10 dt", "fraction V > 1e-3 m/s"],
                target="log10(accepted step n+1 / rkqs proposal after step n)")
    save(a.out, mlp, meta)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
