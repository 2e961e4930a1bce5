"""Binary problem files read by the Fortran host drivers (fortran/drive_solver.f90,
fortran/drive_ensemble.f90) and the result file of the ensemble driver."""
import numpy as np

from .problems import PER_ELEMENT


def write_problem(problem, path, nsteps):
    src = np.array(problem.sources, np.float64).reshape(-1, 5)
    with open(path, "wb") as f:
        np.array([problem.N, problem.nb_fault, problem.fracture_mode, problem.static_kernel, len(src), nsteps],
                 np.int64).tofile(f)
        problem.fault_nb_element.astype(np.int64).tofile(f)
        np.ascontiguousarray(problem.node, np.float64).tofile(f)
        np.stack([problem.arrays[k] for k in PER_ELEMENT], axis=1).astype(np.float64).T.copy().tofile(f)
        np.array([problem.mu, problem.cs, problem.cp, problem.Sigma11_dot, problem.Sigma12_dot, problem.Sigma22_dot,
                  problem.Sigma31_dot, problem.Sigma32_dot, problem.porosity, problem.fluid_comp, problem.rock_comp,
                  problem.dyn_viscosity, problem.tol_solver, problem.eta], np.float64).tofile(f)
        if len(src):
            src[:, 0].astype(np.int64).tofile(f)
            src[:, 1].astype(np.int64).tofile(f)
            src[:, 2].tofile(f)
            src[:, 3].tofile(f)
            src[:, 4].tofile(f)
        # H-matrix parameters and the coupling switch (read by the ensemble driver)
        np.array([problem.eps_aca, problem.nleaf, problem.max_rank, problem.permeability_coupling],
                 np.float64).tofile(f)


def write_weights(path, params, norm):
    """Raw FP32 parameters (DFN_MLP_PARAMS) then normalisation (18 values)."""
    np.concatenate([np.asarray(params, np.float32).ravel(), np.asarray(norm, np.float32).ravel()]).tofile(path)


def read_ensemble_result(path, Ns, n_params):
    """The ensemble driver's result file: per network (time, Y, P, accepted steps,
    rejected attempts) and the count of recorded pairs in FP64, then the pairs
    (X, T) in FP32; with a proposal network, its FP32 outputs on X, the FP64 loss
    after one training epoch and the trained FP32 parameters."""
    with open(path, "rb") as f:
        raw = f.read()
    pos = 0

    def take(dtype, n):
        nonlocal pos
        a = np.frombuffer(raw, dtype, n, pos)
        pos += a.nbytes
        return a

    assert int(take(np.float64, 1)[0]) == len(Ns)
    out = dict(nets=[])
    for N in Ns:
        t = float(take(np.float64, 1)[0])
        Y, P = take(np.float64, 3 * N), take(np.float64, N)
        acc, rej = take(np.float64, 2)
        out["nets"].append(dict(time=t, Y=Y, P=P, accepted=int(acc), rejected=int(rej)))
    cnt = int(take(np.float64, 1)[0])
    out["X"] = take(np.float32, 8 * cnt).reshape(cnt, 8)
    out["T"] = take(np.float32, cnt)
    if pos < len(raw):
        out["Y"] = take(np.float32, cnt)
        out["loss"] = float(take(np.float64, 1)[0])
        out["params"] = take(np.float32, n_params)
    assert pos == len(raw), (pos, len(raw))
    return out


def read_result(path, N):
    """The single-problem driver's result file: time, Y (3N), P (N)."""
    v = np.fromfile(path, np.float64)
    return float(v[0]), v[1:1 + 3 * N], v[1 + 3 * N:1 + 4 * N]
