"""Writes the golden fixtures from the reference's own Fortran modules.

Run in a container where /root/reference exists and oracle/_ref/libdfn_ref.so has
been built (make -C oracle -f Makefile.ref).  Produces tests/golden/*.npz: inputs
and the reference's outputs for the kernel, the friction law, compute_P, rkqs,
20-30 solver steps, the cluster / block cluster trees, the kernel-entry functions
of the H-matrix (eval_kernel_hmatrix.f90) and the permeability-coupled pressure
solve (permeability_coupling = 1).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import oracle as O  # noqa: E402
import cases  # noqa: E402


def run_case(prob, xs, dts, nsteps, seed):
    r = O.RefSimulation(prob)
    rng = np.random.default_rng(seed)
    V = rng.random(prob.N) * 1e-9
    out = dict(kernel_in=V)
    out["kernel_T"], out["kernel_N"] = r.kernel(V)
    Y0 = prob.initial_Y()
    out["Y0"] = Y0
    out["rhs_x"] = np.array(xs)
    out["rhs"] = np.stack([r.friction_law(x, Y0) for x in xs])
    out["P_dt"] = np.array(dts)
    res = [r.compute_P(dt) for dt in dts]
    out["P_temp"] = np.stack([a for a, _, _ in res])
    out["P_dot"] = np.stack([b for _, b, _ in res])
    out["P_ier"] = np.array([c for _, _, c in res])
    htry = 200.0
    yd = r.friction_law(htry, Y0)
    yscal = np.abs(Y0) + np.abs(yd * htry) + prob.tol_solver
    y1, x1, hdid, hnext = r.rkqs(Y0, yd, 0.0, htry, prob.tol_solver, yscal)
    out.update(rkqs_dydx=yd, rkqs_yscal=yscal, rkqs_y=y1, rkqs_x=x1, rkqs_hdid=hdid, rkqs_hnext=hnext)
    r2 = O.RefSimulation(prob)
    out["solve_steps"] = np.array(nsteps)
    r2.solve(nsteps)
    Y, P, Pd, t = r2.get_state()
    out.update(solve_Y=Y, solve_P=P, solve_time=t)
    return out


def kernel_entries(prob, seed, nsample):
    """Entries of eval_kernel_ModeIII (Mode III) or eval_kernel_ModeII_T / _N
    (Mode II) at the diagonal, both neighbour diagonals, 64 seeded entries in the
    row and in the column of every fake element and nsample seeded (receiver,
    source) pairs; the whole matrix when it has at most 2000 entries."""
    r = O.RefSimulation(prob)
    N = prob.N
    if N * N <= 2000:
        k, l = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
        k, l = k.ravel(), l.ravel()
    else:
        rng = np.random.default_rng(seed)
        fake = prob.fake_indices()
        ar = np.arange(N)
        fk = np.repeat(fake, 64)
        k = np.concatenate([ar, ar[:-1], ar[1:], fk, rng.integers(0, N, fk.size), rng.integers(0, N, nsample)])
        l = np.concatenate([ar, ar[1:], ar[:-1], rng.integers(0, N, fk.size), fk, rng.integers(0, N, nsample)])
    assert N < 32768
    out = {"k": k.astype(np.int16), "l": l.astype(np.int16)}
    for op in ((0,) if prob.fracture_mode == 3 else (0, 1)):
        out[f"op{op}"] = r.eval_matrix(op)[k, l]
    return out


def main():
    assert O.RefSimulation.available(), "build oracle/_ref first (make -C oracle -f Makefile.ref)"
    p1 = cases.single_fault_small()
    np.savez_compressed(os.path.join(HERE, "single_fault_300.npz"),
                        **run_case(p1, [1.0, 1000.0, 20000.0], [10.0, 5000.0, 86400.0], 20, 0))
    p2 = cases.crossing_network()
    np.savez_compressed(os.path.join(HERE, "crossing_network.npz"),
                        **run_case(p2, [1.0, 1000.0, 20000.0], [10.0, 5000.0, 86400.0], 30, 1))
    q = cases.tree_network()
    osim = O.OracleSimulation(q)
    g = O.geometry(osim)
    t = O.RefSimulation.trees(g["elem"], q.nleaf, q.eta)
    np.savez_compressed(os.path.join(HERE, "trees_network20.npz"), elem=g["elem"], **t)
    # the reference's own network input file, read with the library's HDF5 reader,
    # run by the reference's Fortran with the dense Mode II kernel
    cfg = "/root/reference/problems/network_rate_strengthening/config"
    if os.path.exists(cfg):
        from demystification_dfn_amd.config_io import read_config
        p3 = read_config(cfg)
        p3.static_kernel = 0
        out = cases.problem_arrays(p3)
        out.update(run_case(p3, [1.0, 1000.0, 20000.0], [10.0, 5000.0, 86400.0], 20, 2))
        np.savez_compressed(os.path.join(HERE, "network_rate_strengthening.npz"), **out)
    ent = {}
    for name, prob, seed in (("single_fault_300", p1, 3), ("crossing_network", p2, 4), ("network20", q, 5),
                             ("network20_mode3", cases.mode3_network(), 6)):
        for key, v in kernel_entries(prob, seed, 1500).items():
            ent[f"{name}__{key}"] = v
    if os.path.exists(cfg):
        for key, v in kernel_entries(cases.reference_network(), 7, 1500).items():
            ent[f"network_rate_strengthening__{key}"] = v
    np.savez_compressed(os.path.join(HERE, "kernel_entries.npz"), **ent)
    # permeability_coupling = 1: compute_P at three steps and 30 solver steps
    pc = cases.coupling_network()
    np.savez_compressed(os.path.join(HERE, "coupling_network.npz"),
                        **run_case(pc, [1.0, 1000.0, 20000.0], [10.0, 5000.0, 86400.0], 30, 8))
    print("written")


if __name__ == "__main__":
    main()
