"""Writes the golden fixtures from the reference's own Fortran modules.

Run in a container where /root/reference exists and oracle/_ref/libdfn_ref.so has
been built (make -C oracle -f Makefile.ref).  Produces tests/golden/*.npz: inputs
and the reference's outputs for the kernel, the friction law, compute_P, rkqs,
20-30 solver steps and the cluster / block cluster trees.
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
    print("written")


if __name__ == "__main__":
    main()
