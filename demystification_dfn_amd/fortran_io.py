"""Binary problem file read by the Fortran host driver (fortran/drive_solver.f90)."""
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


def read_result(path, N):
    v = np.fromfile(path, np.float64)
    return float(v[0]), v[1:1 + 3 * N], v[1 + 3 * N:1 + 4 * N]
