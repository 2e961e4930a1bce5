"""Problems the golden fixtures are generated on (shared by make_golden.py and the tests)."""
import os

import numpy as np

from demystification_dfn_amd import problems


def config1_single_fault():
    """Config 1: BP6-QD-A-like fault, 4000 elements over 40 km, Mode III, dense
    kernel, tol 1e-8, injection at the centre for 100 days."""
    return problems.single_fault(n=4000, length=40e3, tol=1e-8, name="config1_single_fault_4000")


def single_fault_small():
    """BP6-QD-A-like fault, 300 elements over 3 km, dense Mode III, tol 1e-8."""
    return problems.single_fault(n=300, length=3000.0, tol=1e-8, name="golden_single_fault_300")


def _seg(x0, y0, x1, y1, n):
    t = np.linspace(0, 1, n + 1)
    return np.stack([x0 + (x1 - x0) * t, y0 + (y1 - y0) * t], 1)


def crossing_network():
    """Three Mode II faults, two of them sharing the node (0,0) (one crossing), an
    injection source on fault 2, a 10 MPa pulse and graded permeability."""
    fs = []
    for nd, a in [(_seg(-50, 0, 50, 0, 10), 0.01), (_seg(0, -50, 0, 50, 10), 0.011),
                  (_seg(30, 40, 120, 60, 9), 0.009)]:
        n = nd.shape[0] - 1
        one = np.ones(n)
        fs.append(dict(node=nd, a=a * one, b=0.012 * one, Dc=1e-3 * one, f0=0.6 * one, V0=1e-6 * one,
                       sigmaN=-30e6 * one, V=1e-12 * one, theta=np.log(1e6) * one, P=0 * one,
                       permeability=1e-15 * one * (1 + np.arange(n) / 10), Snk=one, Lk=one, Tk=one))
    nels, node, arr = problems.concatenate_faults(fs)
    arr["P"][5] = 10e6
    return problems.Problem("golden_crossing_network", 2, 0, nels, node, arr, tol_solver=1e-6,
                            sources=[(2, 3, 1e-5, 0.0, 1e5)])


def coupling_network():
    """The crossing network with permeability_coupling = 1 (src/fluid_diffusion.f90:68-70,
    src/initialisation.f90:252-266): each compute_P sets the permeability from the
    effective normal stress, k = (k* - kmin) exp(-|sigmaN + P| / Snk) + kmin, with
    Snk = 20 MPa and kmin a tenth of the initial permeability on the real elements."""
    p = crossing_network()
    p.name = "golden_coupling_network"
    p.permeability_coupling = 1
    real = np.ones(p.N, bool)
    real[p.fake_indices()] = False
    p.arrays["kmin"] = np.where(real, 0.1 * p.arrays["permeability"], 0.0)
    p.arrays["Snk"] = np.where(real, 20e6, 1.0)
    return p


def mode3_network():
    """The 20-fault seeded network of tree_network in Mode III (kernel entries of
    eval_kernel_ModeIII on faults of every strike)."""
    return problems.network_small(n_faults=20, seed=3, mode=3, name="network_small_20_seed3_mode3")


def tree_network():
    """Seeded 20-fault network for the cluster / block cluster trees (leaf 16, eta 2)."""
    q = problems.network_small(n_faults=20, seed=3)
    q.nleaf = 16
    q.eta = 2.0
    return q


def grid_network(ds=2.0):
    """Three horizontal and three vertical Mode II faults whose nodes coincide at
    the nine intersections (crossings), ~2.7k elements: the pressure solve spans
    several device workgroups and carries crossing terms.  A source on fault 1."""
    fs = []
    for y in (0.0, 300.0, 600.0):
        fs.append(_seg(0.0, y, 1000.0, y, int(round(1000.0 / ds))))
    for x in (100.0, 500.0, 900.0):
        fs.append(_seg(x, -100.0, x, 700.0, int(round(800.0 / ds))))
    faults = []
    for i, nd in enumerate(fs):
        n = nd.shape[0] - 1
        one = np.ones(n)
        faults.append(dict(node=nd, a=(0.009 + 0.0005 * i) * one, b=0.012 * one, Dc=1e-3 * one, f0=0.6 * one,
                           V0=1e-6 * one, sigmaN=-30e6 * one, V=1e-12 * one, theta=np.log(1e6) * one, P=0 * one,
                           permeability=1e-14 * one * (1 + 0.5 * np.sin(np.arange(n) / 37.0)), Snk=one, Lk=one,
                           Tk=one))
    nels, node, arr = problems.concatenate_faults(faults)
    return problems.Problem("grid_network", 2, 0, nels, node, arr, tol_solver=1e-6,
                            sources=[(1, 75, 2e-4, 0.0, 1e6)])


SCALARS = ("mu", "cs", "cp", "Sigma11_dot", "Sigma12_dot", "Sigma22_dot", "Sigma31_dot", "Sigma32_dot", "porosity",
           "fluid_comp", "rock_comp", "dyn_viscosity", "fluid_density", "tol_solver", "eta", "eps_aca")


def problem_arrays(p):
    """A Problem as plain arrays (for a fixture file)."""
    out = {"in_node": p.node, "in_fnel": np.asarray(p.fault_nb_element, np.int64),
           "in_ints": np.array([p.fracture_mode, p.static_kernel, p.permeability_coupling, p.nleaf, p.max_rank]),
           "in_scalars": np.array([float(getattr(p, k)) for k in SCALARS]),
           "in_sources": np.array(p.sources, np.float64).reshape(-1, 5)}
    for k, v in p.arrays.items():
        out["in_" + k] = v
    return out


def problem_from_arrays(z, name):
    """Inverse of problem_arrays."""
    arrays = {k[3:]: np.asarray(z[k]) for k in z.files if k.startswith("in_") and k[3:] in problems.PER_ELEMENT}
    ints = z["in_ints"]
    sc = dict(zip(SCALARS, (float(x) for x in z["in_scalars"])))
    src = [(int(r[0]), int(r[1]), float(r[2]), float(r[3]), float(r[4])) for r in z["in_sources"]]
    return problems.Problem(name, int(ints[0]), int(ints[1]), np.asarray(z["in_fnel"]), np.asarray(z["in_node"]),
                            arrays, permeability_coupling=int(ints[2]), nleaf=int(ints[3]), max_rank=int(ints[4]),
                            sources=src, **sc)


def reference_network():
    """The reference's problems/network_rate_strengthening/config (10 Mode II faults,
    1029 elements), as read by the library's HDF5 reader into the fixture."""
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "network_rate_strengthening.npz"))
    return problem_from_arrays(z, "network_rate_strengthening")


def connected_network(static_kernel=0):
    """40 straight Mode II faults in 600 m x 600 m with every intersection snapped
    to a shared node (problems.snap_intersections): 944 elements, 52 hydraulic
    crossings, so the pulse diffuses from fault to fault."""
    return problems.network_small(n_faults=40, seed=3, box=600.0, lmin=100.0, lmax=400.0, alpha=0.5, snap=True,
                                  static_kernel=static_kernel, eta=2.0, eps_aca=1e-6, nleaf=32,
                                  name="connected_network_40")


def connected_network_spans(static_kernel=0):
    """120 snapped Mode II faults in 2 km x 2 km: 4855 elements, 143 hydraulic
    crossings, so the cooperative compute_P splits each system over three
    workgroup spans of 2048 rows with crossing terms between spans."""
    return problems.fault_network(n_faults=120, box=2000.0, lmin=100.0, lmax=800.0, alpha=0.5, seed=3, eta=2.0,
                                  eps_aca=1e-6, nleaf=32, tol=1e-8, snap=True, static_kernel=static_kernel,
                                  name="connected_network_120")


def _strike_segment(theta, x0, y0, length, n):
    t = np.linspace(0.0, 1.0, n + 1)
    return np.stack([x0 + length * np.cos(theta) * t, y0 + length * np.sin(theta) * t], 1)


def _faults_problem(segs, mode, name, seed=0, **kw):
    rng = np.random.default_rng(seed)
    faults = []
    for nd in segs:
        n = nd.shape[0] - 1
        one = np.ones(n)
        faults.append(dict(node=nd, a=0.01 * one, b=0.012 * one, Dc=1e-3 * one, f0=0.6 * one, V0=1e-6 * one,
                           sigmaN=-30e6 * (1.0 + 0.1 * rng.random(n)), V=1e-12 * one, theta=np.log(1e6) * one,
                           P=0 * one, permeability=1e-15 * one, Snk=one, Lk=one, Tk=one,
                           shear_loading_dot=1e-2 * rng.standard_normal(n),
                           normal_loading_dot=1e-2 * rng.standard_normal(n)))
    nels, node, arr = problems.concatenate_faults(faults)
    return problems.Problem(name, mode, 0, nels, node, arr, **kw)


STRIKE_FAN_RATES = dict(Sigma11_dot=0.31, Sigma12_dot=-0.17, Sigma22_dot=0.05, Sigma31_dot=0.23, Sigma32_dot=-0.41)


def strike_fan(mode):
    """Seven straight faults of strikes 0 to 3.1 rad (fake elements between them at
    other strikes) under all five uniform stress rates, per-element loading rates
    and normal stresses: the loading projections of src/initialisation.f90:159-200."""
    thetas = [0.0, 0.3, 1.1, np.pi / 2, 2.0, 2.9, 3.1]
    segs = [_strike_segment(th, 100.0 * i, 37.0 * i, 200.0 + 13.0 * i, 8 + i) for i, th in enumerate(thetas)]
    return _faults_problem(segs, mode, f"strike_fan_mode{mode}", **STRIKE_FAN_RATES)


def junction_network():
    """Five Mode II faults of 10 m elements with a T junction (fault 2 starts on
    node 6 of fault 1), an X crossing (node 2 of fault 5 on node 5 of fault 2) and
    two near misses of node 11 and node 3 of fault 1 at 0.099999999 m and
    0.099999997 m, either side of the reference's threshold 0.01f * min ds
    (src/initialisation.f90:299): three crossings.  A 10 MPa pulse on element 5,
    beside the junction."""
    f1 = _strike_segment(0.0, 0.0, 0.0, 100.0, 10)
    f2 = _strike_segment(np.pi / 2, 50.0, 0.0, 80.0, 8)
    f3 = np.stack([np.full(7, 100.0), -0.099999999 - 10.0 * np.arange(7)], 1)
    f4 = np.stack([np.full(6, 20.0), -0.099999997 - 10.0 * np.arange(6)], 1)
    f5 = _strike_segment(0.0, 40.0, 40.0, 20.0, 2)
    p = _faults_problem([f1, f2, f3, f4, f5], 2, "junction_network", permeability_coupling=0)
    p.arrays["P"][4] = 10e6
    p.arrays["permeability"][:] = np.where(p.arrays["permeability"] > 0, 1e-13, 0.0)
    return p
