"""Problem descriptions and the seeded workload generators.

A :class:`Problem` holds the content of a DemystiFicatioN ``problems/<name>/config``
(src/configuration.f90:70-379, written by wrapper_python/DemystiFicatioN.py:330-422)
as numpy arrays in the reference's concatenated element order: the faults one after
another with one fake element between consecutive faults, N = sum(nb_el + 1) - 1.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi

PER_ELEMENT = ("a", "b", "Dc", "f0", "V0", "normal_loading_dot", "shear_loading_dot", "V", "theta",
               "sigmaN", "P", "permeability", "kmin", "kmax", "Snk", "Lk", "Tk")
# values the reference gives the fake slots (src/configuration.f90:199-203, :226-232)
FAKE_ONE = ("a", "b", "Dc", "f0", "V0", "Snk", "Lk", "Tk")


@dataclass
class Problem:
    name: str
    fracture_mode: int
    static_kernel: int
    fault_nb_element: np.ndarray
    node: np.ndarray                      # (N+1, 2)
    arrays: dict                          # name -> (N,) float64, fake slots included
    mu: float = 32.04e9
    cs: float = 3.464e3
    cp: float = 5000.0
    Sigma11_dot: float = 0.0
    Sigma12_dot: float = 0.0
    Sigma22_dot: float = 0.0
    Sigma31_dot: float = 0.0
    Sigma32_dot: float = 0.0
    porosity: float = 0.1
    fluid_comp: float = 0.9e-8
    rock_comp: float = 0.1e-8
    dyn_viscosity: float = 1e-3
    fluid_density: float = 1000.0
    tol_solver: float = 1e-6
    permeability_coupling: int = 0
    sources: list = field(default_factory=list)   # (fault 1-based, index 1-based, Q, t_beg, t_end)
    eta: float = 3.0
    eps_aca: float = 1e-8
    nleaf: int = 100
    max_rank: int = 20
    meta: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.node.shape[0] - 1)

    @property
    def nb_fault(self) -> int:
        return int(len(self.fault_nb_element))

    def fake_indices(self) -> np.ndarray:
        """0-based fake element indices (RateStateAgingModeIII.f90:105-113 rule)."""
        idx = np.cumsum(self.fault_nb_element + 1)
        return (idx[idx <= self.N] - 1).astype(np.int64)

    def initial_Y(self) -> np.ndarray:
        return np.concatenate([self.arrays["V"], self.arrays["theta"], self.arrays["sigmaN"]]).astype(np.float64)

    def to_c(self):
        """Return (DfnProblem, keepalive) for the C ABI."""
        keep = []

        def arr(a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        p = _abi.DfnProblem()
        p.fracture_mode = self.fracture_mode
        p.static_kernel = self.static_kernel
        p.nb_fault = self.nb_fault
        p.nb_element = self.N
        p.fault_nb_element = _abi.i64ptr(arr(self.fault_nb_element, np.int64))
        p.node = _abi.dptr(arr(self.node.reshape(-1)))
        for nm in PER_ELEMENT:
            setattr(p, nm, _abi.dptr(arr(self.arrays[nm])))
        for nm in ("mu", "cs", "cp", "Sigma11_dot", "Sigma12_dot", "Sigma22_dot", "Sigma31_dot",
                   "Sigma32_dot", "porosity", "fluid_comp", "rock_comp", "dyn_viscosity", "fluid_density",
                   "tol_solver", "eta", "eps_aca"):
            setattr(p, nm, float(getattr(self, nm)))
        p.permeability_coupling = int(self.permeability_coupling)
        p.nleaf = int(self.nleaf)
        p.max_rank = int(self.max_rank)
        ns = len(self.sources)
        p.nb_source = ns
        src = np.array(self.sources, dtype=np.float64).reshape(ns, 5) if ns else np.zeros((0, 5))
        p.source_fault = _abi.i64ptr(arr(src[:, 0].astype(np.int64), np.int64))
        p.source_index = _abi.i64ptr(arr(src[:, 1].astype(np.int64), np.int64))
        p.source_Q = _abi.dptr(arr(src[:, 2]))
        p.source_t_beg = _abi.dptr(arr(src[:, 3]))
        p.source_t_end = _abi.dptr(arr(src[:, 4]))
        return p, keep


def concatenate_faults(faults):
    """faults: list of dicts with 'node' (nel+1, 2) and per-element arrays (nel,).

    Returns (fault_nb_element, node (N+1,2), arrays) in the reference layout, the
    fake slots filled with the reference values.
    """
    nels = np.array([f["node"].shape[0] - 1 for f in faults], dtype=np.int64)
    N = int(np.sum(nels + 1) - 1)
    node = np.concatenate([np.asarray(f["node"], dtype=np.float64) for f in faults], axis=0)
    assert node.shape == (N + 1, 2)
    arrays = {}
    for nm in PER_ELEMENT:
        out = np.full(N, 1.0 if nm in FAKE_ONE else 0.0)
        pos = 0
        for f, nel in zip(faults, nels):
            out[pos:pos + nel] = f.get(nm, 0.0)
            pos += nel + 1
        arrays[nm] = out
    return nels, node, arrays


def psi_from_stress(tau, a, b, f0, V0, V, Dc, sigma):
    """Log state ln(theta V0/Dc) for steady sliding at shear stress tau
    (wrapper_python/BP6QD_A.py:84-87)."""
    theta = Dc / V0 * np.exp(a / b * np.log(2.0 * V0 / V * np.sinh(tau / (a * np.abs(sigma)))) - f0 / b)
    return np.log(V0 * theta / Dc)


def single_fault(n=4000, length=40e3, static_kernel=_abi.KERNEL_DIRECT, mode=_abi.MODE_III, tol=1e-8,
                 a=0.007, b=0.005, Dc=0.004, f0=0.6, V0=1e-6, sigmaN=-50e6, tau0=29.2e6, V_init=1e-12,
                 permeability=1e-13, Q=1.25e-6, t_inj=100 * 86400.0, name=None, **hpar):
    """BP6-QD-A-like straight fault (wrapper_python/BP6QD_A.py:13-135).

    With the defaults this is the plain-FP64 CPU-runnable workload: 4000 equal
    elements over 40 km, Mode III, aging law, injection at the fault centre for
    100 days, 1D diffusivity k/(phi eta beta) = 0.1 m^2/s, dense kernel, tol 1e-8.
    """
    ds = length / n
    x = np.linspace(-length / 2 + ds / 2, length / 2 + ds / 2, n + 1)
    node = np.stack([x, np.zeros(n + 1)], axis=1)
    one = np.ones(n)
    psi = psi_from_stress(tau0, a, b, f0, V0, V_init, Dc, sigmaN)
    fault = dict(node=node, a=a * one, b=b * one, Dc=Dc * one, f0=f0 * one, V0=V0 * one,
                 sigmaN=sigmaN * one, V=V_init * one, theta=psi * one, P=0 * one,
                 permeability=permeability * one, Lk=0.0005 * one, Tk=1e8 * one, kmin=8e-16 * one,
                 kmax=1e-13 * one, Snk=5e6 * one)
    nels, node, arrays = concatenate_faults([fault])
    element_x = 0.5 * (node[1:, 0] + node[:-1, 0])
    idx = int(np.argmin(np.abs(element_x))) + 1
    return Problem(name=name or f"single_fault_{n}", fracture_mode=mode, static_kernel=static_kernel,
                   fault_nb_element=nels, node=node, arrays=arrays, mu=32.04e9, cs=3.464e3, cp=1.0,
                   tol_solver=tol, sources=[(1, idx, Q, 0.0, t_inj)],
                   meta=dict(length=length, ds=ds), **hpar)


def snap_intersections(nodes):
    """Moves, for every pair of intersecting straight faults (node arrays, in
    order), the nearest interior node of each onto the intersection point; a node
    takes at most one crossing.  Returns the number of crossings made."""
    nf = len(nodes)
    p0 = np.array([n[0] for n in nodes])
    p1 = np.array([n[-1] for n in nodes])
    d = p1 - p0
    used = [set() for _ in range(nf)]
    made = 0
    for i in range(nf):
        # segment i against all later segments j: p0_i + s d_i = p0_j + t d_j, s, t in (0, 1)
        dj = d[i + 1:]
        den = d[i, 0] * dj[:, 1] - d[i, 1] * dj[:, 0]
        ok = np.abs(den) > 1e-12
        q = p0[i + 1:] - p0[i]
        with np.errstate(divide="ignore", invalid="ignore"):
            sv = (q[:, 0] * dj[:, 1] - q[:, 1] * dj[:, 0]) / den
            tv = (q[:, 0] * d[i, 1] - q[:, 1] * d[i, 0]) / den
        hit = ok & (sv > 0) & (sv < 1) & (tv > 0) & (tv < 1)
        for jj in np.nonzero(hit)[0]:
            j = i + 1 + int(jj)
            x = p0[i] + sv[jj] * d[i]
            ni, nj = nodes[i].shape[0] - 1, nodes[j].shape[0] - 1
            ki, kj = int(round(sv[jj] * ni)), int(round(tv[jj] * nj))
            if not (1 <= ki <= ni - 1 and 1 <= kj <= nj - 1) or ki in used[i] or kj in used[j]:
                continue
            nodes[i][ki] = x
            nodes[j][kj] = x
            used[i].add(ki)
            used[j].add(kj)
            made += 1
    return made


def power_law_lengths(rng, n, lmin, lmax, alpha):
    """Inverse-CDF sample of p(L) ~ L^-alpha on [lmin, lmax]."""
    u = rng.random(n)
    if abs(alpha - 1.0) < 1e-12:
        return lmin * (lmax / lmin) ** u
    e = 1.0 - alpha
    return (lmin ** e + u * (lmax ** e - lmin ** e)) ** (1.0 / e)


def fault_network(n_faults=800, box=7000.0, lmin=50.0, lmax=2000.0, alpha=2.5, ds=10.0, seed=0,
                  mode=_abi.MODE_II, static_kernel=_abi.KERNEL_HMATRIX, a_range=(0.008, 0.012),
                  b_weak=0.01, b_strong=0.006, frac_weak=0.7, Dc=1e-3, sigmaN=-30e6, f0=0.6, V0=1e-6,
                  V_init=1e-12, permeability=1e-15, pulse=10e6, tol=1e-6, eta=2.0, eps_aca=1e-6,
                  nleaf=64, max_rank=20, snap=False, name=None):
    """Seeded random network of straight faults in a box.

    Lengths follow p(L) ~ L^-alpha on [lmin, lmax], strikes are uniform in
    [0, pi), centres uniform in the box, elements of size ~ds.  Friction a is
    uniform per fault in a_range; b = b_weak on a fraction frac_weak of the
    faults, else b_strong.  No tectonic loading and no injection source: a
    pore-pressure pulse of `pulse` Pa sits on one seeded element.  Material and
    hydraulic values as wrapper_python/LSBB_noloading.py:141-200.

    The reference connects faults hydraulically only where two faults share a node
    (distance below 0.01 min ds, src/initialisation.f90:279-344), which random
    straight faults never do.  With snap=True, every intersection of two faults
    moves the nearest interior node of each onto the intersection point (at most
    one crossing per node), so intersecting faults become hydraulic crossings.
    """
    rng = np.random.default_rng(seed)
    lengths = power_law_lengths(rng, n_faults, lmin, lmax, alpha)
    strikes = rng.random(n_faults) * np.pi
    centres = rng.random((n_faults, 2)) * box
    avals = rng.uniform(a_range[0], a_range[1], n_faults)
    weak = rng.random(n_faults) < frac_weak
    faults = []
    nodes = []
    for i in range(n_faults):
        nel = max(1, int(round(lengths[i] / ds)))
        s = np.linspace(-lengths[i] / 2, lengths[i] / 2, nel + 1)
        d = np.array([np.cos(strikes[i]), np.sin(strikes[i])])
        nodes.append(centres[i][None, :] + s[:, None] * d[None, :])
    crossings = snap_intersections(nodes) if snap else 0
    for i in range(n_faults):
        node = nodes[i]
        nel = node.shape[0] - 1
        one = np.ones(nel)
        faults.append(dict(node=node, a=avals[i] * one, b=(b_weak if weak[i] else b_strong) * one,
                           Dc=Dc * one, f0=f0 * one, V0=V0 * one, sigmaN=sigmaN * one, V=V_init * one,
                           theta=np.log(V0 / V_init) * one, P=0 * one, permeability=permeability * one,
                           Lk=one, Tk=one, Snk=one, kmin=0 * one, kmax=0 * one))
    nels, node, arrays = concatenate_faults(faults)
    N = int(node.shape[0] - 1)
    fake = set((np.cumsum(nels + 1)[np.cumsum(nels + 1) <= N] - 1).tolist())
    real = np.array([k for k in range(N) if k not in fake])
    k_pulse = int(real[rng.integers(len(real))])
    arrays["P"][k_pulse] = pulse
    prob = Problem(name=name or f"network_{n_faults}_seed{seed}", fracture_mode=mode,
                   static_kernel=static_kernel, fault_nb_element=nels, node=node, arrays=arrays,
                   mu=32.04e9, cs=3.464e3, cp=5000.0, tol_solver=tol, eta=eta, eps_aca=eps_aca,
                   nleaf=nleaf, max_rank=max_rank,
                   meta=dict(seed=seed, n_faults=n_faults, box=box, lmin=lmin, lmax=lmax, alpha=alpha,
                             ds=ds, pulse_element=k_pulse, weak_fraction=float(np.mean(weak)),
                             snapped_crossings=crossings))
    return prob


# Workloads named by what they are (used by bench.py and the tests)
def network_80k(seed=2024, **kw):
    """800 straight faults in a 7 km x 7 km box, 10 m elements, ~80k elements, Mode II,
    H-matrix (eta 2, ACA 1e-6, leaf 64), 10 MPa pore-pressure pulse, no loading.

    The length law keeps the [50 m, 2 km] range; its exponent is 0.1 so that 800
    faults carry ~80k elements (the literal exponent 2.5 has a mean length of 127 m
    and gives ~11k elements)."""
    kw.setdefault("alpha", 0.1)
    return fault_network(n_faults=800, box=7000.0, lmin=50.0, lmax=2000.0, ds=10.0, seed=seed,
                         name=kw.pop("name", f"network800_80k_seed{seed}"), **kw)


def network_small(n_faults=12, seed=1, box=800.0, lmin=40.0, lmax=300.0, alpha=1.5, ds=10.0, **kw):
    """Small seeded Mode II network for parity tests (a few hundred elements)."""
    return fault_network(n_faults=n_faults, box=box, lmin=lmin, lmax=lmax, alpha=alpha, ds=ds,
                         seed=seed, name=kw.pop("name", f"network_small_{n_faults}_seed{seed}"), **kw)


def network_ensemble(n_networks=64, faults=(50, 400), seed=7, **kw):
    """Independent seeded networks for the ensemble engine: network i has a fault
    count drawn uniformly in `faults`, its own geometry seed and therefore its own
    friction draws (a per fault, weak/strong b), and otherwise the parameters of
    fault_network (defaults: the 80k network's length law on [50 m, 2 km] in a
    7 km box, so 50-400 faults carry ~5k-40k elements)."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(faults[0], faults[1] + 1, n_networks)
    seeds = rng.integers(0, 2**31 - 1, n_networks)
    kw.setdefault("alpha", 0.1)
    return [fault_network(n_faults=int(c), seed=int(s), name=f"ensemble{seed}_net{i}_{int(c)}faults", **kw)
            for i, (c, s) in enumerate(zip(counts, seeds))]


def loaded_faults(n_faults=2, n=200, length=2000.0, spacing=600.0, a=0.010, b=0.015, Dc=5e-4, sigmaN=-30e6,
                  tau0=18.0e6, V_init=1e-9, tau_dot=0.1, tol=1e-6, f0=0.6, V0=1e-6, name=None):
    """Parallel straight velocity-weakening faults (Mode III, dense kernel) under a
    uniform shear stress rate tau_dot (Sigma32_dot): an earthquake cycle with events
    within a few simulated years, for the event catalog and moment.  No injection."""
    ds = length / n
    faults = []
    psi = psi_from_stress(tau0, a, b, f0, V0, V_init, Dc, sigmaN)
    for i in range(n_faults):
        x = np.linspace(-length / 2, length / 2, n + 1) + 0.37 * ds * i
        node = np.stack([x, np.full(n + 1, spacing * i)], axis=1)
        one = np.ones(n)
        faults.append(dict(node=node, a=a * one, b=b * one, Dc=Dc * one, f0=f0 * one, V0=V0 * one,
                           sigmaN=sigmaN * one, V=V_init * (1.0 + 0.1 * i) * one, theta=psi * one, P=0 * one,
                           permeability=1e-15 * one, Lk=one, Tk=one, Snk=one, kmin=0 * one, kmax=0 * one))
    nels, node, arrays = concatenate_faults(faults)
    return Problem(name=name or f"loaded_faults_{n_faults}x{n}", fracture_mode=_abi.MODE_III,
                   static_kernel=_abi.KERNEL_DIRECT, fault_nb_element=nels, node=node, arrays=arrays, mu=32.04e9,
                   cs=3.464e3, cp=1.0, tol_solver=tol, Sigma32_dot=tau_dot, meta=dict(length=length, ds=ds))
