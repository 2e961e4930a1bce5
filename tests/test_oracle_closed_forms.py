"""The oracle against what the mathematics fixes: closed-form kernels, the
H-matrix error bound and partition invariants, discrete mass conservation of the
pressure step, and the Cash-Karp tableau.  Each check is chosen so that a dropped
term, a wrong sign or a transposed index fails it."""
import numpy as np
import pytest

import cases
from demystification_dfn_amd import problems
from oracle.oracle import OracleSimulation, eval_matrix, geometry, hmat_tree


def _straight(n=40, length=400.0, mode=3):
    p = problems.single_fault(n=n, length=length, mode=mode, static_kernel=0)
    return p


def test_mode3_self_stiffness_and_uniform_slip():
    p = _straight()
    o = OracleSimulation(p)
    ds = p.meta["ds"]
    # one element slipping: stress at its own centre is -2 mu / (pi ds)
    for j in (0, 17, p.N - 1):
        s = np.zeros(p.N)
        s[j] = 1.0
        sT, _ = o.kernel(s)
        assert sT[j] == pytest.approx(-2 * p.mu / (np.pi * ds), rel=1e-12)
    # uniform slip D on elements [i0, i1): tau(x) = mu D/(2 pi) (1/(x - x_b) - 1/(x - x_a))
    i0, i1, D = 10, 25, 1e-3
    s = np.zeros(p.N)
    s[i0:i1] = D
    sT, _ = o.kernel(s)
    xa, xb = p.node[i0, 0], p.node[i1, 0]
    xc = 0.5 * (p.node[1:, 0] + p.node[:-1, 0])
    ref = p.mu * D / (2 * np.pi) * (1 / (xc - xb) - 1 / (xc - xa))
    np.testing.assert_allclose(sT, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_mode3_dense_matrix_equals_direct_kernel():
    """directIII (dislocation form) and the H-matrix entry function are two algebraic
    forms of one operator: K s == directIII(s)."""
    p = _straight(n=60)
    o = OracleSimulation(p)
    K = eval_matrix(o, 0)
    s = np.random.default_rng(0).standard_normal(p.N)
    sT, _ = o.kernel(s)
    np.testing.assert_allclose(p.mu / (2 * np.pi) * (K @ s), sT, rtol=1e-9, atol=1e-9 * np.abs(sT).max())


def test_mode2_on_plane_normal_stress_vanishes_and_shear_closed_form():
    p = _straight(n=50, mode=2)
    p.cp = 5000.0
    o = OracleSimulation(p)
    i0, i1, D = 12, 30, 1e-3
    s = np.zeros(p.N)
    s[i0:i1] = D
    sT, sN = o.kernel(s)
    assert np.max(np.abs(sN)) <= 1e-9 * np.max(np.abs(sT))
    nu_fac = 1 - p.cs ** 2 / p.cp ** 2
    xa, xb = p.node[i0, 0], p.node[i1, 0]
    xc = 0.5 * (p.node[1:, 0] + p.node[:-1, 0])
    ref = p.mu * nu_fac * D / np.pi * (1 / (xc - xb) - 1 / (xc - xa))
    np.testing.assert_allclose(sT, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def _dense_from_export(d, n):
    H = np.zeros((n, n))
    for b in range(len(d["dense_m"])):
        r0, c0, m, k = d["dense_row0"][b], d["dense_col0"][b], d["dense_m"][b], d["dense_n"][b]
        H[r0:r0 + m, c0:c0 + k] += d["dense_data"][d["dense_off"][b]:d["dense_off"][b] + m * k].reshape(m, k, order="F")
    for b in range(len(d["lr_m"])):
        r0, c0, m, k, r = d["lr_row0"][b], d["lr_col0"][b], d["lr_m"][b], d["lr_n"][b], d["lr_k"][b]
        U = d["lr_u"][d["lr_uoff"][b]:d["lr_uoff"][b] + m * r].reshape(m, r, order="F")
        V = d["lr_v"][d["lr_voff"][b]:d["lr_voff"][b] + k * r].reshape(k, r, order="F")
        H[r0:r0 + m, c0:c0 + k] += U @ V.T
    return H


@pytest.mark.parametrize("mode", [2, 3])
def test_hmatrix_partition_and_aca_error(mode):
    q = problems.network_small(n_faults=16, seed=11, mode=mode, eps_aca=1e-6, eta=2.0, nleaf=16)
    o = OracleSimulation(q)
    o.initiate_hmat()
    n = q.N
    ops = (0, 1) if mode == 2 else (0,)
    for op in ops:
        d = o.hmat_export(op)
        perm = d["perm"]
        assert sorted(perm.tolist()) == list(range(n))        # bijection
        area = np.zeros((n, n), np.int32)
        for b in range(len(d["dense_m"])):
            area[d["dense_row0"][b]:d["dense_row0"][b] + d["dense_m"][b],
                 d["dense_col0"][b]:d["dense_col0"][b] + d["dense_n"][b]] += 1
        for b in range(len(d["lr_m"])):
            area[d["lr_row0"][b]:d["lr_row0"][b] + d["lr_m"][b], d["lr_col0"][b]:d["lr_col0"][b] + d["lr_n"][b]] += 1
        assert np.all(area == 1)                              # leaves partition the matrix
        assert d["lr_k"].max() <= q.max_rank
        Hh = _dense_from_export(d, n)
        K = eval_matrix(o, op)[np.ix_(perm, perm)]             # dense kernel in H order
        err = np.linalg.norm(Hh - K) / np.linalg.norm(K)
        assert err < 20 * q.eps_aca, err
        # the admissible blocks satisfy eta*dist >= max(diam) on element-centre boxes
        el = geometry(o)["elem"].reshape(-1, 2)[perm]
        for b in range(len(d["lr_m"])):
            R = el[d["lr_row0"][b]:d["lr_row0"][b] + d["lr_m"][b]]
            C = el[d["lr_col0"][b]:d["lr_col0"][b] + d["lr_n"][b]]
            dr = np.linalg.norm(R.max(0) - R.min(0))
            dc = np.linalg.norm(C.max(0) - C.min(0))
            gap = np.maximum(0, np.maximum(R.min(0) - C.max(0), C.min(0) - R.max(0)))
            assert q.eta * np.linalg.norm(gap) >= max(dr, dc) * (1 - 1e-12)
        # recursive Hgemv_right == flat block-list product == dense reconstruction
        x = np.random.default_rng(op).standard_normal(n)
        y = o.Hgemv_right(x, op=op)
        np.testing.assert_allclose(y, Hh @ x, rtol=1e-12, atol=1e-12 * np.abs(Hh @ x).max())
        o2 = OracleSimulation(q)
        o2.hmat_import(op, d)
        np.testing.assert_allclose(o2.Hgemv_right(x, op=op), y, rtol=1e-12, atol=1e-13 * np.abs(y).max())


def test_cluster_tree_invariants():
    q = cases.tree_network()
    o = OracleSimulation(q)
    o.initiate_hmat()
    t = hmat_tree(o, 0)
    ind = t["ind_clusters"].reshape(-1, 2)
    sizes = ind[:, 1] - ind[:, 0] + 1
    assert ind[0, 0] == 1 and ind[0, 1] == q.N
    # leaves (size <= nleaf) tile 1..N in order
    bt = t["bltree"]
    assert set(np.unique(bt[:, 0])) <= {0, 1, 2}
    leaves = sorted({tuple(r) for r in ind[sizes <= q.nleaf]})
    pos = 1
    for a, b in leaves:
        if a == pos:
            pos = b + 1
    assert pos == q.N + 1


def test_compute_P_mass_conservation_and_stop_rule():
    """Uniform ds, no-flux ends, no crossings: sum(x - P) ds phi beta = dt Q."""
    p = problems.single_fault(n=200, length=2000.0, tol=1e-8)
    o = OracleSimulation(p)
    dt = 3000.0
    x, pdot, ier = o.compute_P(dt)
    assert ier == 0
    P0 = p.arrays["P"]
    beta = p.fluid_comp + p.rock_comp
    ds = p.meta["ds"]
    Q = p.sources[0][2]
    assert np.sum(x - P0) * ds * p.porosity * beta == pytest.approx(dt * Q, rel=1e-6)
    np.testing.assert_allclose(pdot, (x - P0) / dt, rtol=1e-14, atol=0)
    # the diffusion front: a symmetric profile peaked at the source element
    k = p.sources[0][1] - 1
    assert np.argmax(x) == k
    np.testing.assert_allclose(x[k - 20:k], x[k + 20:k:-1], rtol=1e-6)


def test_cash_karp_tableau():
    """Consistency of the Cash-Karp (1990) coefficients used by rkck
    (src/ode_solver/ode_bs.f90:200-210): row sums, order-5 and embedded order-4
    quadrature conditions."""
    A = np.array([0, 0.2, 0.3, 0.6, 1.0, 0.875])
    B = [[], [0.2], [3 / 40, 9 / 40], [0.3, -0.9, 1.2], [-11 / 54, 2.5, -70 / 27, 35 / 27],
         [1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]]
    for i in range(1, 6):
        assert sum(B[i]) == pytest.approx(A[i], abs=1e-15)
    c5 = np.array([37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771])
    c4 = np.array([2825 / 27648, 0, 18575 / 48384, 13525 / 55296, 277 / 14336, 0.25])
    for q in range(5):
        assert np.dot(c5, A ** q) == pytest.approx(1 / (q + 1), abs=1e-14)
    for q in range(4):
        assert np.dot(c4, A ** q) == pytest.approx(1 / (q + 1), abs=1e-14)
