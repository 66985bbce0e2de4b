"""Pins of the oracle functions that the reference's own tests leave open:
Legendre functions against scipy, the dipole P2M against finite differences of
the monopole P2M, M2M / L2L exactness identities, quadrature rules against
closed forms, and the double-layer operator against the solid angles of a
cube (1/2 on faces, 1/4 on edges, 1/8 at corners)."""
import math

import numpy as np
import pytest
from scipy.special import lpmv


def test_legendre_against_scipy(oracle_mod):
    for x in (-0.97, -0.3, 0.0, 0.41, 0.999):
        P, dP = oracle_mod.legendre(20, x, deriv=True)
        for n in range(21):
            for m in range(n + 1):
                norm = math.sqrt(math.factorial(n - m) / math.factorial(n + m))
                ref = norm * lpmv(m, n, x)
                assert abs(P[n, m] - ref) <= 1e-12 * max(1.0, abs(ref)), (n, m, x)
                h = 1e-6
                fd = norm * (lpmv(m, n, x + h) - lpmv(m, n, x - h)) / (2 * h)
                assert abs(dP[n, m] - fd) <= 2e-5 * max(1.0, abs(fd)), (n, m, x)


def test_legendre_closed_forms(oracle_mod):
    # source/ass_leg_function.cc: P_1_1 = -(sqrt2/2) sqrt(1-x^2), P_2_1 = -(sqrt6/2) x sqrt(1-x^2)
    x = 0.37
    P = oracle_mod.legendre(4, x)
    s = math.sqrt(1 - x * x)
    assert P[1, 1] == pytest.approx(-math.sqrt(2) / 2 * s, rel=1e-15)
    assert P[2, 1] == pytest.approx(-math.sqrt(6) / 2 * x * s, rel=1e-15)
    assert P[2, 0] == pytest.approx(1.5 * x * x - 0.5, rel=1e-15)


def _rand_sources(seed, n=7, c=(0.3, -0.2, 0.5), r=0.4):
    rng = np.random.default_rng(seed)
    y = np.asarray(c) + rng.uniform(-r, r, size=(n, 3))
    q = rng.normal(size=n)
    nr = rng.normal(size=(n, 3))
    nr /= np.linalg.norm(nr, axis=1, keepdims=True)
    return y, q, nr


def test_dipole_p2m_is_source_derivative(oracle_mod):
    O = oracle_mod
    p, c = 8, np.array([0.3, -0.2, 0.5])
    y, q, nr = _rand_sources(1)
    M = O.p2m_dipole(p, c, q, y, nr)
    h = 1e-5
    fd = (O.p2m(p, c, q, y + h * nr) - O.p2m(p, c, q, y - h * nr)) / (2 * h)
    assert np.max(np.abs(M - fd)) <= 1e-7 * np.max(np.abs(M))


def test_m2m_equals_p2m_at_new_centre(oracle_mod):
    # a degree-n coefficient about the new centre only needs degrees <= n about the old one
    O = oracle_mod
    p = 8
    c1, c2 = np.array([0.3, -0.2, 0.5]), np.array([0.55, -0.45, 0.8])
    y, q, _ = _rand_sources(2)
    direct = O.p2m(p, c2, q, y)
    trans = O.m2m(p, c1, O.p2m(p, c1, q, y), c2)
    assert np.max(np.abs(direct - trans)) <= 1e-13 * np.max(np.abs(direct))


def test_l2l_preserves_local_polynomial(oracle_mod):
    O = oracle_mod
    p = 7
    y, q, _ = _rand_sources(3, c=(3.0, 3.0, 3.0), r=0.3)
    M = O.p2m(p, [3, 3, 3], q, y)
    L = O.m2l(p, [3, 3, 3], M, np.array([-1.0, -0.5, -0.8]))
    L2 = O.l2l(p, [-1.0, -0.5, -0.8], L, np.array([-0.7, -0.2, -1.1]))
    for x in ([-0.6, -0.1, -1.0], [-0.9, -0.4, -0.9], [-0.75, -0.3, -1.2]):
        a = O.le_eval(p, [-1.0, -0.5, -0.8], L, np.array(x))
        b = O.le_eval(p, [-0.7, -0.2, -1.1], L2, np.array(x))
        assert abs(a - b) <= 1e-13 * abs(a)


def test_m2l_converges_with_order(oracle_mod):
    O = oracle_mod
    y, q, _ = _rand_sources(4, c=(3.0, 3.0, 3.0), r=0.3)
    x = np.array([-0.2, 0.1, -0.1])
    exact = sum(qi / np.linalg.norm(x - yi) for qi, yi in zip(q, y))
    errs = []
    for p in (2, 4, 6, 8, 10):
        L = O.m2l(p, [3, 3, 3], O.p2m(p, [3, 3, 3], q, y), np.zeros(3))
        errs.append(abs(O.le_eval(p, [0, 0, 0], L, x) - exact))
    assert all(errs[i + 1] < errs[i] for i in range(len(errs) - 1))
    assert errs[-1] < 1e-8 * abs(exact)


def test_gauss_exactness(oracle_mod):
    for n in (1, 2, 4, 5, 8):
        x, w = oracle_mod.gauss01(n)
        for k in range(2 * n):
            assert abs((w * x ** k).sum() - 1.0 / (k + 1)) <= 1e-14


def test_singular_rules_corner_closed_form(oracle_mod):
    exact = 2 * math.log(1 + math.sqrt(2))  # int_[0,1]^2 dx dy / sqrt(x^2 + y^2)
    prev = None
    for n in (3, 5, 8, 12):
        x, w = oracle_mod.telles01(n, 0.0)
        X, Y = np.meshgrid(x, x, indexing="xy")
        t = (np.outer(w, w) / np.sqrt(X ** 2 + Y ** 2)).sum()
        err = abs(t - exact)
        assert prev is None or err < prev
        prev = err
    assert prev < 3e-4
    pts, w = oracle_mod.duffy(5, 0.0, 0.0)
    assert abs((w / np.linalg.norm(pts, axis=1)).sum() - exact) < 1e-7


def test_telles_rule_vertex_map(oracle_mod):
    # singularity at 0: eta = ((g+1)^3 - 4)/4, J = 3 (g+1)^2 / 4 on [-1, 1]
    gx, gw = oracle_mod.gauss01(5)
    x, w = oracle_mod.telles01(5, 0.0)
    g = 2 * gx - 1
    assert np.allclose(x, ((((g + 1) ** 3 - 4) / 4) + 1) / 2, atol=1e-15)
    assert np.allclose(w, gw * 3 * (g + 1) ** 2 / 4, atol=1e-15)
    assert abs(w.sum() - 1) < 1e-14


def test_cube_solid_angles(oracle_mod):
    from pibem_amd import mesh as M
    pr = M.config_dcube(6, 4)
    o = oracle_mod.OracleFMM(pr)
    o.generate_octree_blocking()
    mult = np.array([len(s) for s in pr.mesh.double_sets()])
    for rule in (0, 1):
        Nm, Dm = o.dense(rule)
        alpha = -Nm.sum(axis=1)
        for k, ref in ((1, 0.5), (2, 0.25), (3, 0.125)):
            assert np.max(np.abs(alpha[mult == k] - ref)) < 2e-5, (rule, k)
    # flat faces: no double-layer coupling (up to rounding) between DoFs interior to one face
    face_of = np.full(pr.mesh.n_dofs, -1)
    for c, f in zip(pr.mesh.cell_dofs, pr.mesh.cell_face):
        face_of[c] = f
    inner = np.where(mult == 1)[0]
    same = face_of[inner][:, None] == face_of[inner][None, :]
    assert np.max(np.abs(Nm[np.ix_(inner, inner)][same])) < 1e-14


def test_fe_partition_of_unity(oracle_mod):
    # shape values of the quadrature data sum to one and the JxW of the cube sum to its area
    from pibem_amd import mesh as M
    pr = M.config_dcube(4, 4)
    o = oracle_mod.OracleFMM(pr)
    o.generate_octree_blocking()
    nc, nq, dpc = pr.mesh.n_cells, 16, 4
    qp, qn, qw, qs = np.zeros(nc * nq * 3), np.zeros(nc * nq * 3), np.zeros(nc * nq), np.zeros(nc * nq * dpc)
    oracle_mod.lib().orc_quad_export(o.h, *(a.ctypes.data_as(oracle_mod._d) for a in (qp, qn, qw, qs)))
    assert np.allclose(qs.reshape(-1, dpc).sum(1), 1.0, atol=1e-15)
    assert abs(qw.sum() - 24.0) < 1e-12
    # outward normals: n . x > 0 on the cube
    assert np.all((qn.reshape(-1, 3) * qp.reshape(-1, 3)).sum(1) > 0)
