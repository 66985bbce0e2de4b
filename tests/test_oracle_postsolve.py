"""Pins of the post-solve oracle (oracle/orc_postsolve.py) against what the
mathematics fixes, chosen so that a dropped term, a wrong sign or a swapped
index fails:

* flat faces (the double-node cube [-1,1]^3): the L2 projection of a
  constant is exact, so projected normals are the face normals, the gradient
  of phi = a.x (with dphi_dn = a.n) is a, and its surface gradient is
  a - (a.n) n, at every DoF;
* curved faces: projected normals of the unit sphere converge to x/|x| at
  second order;
* divergence theorem: volume_integral = -V per component for outward
  normals; Archimedes: the hydrostatic force int p n = -rho g V e_z and the
  moment about c is -rho g V c x e_z (cube centred at the origin);
* centre of pressure of p = rho g (2 - z) on the cube: (0, 0, -5/18)
  (int p = 48 rho g, int z p = -40/3 rho g);
* hydrodynamic pressure of a uniform gradient U with wind W:
  -rho U.(U/2 - W) at every DoF;
* waterline elevation of p = x on the half Wigley hull: points on the z = 0
  lines, elevation x / (rho g), four per line.
"""
import numpy as np
import pytest

from pibem_amd import mesh as M
from orc_postsolve import PostSolve

RHO, GRAV = 1000.0, 9.80665


@pytest.fixture(scope="module")
def cube():
    m = M.double_node_cube(4)
    return m, PostSolve(m)


def test_flat_face_projections_exact(cube):
    m, ps = cube
    n = ps.compute_normals().reshape(3, -1).T
    assert np.abs(n - m.normals).max() < 1e-13
    a = np.array([0.3, -1.2, 0.7])
    phi, dphi = m.support_xyz @ a, m.normals @ a
    g = ps.compute_gradients(phi, dphi).reshape(3, -1).T
    assert np.abs(g - a).max() < 1e-13
    s = ps.compute_surface_gradients(phi).reshape(3, -1).T
    assert np.abs(s - (a - dphi[:, None] * m.normals)).max() < 1e-13
    # the normal-derivative term alone: phi = 0 leaves n dphi_dn
    g0 = ps.compute_gradients(np.zeros(m.n_dofs), dphi).reshape(3, -1).T
    assert np.abs(g0 - dphi[:, None] * m.normals).max() < 1e-13


def test_sphere_normals_second_order():
    err = []
    for k in (8, 16):
        m = M.cube_sphere(k)
        n = PostSolve(m).compute_normals().reshape(3, -1).T
        ex = m.support_xyz / np.linalg.norm(m.support_xyz, axis=1)[:, None]
        err.append(np.abs(n - ex).max())
    assert err[1] < err[0] / 3.0, err


def test_volume_force_moment(cube):
    m, ps = cube
    mask = np.ones(m.n_cells)
    V = 8.0
    assert np.allclose(ps.volume_integral(mask), -V, rtol=0, atol=1e-12)
    p = ps.hydrostatic_pressure(GRAV, RHO, 0.0)
    c = np.array([0.3, -0.2, 0.7])
    F, Mo = ps.pressure_force_and_moment(mask, p, c)
    assert np.allclose(F, [0, 0, -RHO * GRAV * V], rtol=1e-13, atol=1e-8)
    assert np.allclose(Mo, -RHO * GRAV * V * np.cross(c, [0, 0, 1]), rtol=1e-13, atol=1e-8)
    # a body mask that keeps the top face only: force is int p n over z = 1
    top = m.cell_face == np.bincount(m.cell_face).argmax()  # any one face
    Ft, _ = ps.pressure_force_and_moment(top, p, c)
    assert np.count_nonzero(np.abs(Ft) > 1e-9) <= 1


def test_center_of_pressure(cube):
    m, ps = cube
    p = ps.hydrostatic_pressure(GRAV, RHO, 2.0)
    cp = ps.center_of_pressure(np.ones(m.n_cells), p)
    assert np.allclose(cp, [0, 0, -5.0 / 18.0], atol=1e-13)
    # zero pressure leaves the centre unchanged
    keep = np.array([9.0, 8.0, 7.0])
    assert np.array_equal(ps.center_of_pressure(np.ones(m.n_cells), 0 * p, keep), keep)


def test_hydrodynamic_pressure_uniform(cube):
    m, ps = cube
    U, W = np.array([1.5, -0.5, 0.25]), np.array([2.0, 1.0, -1.0])
    grad = np.repeat(U, m.n_dofs)
    wind = np.repeat(W, m.n_dofs)
    p = ps.hydrodynamic_pressure(RHO, grad, wind)
    assert np.allclose(p, -RHO * U @ (0.5 * U - W), rtol=1e-15)


def test_waterline_elevation():
    pr = M.config_wigley_half(16, 4, 4, 1)
    m = M.mark_lines(pr.mesh, lambda x: np.abs(x[:, 2]) < 1e-12, 7)
    ps = PostSolve(m)
    bl = M.boundary_lines(m)
    c, j = np.nonzero(bl & (m.line_ids == 7))
    assert c.size == 2 * 16
    cl = np.stack([c, j], axis=1)
    p = m.support_xyz[:, 0].copy()
    el = ps.free_surface_elevation(GRAV, RHO, cl, p)
    assert el.shape == (4 * c.size, 3)
    # points lie on the waterline chords, elevation is x / (rho g)
    assert np.allclose(el[:, 2], el[:, 0] / (RHO * GRAV), rtol=1e-14, atol=1e-18)
    for k, (cc, jj) in enumerate(cl):
        a, b = (m.support_xyz[m.cell_dofs[cc, u]] for u in M.CELL_LINES[jj])
        t = np.polynomial.legendre.leggauss(4)[0] * 0.5 + 0.5
        chord = a[None, :2] * (1 - t)[:, None] + b[None, :2] * t[:, None]
        assert np.allclose(el[4 * k:4 * k + 4, :2], chord, atol=1e-15)


def flat_q2_patch(n, z0=0.3):
    """n x n Q2 cells on the plane z = z0 over the warped square
    (u, v) -> (u + 0.2 u v, v + 0.1 u^2): non-affine cells whose geometry the
    Q2 interpolant still represents exactly."""
    k = 2 * n + 1
    uu, vv = np.meshgrid(np.linspace(0, 1, k), np.linspace(0, 1, k), indexing="xy")
    x = uu + 0.2 * uu * vv
    y = vv + 0.1 * uu ** 2
    xyz = np.stack([x.ravel(), y.ravel(), np.full(k * k, z0)], axis=1)
    loc = [(0, 0), (2, 0), (0, 2), (2, 2), (0, 1), (2, 1), (1, 0), (1, 2), (1, 1)]  # FE_Q(2) order (i, j)
    cells = []
    for cj in range(n):
        for ci in range(n):
            cells.append([(2 * cj + j) * k + (2 * ci + i) for i, j in loc])
    nrm = np.tile([0.0, 0.0, 1.0], (k * k, 1))
    return M.Mesh(xyz, np.asarray(cells, dtype=np.int64), 9, nrm, np.zeros(n * n, dtype=np.int32), name="q2patch")


def test_q2_warped_patch_exact():
    m = flat_q2_patch(3)
    ps = PostSolve(m)
    n = ps.compute_normals().reshape(3, -1).T
    assert np.abs(n - [0, 0, 1]).max() < 1e-13
    a = np.array([0.3, -1.2, 0.7])
    phi, dphi = m.support_xyz @ a, np.full(m.n_dofs, a[2])
    g = ps.compute_gradients(phi, dphi).reshape(3, -1).T
    assert np.abs(g - a).max() < 1e-12
    s = ps.compute_surface_gradients(phi).reshape(3, -1).T
    assert np.abs(s - [a[0], a[1], 0.0]).max() < 1e-12
    # area of the warped square: int (1 + 0.2 v)(1) - (0.2 u)(0.2 u) du dv = 1.1 - 0.04/3
    assert abs(-ps.volume_integral(np.ones(m.n_cells))[2] / 0.3 - (1.1 - 0.04 / 3)) < 1e-13


# -- double-node constraints (BEMProblem::compute_constraints) --------------
SHEAR = np.array([[1.0, 0.3, 0.0], [0.0, 1.0, 0.25], [0.15, 0.0, 1.0]])


def _rowdict(cons):
    ptr, idx, val, inh, isc = cons
    return {int(i): (sorted(zip(idx[ptr[i]:ptr[i + 1]].tolist(), val[ptr[i]:ptr[i + 1]].tolist())), inh[i])
            for i in np.nonzero(isc)[0]}


@pytest.mark.parametrize("shear", [False, True], ids=["cube", "sheared"])
def test_constraints_linear_field_exact(shear):
    """phi = a.x with Dirichlet data everywhere: the projections are exact on
    flat faces, so both normal derivatives of every Dirichlet pair equal
    a.n_i and a.n_it.  On the sheared cube n_i.n_it != 0, so a dropped
    (s_i.n_it)(n_i.n_it) term, a swapped pair of indices or a wrong
    denominator fails."""
    m = M.linear_map(M.double_node_cube(4), SHEAR if shear else np.eye(3))
    a = np.array([0.3, -0.5, 0.8])
    dmask = np.ones(m.n_dofs, dtype=np.uint8)
    got = _rowdict(PostSolve(m).compute_constraints(m.support_xyz @ a, dmask))
    dsets = m.double_sets()
    want = {i for i in range(m.n_dofs) if len(dsets[i]) > 1}
    assert set(got) == want
    nn = [m.normals[i] @ m.normals[j] for i in want for j in dsets[i] if j != i]
    assert (max(np.abs(nn)) > 0.1) == shear
    err = max(abs(inh - a @ m.normals[i]) for i, (ent, inh) in got.items())
    assert all(not ent for ent, _ in got.values()) and err < 1e-12


def test_constraints_mixed_rows():
    """Dirichlet on the upper cells of the cube, seeded 8-source field: the
    row pattern is fixed by the double sets and the Dirichlet mask alone
    (equal to the rows of the exact-gradient constraints), Neumann partners
    read (master, 1), Neumann partners of a Dirichlet master carry the
    master's tmp_rhs, and the Dirichlet pairs whose cells hold Dirichlet data
    only approach the exact normal derivatives as the mesh is refined."""
    errs = []
    for n in (4, 8):
        pr = M.config_dcube(n, 4)
        m = pr.mesh
        got = _rowdict(PostSolve(m).compute_constraints(pr.tmp_rhs, pr.dirichlet))
        want = _rowdict(M.constraints(pr))
        assert set(got) == set(want)
        dsets = m.double_sets()
        clean = np.ones(m.n_dofs, bool)      # every cell around the DoF is all-Dirichlet
        for c in m.cell_dofs:
            if not pr.dirichlet[c].all():
                clean[c] = False
        dn = (pr.exact.grad(m.support_xyz) * m.normals).sum(1)
        e = 0.0
        for i, (ent, inh) in got.items():
            assert [k for k, _ in ent] == [k for k, _ in want[i][0]]
            assert all(v == 1.0 for _, v in ent)
            master = min(k for k in dsets[i] if pr.dirichlet[k]) if pr.dirichlet[list(dsets[i])].any() else None
            if pr.dirichlet[i] == 0 and master is not None:
                assert inh == pr.tmp_rhs[master]
            if pr.dirichlet[i] and all(clean[k] for k in dsets[i]):
                e = max(e, abs(inh - dn[i]))
        errs.append(e / np.abs(dn).max())
    assert errs[1] < 0.6 * errs[0] and errs[1] < 0.2, errs
