"""The double-body operator (BEMProblem::_assemble_system_double_body,
bem_problem.cc:957-1135) in the oracle, CPU only.

For a body strictly below the mirror plane, every reflected cell is a regular
cell of a second, explicitly mirrored body.  So the dense double-body matrices
must equal the one-body matrices of the union mesh (body + explicit mirror,
own DoFs) with the mirror's columns folded onto the original DoFs, to rounding;
and the explicit mirror's solid angles (alpha = -N.1 on its own nodes) must be
1/2, which pins that the image cells face outwards (their normal is the
reflected one, (n_x, n_y, -n_z)).  The FMA with images then matches the dense
double-body matvec within the truncation error."""
import dataclasses

import numpy as np

from pibem_amd import mesh as M

PERM_Q1 = [0, 2, 1, 3]


def _below(n=6):
    pr = M.config_sphere(n, 6)
    m = pr.mesh
    m = dataclasses.replace(m, support_xyz=m.support_xyz + np.array([0.3, -0.2, -2.5]))
    sol = M.PointSources([[0.3, -0.2, 1.0]], [1.0])
    dmask = pr.dirichlet
    return M.Problem(m, dmask, sol, M.bc_vector(m, dmask, sol), 6, mirror_z0=0.0)


def _union(pr):
    m = pr.mesh
    n = m.n_dofs
    xm = m.support_xyz * np.array([1.0, 1.0, -1.0])
    cells = np.concatenate([m.cell_dofs, m.cell_dofs[:, PERM_Q1] + n])
    nrm = np.concatenate([m.normals, m.normals * np.array([1.0, 1.0, -1.0])])
    mu = M.Mesh(np.concatenate([m.support_xyz, xm]), cells, 4, nrm,
                np.concatenate([m.cell_face, m.cell_face]), name="union")
    mu.dbl_ptr, mu.dbl_idx = M._double_nodes(mu.support_xyz)
    dm = np.concatenate([pr.dirichlet, pr.dirichlet])
    return M.Problem(mu, dm, pr.exact, np.zeros(2 * n), pr.p)


def test_dense_double_body_folds_union(oracle_mod):
    pr = _below()
    n = pr.mesh.n_dofs
    Nd, Dd = oracle_mod.OracleFMM(pr).dense()
    ou = oracle_mod.OracleFMM(_union(pr))
    Nu, Du = ou.dense()
    Nf = Nu[:n, :n] + Nu[:n, n:]
    Df = Du[:n, :n] + Du[:n, n:]
    assert np.max(np.abs(Nd - Nf)) <= 1e-13 * np.max(np.abs(Nf))
    assert np.max(np.abs(Dd - Df)) <= 1e-13 * np.max(np.abs(Df))
    # the explicit mirror is congruent to the body and must face outwards as
    # well: the same solid angle alpha = -N.1 at each mirrored node as at its
    # original (just under 1/2 on this faceted sphere; an inward mirror would
    # flip its sign; the singular entries follow the last bit of the reflected
    # coordinates, hence 1e-9 rather than 1e-15)
    alpha = -Nu.sum(axis=1)
    assert np.max(np.abs(alpha[n:] - alpha[:n])) <= 1e-9
    assert 0.35 < np.median(alpha[:n]) < 0.5


def test_fma_double_body_vs_dense(oracle_mod):
    """The FMA with mirror images converges to the dense double-body matvec
    with the expansion order like the one-body FMA does (measured: N 5.1e-4 at
    p = 6, 1.5e-5 at p = 10; D 2.5e-5, 5.0e-7)."""
    err = {}
    for p in (6, 10):
        pr = _below(8)
        pr.p = p
        o = oracle_mod.OracleFMM(pr).setup()
        Nd, Dd = o.dense()
        rng = np.random.default_rng(5)
        a, b = rng.uniform(-1, 1, pr.mesh.n_dofs), rng.uniform(-1, 1, pr.mesh.n_dofs)
        N1, D1 = o.matvec(a, b)
        err[p] = (np.linalg.norm(N1 - Nd @ a) / np.linalg.norm(Nd @ a),
                  np.linalg.norm(D1 - Dd @ b) / np.linalg.norm(Dd @ b))
    assert err[10][0] < 5e-5 and err[10][1] < 2e-6
    assert err[10][0] < err[6][0] / 10 and err[10][1] < err[6][1] / 10
