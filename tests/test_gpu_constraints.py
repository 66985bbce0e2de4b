"""Native double-node constraints (pbfmm_compute_constraints,
BEMProblem::compute_constraints, bem_problem.cc:1376-1653) on GPU
projections of the normals and surface gradients, against the oracle's
restatement on its own projections (orc_postsolve.PostSolve.compute_constraints,
pinned in test_oracle_postsolve.py).

Bars:
* phi = a.x with Dirichlet data everywhere on the double-node cube and on a
  sheared copy (non-orthogonal normals at the edges): the projections are
  exact on flat faces, so every Dirichlet pair's inhomogeneities equal a.n
  to 1e-10 (the projections' CG tolerance);
* the seeded 8-source field with Dirichlet data on the upper cells: the
  same rows as the oracle (pattern and entries exactly), inhomogeneities
  within 1e-9 of the largest (the reference projects the surface gradient
  of tmp_rhs, which holds dphi/dn at Neumann DoFs, so rows next to Neumann
  cells are not gradients of phi; both sides do the same);
* the solve with the native rows reaches 1e-8 and agrees with the oracle's
  solve under the same rows to 1e-6.
"""
import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import mesh as M

pytestmark = pytest.mark.gpu

SHEAR = np.array([[1.0, 0.3, 0.0], [0.0, 1.0, 0.25], [0.15, 0.0, 1.0]])


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _rows(cons):
    ptr, idx, val, inh, isc = cons
    return {int(i): (sorted(zip(idx[ptr[i]:ptr[i + 1]].tolist(), val[ptr[i]:ptr[i + 1]].tolist())), inh[i])
            for i in np.nonzero(isc)[0]}


@pytest.mark.parametrize("shear", [False, True], ids=["cube", "sheared"])
def test_linear_field_rows_exact(shear):
    m = M.linear_map(M.double_node_cube(5), SHEAR if shear else np.eye(3))
    print(f"\n[case] linear field, all Dirichlet, {m.name} N={m.n_dofs}", flush=True)
    dmask = np.ones(m.n_dofs, dtype=np.uint8)
    a = np.array([0.3, -0.5, 0.8])
    sol = M.PointSources(np.zeros((0, 3)), [], linear=a)
    pr = M.Problem(m, dmask, sol, M.bc_vector(m, dmask, sol), 5)
    f = P.setup(pr, P.PBFMM_FP64)
    bp = P.BEMProblem(f)
    got = _rows(bp.compute_constraints(pr.tmp_rhs))
    dsets = m.double_sets()
    assert set(got) == {i for i in range(m.n_dofs) if len(dsets[i]) > 1}
    worst = max(abs(inh - a @ m.normals[i]) for i, (ent, inh) in got.items())
    print(f"  {len(got)} constrained rows, max inhomogeneity error {worst:.2e}")
    assert all(not ent for ent, _ in got.values()) and worst <= 1e-10
    f.close()


def test_mixed_rows_vs_oracle():
    from orc_postsolve import PostSolve
    pr = M.config_dcube(8, 5)
    print(f"\n[case] mixed cube N={pr.mesh.n_dofs}", flush=True)
    f = P.setup(pr, P.PBFMM_FP64)
    bp = P.BEMProblem(f)
    got = _rows(bp.compute_constraints(pr.tmp_rhs))
    want = _rows(PostSolve(pr.mesh).compute_constraints(pr.tmp_rhs, pr.dirichlet))
    assert set(got) == set(want) and len(got) > 0
    scale = max(abs(v[1]) for v in want.values())
    worst = 0.0
    for i, (ent, inh) in want.items():
        assert got[i][0] == ent
        worst = max(worst, abs(got[i][1] - inh))
    print(f"  {len(got)} constrained rows, max inhomogeneity difference {worst / scale:.2e} of the largest")
    assert worst <= 1e-9 * scale
    # without continuous gradients the Dirichlet pairs carry no rows
    ptr, idx, val, inh, isc = bp.compute_constraints(pr.tmp_rhs, continuous_gradient=False)
    dd = [i for i in want if pr.dirichlet[i] and not want[i][0] and
          any(pr.dirichlet[k] for k in pr.mesh.double_sets()[i] if k != i)]
    assert dd and not isc[dd].any()
    f.close()


def test_source_field_solve(oracle_mod):
    pr = M.config_dcube(8, 5)
    f = P.setup(pr, P.PBFMM_FP64)
    bp = P.BEMProblem(f)
    cons = bp.compute_constraints(pr.tmp_rhs)
    phi, dphi, it, rr = bp.solve_system(pr.tmp_rhs, tol=1e-8)
    assert rr <= 1e-8
    ph0, dph0, _, _ = oracle_mod.OracleFMM(pr).setup().solve_system(pr.tmp_rhs, cons, tol=1e-10)
    neu = pr.dirichlet == 0
    e_ph, e_dn = rel(phi[neu], ph0[neu]), rel(dphi[~neu], dph0[~neu])
    print(f"\n  solve vs oracle with the same rows: phi {e_ph:.2e} dphi {e_dn:.2e}")
    assert e_ph <= 1e-6 and e_dn <= 1e-6
    f.close()
