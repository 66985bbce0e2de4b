"""The double-body operator (mirror images about z = z0,
BEMProblem::_assemble_system_double_body, bem_problem.cc:957-1135) on the GPU.

Bars:
* a sphere below the plane: FP64 FMM matvec, alpha, vmult and compute_rhs
  within 1e-10 of the oracle with the same images, the FP32 operator within
  2e-5, and the GMRES solve within 1e-6 of the oracle's solve;
* the submerged half of the Wigley hull mirrored about the waterline against
  the explicitly closed hull (the same geometry as the 81,002-DoF Wigley
  double body) with horizontal inflows: same demeaned solution within 1e-3
  (measured 2-3e-4, the size of either one's discretisation error), the
  difference being the quadrature of the image cells that touch waterline
  nodes, which the double body integrates with the regular rule (the
  reference's choice) and the closed hull with the singular one."""
import dataclasses

import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import mesh as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _below(n=8, p=6):
    pr = M.config_sphere(n, p)
    m = dataclasses.replace(pr.mesh, support_xyz=pr.mesh.support_xyz + np.array([0.3, -0.2, -2.5]))
    sol = M.PointSources([[0.3, -0.2, 1.0], [1.5, 0.5, -4.0]], [1.0, -0.5])
    return M.Problem(m, pr.dirichlet, sol, M.bc_vector(m, pr.dirichlet, sol), p, mirror_z0=0.0)


def test_double_body_parity(oracle_mod):
    from pibem_amd import hip
    pr = _below()
    n = pr.mesh.n_dofs
    print(f"\n[case] double-body sphere N={n}", flush=True)
    o = oracle_mod.OracleFMM(pr).setup()
    f = P.setup(pr, P.PBFMM_MIXED)
    rng = np.random.default_rng(1234)
    a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    N0, D0 = o.matvec(a, b)
    f.generate_multipole_expansions(a, b)
    N1, D1 = f.multipole_matr_vect_products(a, b)
    da, db = hip.DeviceArray.from_numpy(a), hip.DeviceArray.from_numpy(b)
    oN, oD = hip.DeviceArray(n), hip.DeviceArray(n)
    f.matvec_dev(da.value, db.value, oN.value, oD.value, P.PBFMM_MIXED)
    hip.synchronize()
    e32 = max(rel(oN.numpy(), N0), rel(oD.numpy(), D0))
    print(f"  fp64 N {rel(N1, N0):.2e} D {rel(D1, D0):.2e}  fp32 {e32:.2e}")
    assert rel(N1, N0) <= 1e-10 and rel(D1, D0) <= 1e-10 and e32 <= 2e-5
    bp = P.BEMProblem(f)
    assert rel(bp.compute_alpha(), o.compute_alpha()) <= 1e-10
    assert rel(bp.vmult(a), o.vmult(a)) <= 1e-10
    assert rel(bp.compute_rhs(pr.tmp_rhs), o.compute_rhs(pr.tmp_rhs)) <= 1e-10
    for prec in (P.PBFMM_FP64, P.PBFMM_MIXED):
        phi, dphi, it, rr = bp.solve_system(pr.tmp_rhs, tol=1e-8) if prec == P.PBFMM_MIXED else \
            P.BEMProblem(P.setup(pr, P.PBFMM_FP64)).solve_system(pr.tmp_rhs, tol=1e-8)
        ph0, dph0, _, _ = o.solve_system(pr.tmp_rhs, tol=1e-10)
        neu = pr.dirichlet == 0
        print(f"  solve precision={prec}: iters {it} relres {rr:.2e} vs oracle {rel(phi[neu], ph0[neu]):.2e}")
        assert rr <= 1e-8 and rel(phi[neu], ph0[neu]) <= 1e-6 and rel(dphi[~neu], dph0[~neu]) <= 1e-5


def test_half_hull_mirror_vs_closed_hull():
    from pibem_amd import hip
    nx, nzh = 80, 10
    half = M.config_wigley_half(nx, nzh, 6, 2)
    full = M.config_wigley(nx, 2 * nzh, 6, 1)
    sols = []
    for pr, exact in ((half, half.exact_set), (full, half.exact_set)):
        n = pr.mesh.n_dofs
        f = P.setup(pr, P.PBFMM_FP64)
        bp = P.BEMProblem(f)
        bp.compute_alpha()
        cons = M.constraints(pr)
        bp.set_constraints(cons)
        X = []
        for e in exact:
            g = M.bc_vector(pr.mesh, pr.dirichlet, e)
            rhs = bp.compute_rhs(g)
            rhs[cons[4] == 1] = cons[3][cons[4] == 1]
            dr, dx = hip.DeviceArray.from_numpy(rhs), hip.DeviceArray(n)
            it, rr = f.gmres_dev(dr.value, dx.value, 1e-8, 100, 1000, P.PBFMM_FP64)
            assert rr <= 1e-8
            X.append(dx.numpy())
        sols.append(np.array(X))
    # lower-half DoFs of the closed hull, in the half hull's numbering
    nper_h, nper_f = (nx + 1) * (nzh + 1), (nx + 1) * (2 * nzh + 1)
    idx = np.concatenate([s * nper_f + np.arange(nx + 1)[:, None] * (2 * nzh + 1) + np.arange(nzh + 1)[None, :]
                          for s in (0, 1)]).ravel()
    assert np.allclose(full.mesh.support_xyz[idx], half.mesh.support_xyz, atol=1e-14)
    for r in range(2):
        h, c = sols[0][r], sols[1][r][idx]
        ex = half.exact_set[r].phi(half.mesh.support_xyz)
        d = rel(h - h.mean(), c - c.mean())
        print(f"  rhs {r}: half+mirror vs closed {d:.2e}; vs exact: half {rel(h - h.mean(), ex - ex.mean()):.2e} "
              f"closed {rel(c - c.mean(), ex - ex.mean()):.2e}")
        assert d <= 1e-3
