"""The near field's two singular rules on the GPU, against the oracle, and
the Q2 accuracy they give.

Bars:
* near-field entries (N, D on the shared pattern) under the Telles rule and
  under the Duffy split, Q1 (configuration 1) and Q2 (spheroid): identical
  pattern, FP64 values within 1e-12 of the largest entry;
* the GPU FMM under the Duffy split against the oracle's dense collocation
  matvec with the same rule: within 2x the oracle FMM's own truncation error
  there (and within 1e-10 of the oracle FMM);
* a mid-size Q2 spheroid (24,578 DoFs, p = 8): FP64 FMM matvec within 1e-10
  of the oracle, the FP32 operator within 2e-5;
* Q2 solves with the Duffy split: true relative residual 1e-8 in both
  precisions, the oracle's solution within 1e-7 (phi) / 1e-6 (dphi/dn), and
  the analytic dphi/dn within 3e-2 (the Telles rule leaves 0.12 there).
"""
import dataclasses

import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import mesh as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _problem(name, rule):
    pr = {"c1": lambda: M.config_sphere(16, 4), "q2": lambda: M.config_spheroid(8, 6)}[name]()
    return dataclasses.replace(pr, sing_rule=rule)


@pytest.mark.parametrize("rule", [P.PBFMM_SING_TELLES, P.PBFMM_SING_DUFFY], ids=["telles", "duffy"])
@pytest.mark.parametrize("name", ["c1", "q2"])
def test_nearfield_rule(oracle_mod, name, rule):
    pr = _problem(name, rule)
    print(f"\n[case] {pr.mesh.name} N={pr.mesh.n_dofs} rule={rule}", flush=True)
    o = oracle_mod.OracleFMM(pr).setup()
    f = P.setup(pr, P.PBFMM_FP64)
    rp, col, vN, vD, _ = o.csr()
    grp, gcol, gN, gD = f.nearfield(P.PBFMM_FP64)
    assert np.array_equal(rp, grp) and np.array_equal(col, gcol)
    eN = np.max(np.abs(gN - vN)) / np.max(np.abs(vN))
    eD = np.max(np.abs(gD - vD)) / np.max(np.abs(vD))
    print(f"  near field max rel err N {eN:.2e} D {eD:.2e}")
    assert eN <= 1e-12 and eD <= 1e-12
    f.close()


def test_rules_differ_on_singular_entries_only():
    """Same pattern under both rules, and the entries differ only at (row i,
    column j) where j is a DoF of a cell that holds i or one of its double
    nodes: the cells the singular rule integrates."""
    pr = _problem("q2", P.PBFMM_SING_TELLES)
    f0 = P.setup(pr, P.PBFMM_FP64)
    f1 = P.setup(dataclasses.replace(pr, sing_rule=P.PBFMM_SING_DUFFY), P.PBFMM_FP64)
    r0, c0, n0, d0 = f0.nearfield(P.PBFMM_FP64)
    r1, c1, n1, d1 = f1.nearfield(P.PBFMM_FP64)
    assert np.array_equal(r0, r1) and np.array_equal(c0, c1)
    m = pr.mesh
    dsets = m.double_sets()
    cells_of = [[] for _ in range(m.n_dofs)]
    for c, dofs in enumerate(m.cell_dofs):
        for d in dofs:
            cells_of[d].append(c)
    allowed = np.zeros(len(c0), bool)
    for i in range(m.n_dofs):
        own = {int(d) for k in dsets[i] for c in cells_of[k] for d in m.cell_dofs[c]}
        allowed[r0[i]:r0[i + 1]] = np.isin(c0[r0[i]:r0[i + 1]], list(own))
    tol = 1e-14 * max(np.abs(d1).max(), np.abs(n1).max())
    diff = (np.abs(d0 - d1) > tol) | (np.abs(n0 - n1) > tol)
    print(f"\n  entries that differ: {diff.sum()} of {len(diff)} ({allowed.sum()} singular)")
    assert diff.any() and not (diff & ~allowed).any()
    f0.close()
    f1.close()


@pytest.mark.parametrize("p", [6, 8])
def test_q2_fmm_vs_dense_duffy(oracle_mod, p):
    pr = dataclasses.replace(M.config_spheroid(6, p), sing_rule=P.PBFMM_SING_DUFFY)
    n = pr.mesh.n_dofs
    o = oracle_mod.OracleFMM(pr).setup()
    Nm, Dm = o.dense(P.PBFMM_SING_DUFFY)
    rng = np.random.default_rng(1234)
    a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    No, Do = o.matvec(a, b)
    f = P.setup(pr, P.PBFMM_FP64)
    f.generate_multipole_expansions(a, b)
    Ng, Dg = f.multipole_matr_vect_products(a, b)
    tN, tD = rel(No, Nm @ a), rel(Do, Dm @ b)
    gN, gD = rel(Ng, Nm @ a), rel(Dg, Dm @ b)
    print(f"\n  Q2 N={n} p={p}: oracle FMM vs dense N {tN:.2e} D {tD:.2e}; GPU FMM vs dense N {gN:.2e} D {gD:.2e}")
    assert rel(Ng, No) <= 1e-10 and rel(Dg, Do) <= 1e-10
    assert gN <= 2 * tN and gD <= 2 * tD
    f.close()


@pytest.mark.parametrize("leaf", [20, 80])
def test_q2_mid_size_matvec(oracle_mod, leaf):
    """The FMM matvec on a 24,578-DoF Q2 spheroid at order 8 against the
    oracle's at the same leaf size: 20 (every leaf's unit multipoles fit LDS
    whole) and 80 (leaves of more columns, built one kernel at a time):
    FP64 within 1e-10, FP32 within 2e-5."""
    from pibem_amd import hip
    pr = M.config_spheroid(32, 8)
    pr.max_nodes_per_leaf = leaf
    n = pr.mesh.n_dofs
    print(f"\n[case] {pr.mesh.name} N={n} p=8 leaf {leaf}", flush=True)
    o = oracle_mod.OracleFMM(pr).setup()
    rng = np.random.default_rng(1234)
    a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    N0, D0 = o.matvec(a, b)
    f = P.setup(pr, P.PBFMM_MIXED)
    da, db = hip.DeviceArray.from_numpy(a), hip.DeviceArray.from_numpy(b)
    oN, oD = hip.DeviceArray(n), hip.DeviceArray(n)
    errs = {}
    for prec in (P.PBFMM_FP64, P.PBFMM_MIXED):
        f.matvec_dev(da.value, db.value, oN.value, oD.value, prec)
        hip.synchronize()
        errs[prec] = (rel(oN.numpy(), N0), rel(oD.numpy(), D0))
        print(f"  precision={prec}: N {errs[prec][0]:.2e} D {errs[prec][1]:.2e}")
    assert max(errs[P.PBFMM_FP64]) <= 1e-10 and max(errs[P.PBFMM_MIXED]) <= 2e-5
    f.close()


@pytest.mark.parametrize("precision", [P.PBFMM_FP64, P.PBFMM_MIXED])
def test_q2_solve_duffy(oracle_mod, precision):
    pr = _problem("q2", P.PBFMM_SING_DUFFY)
    f = P.setup(pr, precision)
    phi, dphi, it, rr = P.BEMProblem(f).solve_system(pr.tmp_rhs, tol=1e-8)
    ph0, dph0, _, _ = oracle_mod.OracleFMM(pr).setup().solve_system(pr.tmp_rhs, tol=1e-10)
    x = pr.mesh.support_xyz
    neu = pr.dirichlet == 0
    ex_dn = (pr.exact.grad(x) * pr.mesh.normals).sum(1)
    e_dn = rel(dphi[~neu], ex_dn[~neu])
    print(f"\n  Q2 N={pr.mesh.n_dofs} precision={precision} iters={it} relres={rr:.2e} "
          f"dphi vs oracle {rel(dphi[~neu], dph0[~neu]):.2e} vs analytic {e_dn:.2e}")
    assert rr <= 1e-8
    assert rel(phi[neu], ph0[neu]) <= 1e-7 and rel(dphi[~neu], dph0[~neu]) <= 1e-6
    assert e_dn <= 3e-2
    f.close()
