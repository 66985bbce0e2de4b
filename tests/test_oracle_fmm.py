"""The oracle's whole path on configuration 1 (1538 DoFs): FMM matvec against
the dense collocation matvec with the same singular rule (truncation error
must fall with p), alpha against the dense row sums, Green's identity on the
exact harmonic field, and the GMRES solution against the analytic solution."""
import numpy as np
import pytest

from pibem_amd import mesh as M


@pytest.fixture(scope="module")
def c1(oracle_mod):
    pr = M.config_sphere(16, 4)
    o = oracle_mod.OracleFMM(pr)
    o.generate_octree_blocking()
    Nm, Dm = o.dense()
    return pr, Nm, Dm


def test_fmm_truncation_decays(c1, oracle_mod):
    pr, Nm, Dm = c1
    rng = np.random.default_rng(1234)
    phi, dphi = rng.uniform(-1, 1, pr.mesh.n_dofs), rng.uniform(-1, 1, pr.mesh.n_dofs)
    errs = []
    for p in (2, 4, 6, 8):
        pr.p = p
        o = oracle_mod.OracleFMM(pr).setup()
        N, D = o.matvec(phi, dphi)
        errs.append((np.linalg.norm(N - Nm @ phi) / np.linalg.norm(Nm @ phi),
                     np.linalg.norm(D - Dm @ dphi) / np.linalg.norm(Dm @ dphi)))
    pr.p = 4
    for k in range(len(errs) - 1):
        assert errs[k + 1][0] < 0.5 * errs[k][0] and errs[k + 1][1] < 0.5 * errs[k][1]
    assert errs[-1][0] < 5e-4 and errs[-1][1] < 2e-5


def test_alpha_is_row_sum(c1, oracle_mod):
    pr, Nm, Dm = c1
    pr.p = 8
    o = oracle_mod.OracleFMM(pr).setup()
    a = o.compute_alpha()
    pr.p = 4
    assert np.max(np.abs(a + Nm.sum(1))) < 2e-4
    # solid angle of a vertex of this faceted sphere: just under 1/2
    assert 0.4 < a.min() and a.max() < 0.5


def test_green_identity_and_solve(c1, oracle_mod):
    pr, Nm, Dm = c1
    x = pr.mesh.support_xyz
    ph = pr.exact.phi(x)
    dn = (pr.exact.grad(x) * pr.mesh.normals).sum(1)
    a = -Nm.sum(1)
    res = a * ph + Nm @ ph - Dm @ dn
    assert np.linalg.norm(res) / np.linalg.norm(Dm @ dn) < 1e-2
    o = oracle_mod.OracleFMM(pr).setup()
    phs, dphs, it, rr = o.solve_system(pr.tmp_rhs, tol=1e-8)
    assert rr <= 1e-8
    neu = pr.dirichlet == 0
    assert np.linalg.norm((phs - ph)[neu]) / np.linalg.norm(ph[neu]) < 1e-3
    assert np.linalg.norm((dphs - dn)[~neu]) / np.linalg.norm(dn[~neu]) < 3e-2
