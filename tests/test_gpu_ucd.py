"""The BEM-FMM path on the grids pi-BEM ships, read from UCD and refined
(tests/golden/meshes/): GPU against the oracle on the same ingested mesh.

Bars: FP64 FMM matvec within 1e-10 of the oracle, the FP32 operator within
2e-5 (closed bodies and the open DTC ship hull); on the closed bodies the GMRES
solve reaches true relative residual 1e-8 and matches the oracle's solve to
1e-6 (with the double-node constraints where there are doubles)."""
import os

import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import mesh as M
from pibem_amd import ucd

pytestmark = pytest.mark.gpu
MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


CASES = {
    "sphere": ("coarse_sphere", 4, True, dict(dirichlet_by_z=True)),
    "cube_doubles": ("coarse_cube_double_nodes", 4, False, dict(dirichlet_by_z=True)),
    "sphere_doubles": ("coarse_sphere_double_nodes", 4, True, dict(dirichlet_materials=(0,))),
    "dtc_ship": ("dtc_ship", 2, False, dict(dirichlet_by_z=True)),
}


def _problem(key):
    name, levels, sphere, kw = CASES[key]
    m = ucd.load(os.path.join(MESHES, name + ".inp"), levels=levels,
                 project=ucd.sphere_projection() if sphere else None, name=name)
    return ucd.problem(m, p=6, **kw)


@pytest.mark.parametrize("key", list(CASES))
def test_ingested_matvec(key, oracle_mod):
    from pibem_amd import hip
    pr = _problem(key)
    n = pr.mesh.n_dofs
    print(f"\n[case] {key} N={n} cells={pr.mesh.n_cells}", flush=True)
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


@pytest.mark.parametrize("key", ["sphere", "cube_doubles", "sphere_doubles"])
def test_ingested_solve(key, oracle_mod):
    pr = _problem(key)
    o = oracle_mod.OracleFMM(pr).setup()
    f = P.setup(pr, P.PBFMM_FP64)
    bp = P.BEMProblem(f)
    has_doubles = int(np.diff(pr.mesh.dbl_ptr).max()) > 1
    cons = M.constraints(pr) if has_doubles else None
    if cons is not None:
        bp.set_constraints(cons)
    phi, dphi, it, rr = bp.solve_system(pr.tmp_rhs, tol=1e-8)
    ph0, dph0, _, _ = o.solve_system(pr.tmp_rhs, cons, tol=1e-10)
    neu = pr.dirichlet == 0
    ex = pr.exact.phi(pr.mesh.support_xyz)
    e_o, e_x = rel(phi[neu], ph0[neu]), rel(phi[neu], ex[neu])
    print(f"  {key}: iters {it} relres {rr:.2e} vs oracle {e_o:.2e} vs exact {e_x:.2e}")
    assert rr <= 1e-8 and e_o <= 1e-6
