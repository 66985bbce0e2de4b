"""Iterative refinement with a recycled Krylov space (pbfmm_set_mixed_scheme
2: the second FP32 inner solve keeps the first one's basis, GCRO-style)
against refinement without it (scheme 3), the same with the inner solves'
basis vectors stored in FP32 (scheme 4) and the FP64 solve.

Bars, each solve to true relative residual 1e-8:
* the same solution as the FP64 solve within 1e-6 (both are converged to
  1e-8 with operators 1e-6 apart) and as refinement without recycling;
* no more Arnoldi steps than refinement without recycling, on a Q1 sphere, a
  Q2 spheroid and a constrained double-node cube, and within two steps of
  the FP32-basis refinement;
* with a short restart (12), so that the recycled space is the first inner
  solve's last cycle and the second inner solve restarts too: still 1e-8.
"""
import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import hip
from pibem_amd import mesh as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


CASES = {"sphere_p5": lambda: M.config_sphere(16, 5), "spheroid_q2": lambda: M.config_spheroid(8, 6),
         "dcube": lambda: M.config_dcube(8, 5)}


@pytest.mark.parametrize("name", list(CASES))
def test_recycled_refinement(name):
    pr = CASES[name]()
    print(f"\n[case] {pr.mesh.name} N={pr.mesh.n_dofs}", flush=True)
    f = P.setup(pr, P.PBFMM_MIXED)
    bp = P.BEMProblem(f)
    r = bp.compute_rhs(pr.tmp_rhs)
    if name == "dcube":
        cons = bp.compute_constraints(pr.tmp_rhs)
        r[cons[4] == 1] = cons[3][cons[4] == 1]
    rhs = hip.DeviceArray.from_numpy(r)
    out = {}
    for key, prec, scheme, restart in (("fp64", P.PBFMM_FP64, 0, 100), ("plain", P.PBFMM_MIXED, 3, 100),
                                       ("recycled", P.PBFMM_MIXED, 2, 100), ("recycled_r12", P.PBFMM_MIXED, 2, 12),
                                       ("recycled_fp32_basis", P.PBFMM_MIXED, 4, 100)):
        f.set_mixed_scheme(scheme)
        x = hip.DeviceArray(f.n)
        its, rr = f.gmres_dev(rhs.value, x.value, 1e-8, restart, 2000, prec)
        hip.synchronize()
        out[key] = (x.numpy(), its, rr)
        print(f"  {key}: {its} steps, relres {rr:.2e}")
        assert rr <= 1e-8
    x64 = out["fp64"][0]
    for key in ("plain", "recycled", "recycled_r12", "recycled_fp32_basis"):
        e = rel(out[key][0], x64)
        print(f"  {key} vs FP64 solution {e:.2e}")
        assert e <= 1e-6
    assert out["recycled"][1] <= out["plain"][1]
    assert abs(out["recycled"][1] - out["recycled_fp32_basis"][1]) <= 2
    f.close()


@pytest.mark.parametrize("name", ["sphere_p5", "spheroid_q2"])
def test_gmres_reorth(name):
    """Gram-Schmidt with the second projection always (eta 1, CGS2), when
    the first cancels severely (the default 0.05), when it leaves less than
    0.999 of the norm (the device's test then asks for the second projection
    in almost every step, so the scaling takes the second projection's norm)
    and never (0): FP64 and mixed solves reach true relative residual 1e-8
    with each, the solutions within 1e-10 (FP64) / 1e-6 (mixed) of the CGS2
    ones, the step counts within two."""
    pr = CASES[name]()
    print(f"\n[case] {pr.mesh.name} N={pr.mesh.n_dofs}", flush=True)
    f = P.setup(pr, P.PBFMM_MIXED)
    bp = P.BEMProblem(f)
    rhs = hip.DeviceArray.from_numpy(bp.compute_rhs(pr.tmp_rhs))
    for prec, bar in ((P.PBFMM_FP64, 1e-10), (P.PBFMM_MIXED, 1e-6)):
        ref = None
        for eta in (1.0, -1.0, 0.999, 0.0):
            f.set_gmres_reorth(eta)
            x = hip.DeviceArray(f.n)
            its, rr = f.gmres_dev(rhs.value, x.value, 1e-8, 100, 2000, prec)
            hip.synchronize()
            xs = x.numpy()
            ref = (xs, its) if ref is None else ref
            e = rel(xs, ref[0])
            print(f"  precision={prec} eta {eta}: {its} steps, relres {rr:.2e}, vs CGS2 {e:.2e}")
            assert rr <= 1e-8 and e <= bar and abs(its - ref[1]) <= 2
    f.set_gmres_reorth()
    f.close()
