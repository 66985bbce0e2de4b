"""The GMRES step's operator replayed as a HIP graph (pbfmm_set_graph_mode)
against the same solves with direct launches.

Every Arnoldi step replays the captured graph between the step's own eager
kernels (the Gram-Schmidt projections, the normalisation) and the host's
round trips, so a replay that went wrong would change the Krylov basis and
the solution.  Bars, each solve to true relative residual 1e-8:

* FP64 and mixed precision, every launch plan of the operator (atomic M2L,
  slot M2L, the SpMV on a second stream, whose fork and join the graph
  holds): solutions within 1e-10 (FP64) / 1e-7 (mixed) of the direct ones,
  the same step counts within one;
* two solves in a row with graphs: the second, replaying graphs captured in
  the first, equal to the first;
* constraint rows installed, then removed, between solves: the graph is
  captured again each time, and each solve equals its direct twin;
* the timed default ends on a choice per precision, 0 or 1.
"""
import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import hip
from pibem_amd import mesh as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _solve(f, rhs, prec):
    x = hip.DeviceArray(f.n)
    its, rr = f.gmres_dev(rhs.value, x.value, 1e-8, 100, 2000, prec)
    hip.synchronize()
    return x.numpy(), its, rr


@pytest.mark.parametrize("name", ["sphere_p5", "spheroid_q2"])
def test_graph_solves_equal_direct(name):
    pr = {"sphere_p5": lambda: M.config_sphere(16, 5), "spheroid_q2": lambda: M.config_spheroid(8, 6)}[name]()
    print(f"\n[case] {pr.mesh.name} N={pr.mesh.n_dofs}", flush=True)
    f = P.setup(pr, P.PBFMM_MIXED)
    bp = P.BEMProblem(f)
    rhs = hip.DeviceArray.from_numpy(bp.compute_rhs(pr.tmp_rhs))
    for prec, bar in ((P.PBFMM_FP64, 1e-10), (P.PBFMM_MIXED, 1e-7)):
        for plan in (0, 4, 12, 24, 56):
            f.set_side_fusion(prec, plan)
            f.set_graph_mode(0)
            x0, it0, rr0 = _solve(f, rhs, prec)
            f.set_graph_mode(1)
            x1, it1, rr1 = _solve(f, rhs, prec)
            x2, it2, rr2 = _solve(f, rhs, prec)  # replays of the graphs captured above
            e1, e2 = rel(x1, x0), rel(x2, x1)
            print(f"  precision={prec} plan {plan}: steps {it0}/{it1}/{it2}, graph vs direct {e1:.2e}, "
                  f"second graph solve {e2:.2e}")
            assert max(rr0, rr1, rr2) <= 1e-8
            assert abs(it1 - it0) <= 1 and e1 <= bar and e2 <= bar
    f.set_graph_mode(-1)
    for prec in (P.PBFMM_FP64, P.PBFMM_MIXED):
        f.set_side_fusion(prec, -1)
        _, _, rr = _solve(f, rhs, prec)
        assert rr <= 1e-8
    st = f.tuning_state()
    print(f"  timed choice: graph FP64 {st['graph_fp64']} FP32 {st['graph_fp32']}, ms {f.graph_timings()}")
    assert st["graph_fp64"] in (0, 1) and st["graph_fp32"] in (0, 1)
    f.close()


def test_graph_recaptured_on_new_constraints():
    pr = M.config_dcube(6, 5)
    print(f"\n[case] {pr.mesh.name} N={pr.mesh.n_dofs} with and without constraint rows", flush=True)
    f = P.setup(pr, P.PBFMM_FP64)
    bp = P.BEMProblem(f)
    rhs_plain = hip.DeviceArray.from_numpy(bp.compute_rhs(pr.tmp_rhs))
    cons = M.constraints(pr)
    for step, rows in enumerate((cons, None, cons)):
        bp.set_constraints(rows)
        r = bp.compute_rhs(pr.tmp_rhs)
        if rows is not None:
            r[rows[4] == 1] = rows[3][rows[4] == 1]
        rhs = hip.DeviceArray.from_numpy(r) if rows is not None else rhs_plain
        f.set_graph_mode(0)
        x0, it0, rr0 = _solve(f, rhs, P.PBFMM_FP64)
        f.set_graph_mode(1)
        x1, it1, rr1 = _solve(f, rhs, P.PBFMM_FP64)
        e = rel(x1, x0)
        print(f"  solve {step} ({'constrained' if rows is not None else 'no rows'}): steps {it0}/{it1}, "
              f"graph vs direct {e:.2e}")
        assert rr0 <= 1e-8 and rr1 <= 1e-8 and abs(it1 - it0) <= 1 and e <= 1e-10
    f.close()
