"""Benchmark of the BEM-FMM GMRES solve on one MI355X.

A step is one GMRES(100) solve of configuration 2 to true relative residual
1e-8 (unit sphere from a 6-face cube with 64 x 64 Q1 cells per face, 24,578
DoFs, FMM order 6, leaf <= 20 nodes, mixed Dirichlet (z > 0) / Neumann data
of phi = 1/|x - (2,1,1)|) with the mixed-precision path: FP32 operator
(near-field SpMV, P2M, M2M, M2L, L2L, M2P, L2P) inside FP64 iterative
refinement with an FP64 Arnoldi process and FP64 residuals.  The FP64 GPU path
is timed beside it, as are FMM matvecs per second, per-phase device times from
HIP events on the library's stream, and the CPU oracle (the reference's
algorithm restated in C++, OpenMP) on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "pi-bem_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = json.load(open(os.path.join(ROOT, "BASELINE.json")))["metric"] if os.path.exists(
    os.path.join(ROOT, "BASELINE.json")) else "GMRES time to rel-res 1e-8, mixed vs FP64 GPU; FMM matvecs/s; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFS = 157.3
FP64_PEAK_TFS = 78.6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(wc, s_v):
    """Algorithmic bytes of one fused N/D SpMV on the leaf-blocked layout:
    (N, D) value pairs, one column index per leaf column, the row's node id,
    the gathered phi / dphi entries once each and the two outputs."""
    n = wc["n_dofs"]
    return wc["near_values"] * 2 * s_v + wc["near_cols"] * 4 + wc["near_rows"] * 4 + 2 * n * 8 + 2 * n * 8


def m2l_flops(wc):
    """8 Kt^2 flops per (target, source, kernel): a complex Kt x Kt map; two kernels."""
    p = wc["trunc_order"]
    kt = (p + 1) * (p + 2) // 2
    return wc["m2l_pairs"] * 2 * 8 * kt * kt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this build runs on a single "
              f"MI355X; see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)

    import pibem_amd as P
    from pibem_amd import hip
    from pibem_amd import mesh as M

    if hip.device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    pr = M.config_sphere(64, 6)
    n = pr.mesh.n_dofs
    log(f"config 2: N={n} cells={pr.mesh.n_cells} p={pr.p}")
    stream = hip.Stream()
    t0 = time.perf_counter()
    f = P.BEMFMA(trunc_order=pr.p, num_octree_levels=pr.levels, max_nodes_per_leaf=pr.max_nodes_per_leaf,
                 quad_order=pr.quad_order, sing_quad_order=pr.sing_quad_order, precision=P.PBFMM_MIXED,
                 stream=stream.h)
    f.init_fma(pr.mesh, pr.dirichlet)
    f.generate_octree_blocking()
    t1 = time.perf_counter()
    f.direct_integrals()
    t2 = time.perf_counter()
    f.multipole_integrals()
    bp = P.BEMProblem(f)
    bp.compute_alpha()
    t3 = time.perf_counter()
    log(f"setup: octree {t1 - t0:.2f}s direct {t2 - t1:.2f}s multipole+alpha {t3 - t2:.2f}s")
    wc = f.work_counts()
    rhs_h = bp.compute_rhs(pr.tmp_rhs)
    rhs = hip.DeviceArray.from_numpy(rhs_h)
    x = hip.DeviceArray(n)
    zero = hip.DeviceArray(n)

    def solve(prec):
        x.copy_from(zero)
        return f.gmres_dev(rhs.value, x.value, 1e-8, 100, 2000, prec)

    res = {}
    for name, prec in (("mixed", P.PBFMM_MIXED), ("fp64", P.PBFMM_FP64)):
        for _ in range(args.warmup):
            solve(prec)
        stream.synchronize()
        hip.synchronize()
        t = time.perf_counter()
        its = rrs = None
        for _ in range(args.steps):
            its, rrs = solve(prec)
        stream.synchronize()
        hip.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        res[name] = dict(ms=dt * 1e3, iters=its, relres=rrs)
        log(f"{name}: {dt * 1e3:.2f} ms per solve, {its} operator applications, relres {rrs:.3e}")
    # accuracy of the mixed solution against the analytic field
    sol = x.numpy()
    solve(P.PBFMM_MIXED)
    sol = x.numpy()
    xs = pr.mesh.support_xyz
    neu = pr.dirichlet == 0
    phi_ex = pr.exact.phi(xs)
    dn_ex = (pr.exact.grad(xs) * pr.mesh.normals).sum(1)
    err_phi = float(np.linalg.norm((sol - phi_ex)[neu]) / np.linalg.norm(phi_ex[neu]))
    err_dn = float(np.linalg.norm((sol - dn_ex)[~neu]) / np.linalg.norm(dn_ex[~neu]))

    # FMM matvec throughput and per-phase device times
    rng = np.random.default_rng(1234)
    a = hip.DeviceArray.from_numpy(rng.uniform(-1, 1, n))
    b = hip.DeviceArray.from_numpy(rng.uniform(-1, 1, n))
    oN, oD = hip.DeviceArray(n), hip.DeviceArray(n)
    e0, e1 = hip.Event(), hip.Event()
    mv = {}
    phases = {}
    for name, prec in (("fp32", P.PBFMM_MIXED), ("fp64", P.PBFMM_FP64)):
        for _ in range(3):
            f.matvec_dev(a.value, b.value, oN.value, oD.value, prec)
        reps = 50
        e0.record(stream)
        for _ in range(reps):
            f.matvec_dev(a.value, b.value, oN.value, oD.value, prec)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        mv[name] = 1e3 / ms
        f.set_phase_timing(True)
        acc = None
        for _ in range(10):
            f.matvec_dev(a.value, b.value, oN.value, oD.value, prec)
            t = f.last_timings()
            acc = t if acc is None else {k: acc[k] + t[k] for k in t}
        f.set_phase_timing(False)
        phases[name] = {k: v / 10 for k, v in acc.items()}
        log(f"matvec {name}: {ms:.3f} ms; phases " + " ".join(f"{k}={v:.3f}" for k, v in phases[name].items()))

    spmv = {k: spmv_bytes(wc, 4 if k == "fp32" else 8) / (phases[k]["spmv"] * 1e-3) / 1e9 for k in phases}
    m2l_tfs = {k: m2l_flops(wc) / (phases[k]["m2l"] * 1e-3) / 1e12 for k in phases}
    # dominant kernel of the headline (FP32 operator) matvec
    ph = phases["fp32"]
    dom = max(("spmv", "p2m", "m2l"), key=lambda k: ph[k])
    if dom == "spmv":
        roof = dict(kernel="k_spmv<float>", bound="hbm", achieved=spmv["fp32"], peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=spmv["fp32"] / HBM_PEAK_GBS, traffic=None,
                    algorithmic_per_launch=spmv_bytes(wc, 4))
    elif dom == "m2l":
        roof = dict(kernel="k_m2l<float>", bound="mfma", achieved=m2l_tfs["fp32"], peak=FP32_PEAK_TFS,
                    unit="TFLOP/s", frac=m2l_tfs["fp32"] / FP32_PEAK_TFS, traffic=None,
                    algorithmic_per_launch=m2l_flops(wc))
    else:
        p = wc["trunc_order"]
        kt = (p + 1) * (p + 2) // 2
        fl = wc["quad_points"] * kt * 2 * 60  # ~60 flops per (point, coefficient, kernel) of the recurrences
        ach = fl / (ph["p2m"] * 1e-3) / 1e12
        roof = dict(kernel="k_p2m<float>", bound="mfma", achieved=ach, peak=FP32_PEAK_TFS, unit="TFLOP/s",
                    frac=ach / FP32_PEAK_TFS, traffic=None, algorithmic_per_launch=fl)

    cpu = None
    if not args.no_cpu:
        import oracle
        threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 16)
        oracle.set_threads(threads)
        o = oracle.OracleFMM(pr).setup()
        o.compute_alpha()
        v = rng.uniform(-1, 1, n)
        t = time.perf_counter()
        k = 0
        while k < 3 or (time.perf_counter() - t < args.cpu_seconds and k < 50):
            o.vmult(v)
            k += 1
        per = (time.perf_counter() - t) / k
        ops = res["fp64"]["iters"] + 2
        cpu_solve = per * ops
        cpu = dict(value=1.0 / cpu_solve, unit="solves/s", cores=threads, kind="port",
                   sample=f"{k} oracle vmult calls at config 2 (FP64 FMM operator, reference algorithm restated "
                          f"in C++/OpenMP), {per * 1e3:.1f} ms each, scaled by the {ops} operator applications of "
                          f"the FP64 GMRES solve", ms_per_solve=cpu_solve * 1e3, ms_per_matvec=per * 1e3)
        log(f"cpu oracle: {per * 1e3:.1f} ms per vmult on {threads} threads -> {cpu_solve:.1f} s per solve")

    out = {
        "metric": METRIC,
        "value": 1e3 / res["mixed"]["ms"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["mixed"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "config 2: unit sphere, cube-projected 6x64x64 Q1 cells, 24,578 DoFs, FMM order 6, "
                               "leaf <= 20, GMRES(100) to true rel-res 1e-8, mixed precision (FP32 operator, FP64 "
                               "Arnoldi/residual/refinement)",
                   "n_dofs": n, "trunc_order": pr.p, "parallelism": "single GPU"},
        "fp64_ms_per_solve": res["fp64"]["ms"],
        "mixed_speedup_vs_fp64": res["fp64"]["ms"] / res["mixed"]["ms"],
        "operator_applications": {"mixed": res["mixed"]["iters"], "fp64": res["fp64"]["iters"]},
        "relres": {"mixed": res["mixed"]["relres"], "fp64": res["fp64"]["relres"]},
        "error_vs_analytic": {"phi_neumann_dofs": err_phi, "dphi_dirichlet_dofs": err_dn},
        "fmm_matvecs_per_s": mv,
        "spmv_gbs": spmv,
        "m2l_tflops": m2l_tfs,
        "phase_ms": phases,
        "work_counts": wc,
        "roofline": roof,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (1e3 / res["mixed"]["ms"]) / cpu["value"] if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
