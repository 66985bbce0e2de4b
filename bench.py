"""Benchmark of the BEM-FMM GMRES solve on one MI355X.

A step is one GMRES(100) solve of configuration 3 to true relative residual
1e-8 with the mixed-precision path: prolate spheroid (4,1,1) from a 6-face
cube grid with 128 x 128 Q2 cells per face (98,304 cells, 393,218 DoFs;
equiangular face coordinates c -> tan(pi c / 4) before the projection),
FMM order 8, Dirichlet on cells with centroid z > 0 and
Neumann elsewhere, data from phi = x + 8 seeded exterior point sources at
1.5x the surface (seed 3), singular cells integrated by the Duffy split of the
reference's dense solve.  Mixed precision: FP32 operator (near-field SpMV,
P2M, M2M, M2L, L2L, M2P, L2P) inside an FP64 outer loop with FP64 Arnoldi,
FP64 residuals and FP64 singular quadrature.  The leaf size (max nodes per
leaf box) is the one the project's autotuner picks (pibem_amd.tune_leaf_size:
whole solves timed at leaf sizes on a ladder from 8 to 160, walking from the
reference's default of 20 until the best is bracketed), for the mixed solve
and, separately, for the FP64 solve.  The FP64 GPU path is timed
beside it, as are FMM matvecs per second, per-phase device times from HIP
events on the library's stream, and the CPU oracle (the reference's algorithm
restated in C++, OpenMP) on the host cores.  Configuration 2 (unit sphere,
6 x 64 x 64 Q1 cells, 24,578 DoFs, order 6), the 64-problem batch
(configuration 4) and the 32 right-hand sides of the Wigley hull
(configuration 5) are measured the same way and reported in the "config2",
"config4" and "config5" objects.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--extra 2,4,5]
    python bench.py --mode matvec --config {2,3} --precision {fp32,fp64} --steps K [--leaf L] [--plan P]
    python bench.py --mode copy --mib M --steps K
    python bench.py --mode solve --config {2,3} --precision {fp32,fp64} --steps K [--leaf L]

The default mode prints one JSON line.  "matvec" only runs K FMM matvecs
(separate N and D outputs) followed by K operator applications, and
"solve" K GMRES solves (fp32: the mixed-precision solve), short programs to
run under rocprofv3; "copy" runs K device-to-device copies of M MiB, the
calibration for the HBM byte counters.
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

METRIC = "GMRES time to rel-res 1e-8, mixed vs FP64 GPU; FMM matvecs/s; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFS = 157.3
FP64_PEAK_TFS = 78.6
PHASES = ("spmv", "p2m", "m2m", "l2l", "m2l", "m2p", "l2p", "m2l_sum")
# kernel names as they appear in traces, per phase, operator precision and
# number of expansions per block (2: N and D separately, 1: the operator's
# combined N.phi - D.dphi)
SPMV_KERNEL = {0: "k_spmv_lds<{t}, {k}, false>", 1: "k_spmv_wave<{t}, {k}, 2, false>",
               2: "k_spmv_wave<{t}, {k}, 4, false>", 3: "k_spmv_lds<{t}, {k}, true>",
               4: "k_spmv_wave<{t}, {k}, 2, true>", 5: "k_spmv_wave<{t}, {k}, 4, true>",
               6: "k_spmv_cols<float, 8>"}
SPMV_KERNEL_FP64 = {3: "k_spmv_cols<double, 8>"}  # the FP64 variant numbers past 2
KERNEL = {
    "p2m": "k_p2m<{t}, {k}>", "m2m": "k_translate<{t}, {p}, 0, {k}>", "m2l": "k_translate<{t}, {p}, 1, {k}>",
    "l2l": "k_translate<{t}, {p}, 2, {k}>", "m2p": "k_m2p<{t}, {k}, {p}>", "m2l_sum": "k_m2l_reduce<{t}>",
}
# M2L into per-pair slots (the slot plans; its per-target sum is m2l_sum)
KERNEL_M2L_SLOT = "k_translate<{t}, {p}, 3, {k}>"


def kernel_name(phase, prec, p, nk, tuning):
    """The trace name of the kernel that runs a phase under the tuned
    choices (SpMV variant, M2L plan: bit 4 = slots, bit 1 = M2P/L2P side
    fusion; the operator's L2P is fused with the post-processing)."""
    t = "float" if prec == "fp32" else "double"
    plan = int(tuning.get(f"side_fusion_{prec}", 0)) if nk == 1 else 0
    if phase == "spmv":
        v = int(tuning.get(f"spmv_kernel_{prec}", 0))
        fmt = (((SPMV_KERNEL_FP64 if prec == "fp64" and v > 2 else SPMV_KERNEL)[v]) if nk == 1
               else "k_spmv_lds/k_spmv_wave<{t}, {k}> (variant timed for the matvec)")
    elif phase == "l2p":
        fmt = "k_l2p_post<{t}, {p}>" if nk == 1 else "k_l2p<{t}, {k}, {p}>"
    elif phase == "m2l" and plan & 4:
        fmt = KERNEL_M2L_SLOT
    else:
        fmt = KERNEL[phase]
    name = fmt.format(t=t, p=p, k=nk)
    if plan & 1 and phase in ("m2l", "m2p", "l2p"):
        name += " (side-fused plan: k_translate_side)"
    return name


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def work_model(wc, s_v, nk):
    """Algorithmic bytes and flops of one launch of each phase (s_v: bytes of
    an operator value, 4 for the FP32 operator, 8 for FP64; nk: 2 for the FMM
    matvec with separate N.phi and D.dphi outputs, 1 for the operator GMRES
    applies, whose near field and unit multipoles hold one value per entry (N
    on Neumann, -D on Dirichlet columns) and whose expansions are combined).

    spmv: nk values per leaf-block entry, one column index per leaf column, one
          row id per row, nk input vectors read once, nk outputs written once.
    p2m:  2 kp unit-multipole values per leaf column (nk 2), or the
          operator's compact kpc (its real parts and the imaginary parts of
          m >= 1, padded to a multiple of 4), one DoF index per column, nk
          input vectors once, nk kp values per leaf written.
    m2m, l2l, m2l: 8 Kt^2 flops per (pair, expansion): a complex Kt x Kt map.
    m2p, l2p: ~(8 nk Kt + 8 Kt + 40) flops per evaluation point.
    m2l_sum (slot plans): every pair's 2 Kt slot values read once and its
          4-byte pair index; the targets' writes (at most the expansion
          array) are left out.
    """
    n, p, kp = wc["n_dofs"], wc["trunc_order"], wc["kp"]
    kt = (p + 1) * (p + 2) // 2
    kpc = (2 * kt - (p + 1) + 3) // 4 * 4
    out = {
        "spmv": dict(bound="hbm", bytes=wc["near_values"] * nk * s_v + wc["near_cols"] * 4 + wc["near_rows"] * 4
                     + nk * n * 8 + nk * n * 8),
        "p2m": dict(bound="hbm", bytes=wc["p2m_cols"] * ((2 * kp if nk == 2 else kpc) * s_v + 4) + nk * n * 8
                    + wc["leaves"] * nk * kp * s_v),
    }
    for k, pairs in (("m2m", wc["m2m_translations"]), ("l2l", wc["l2l_translations"]), ("m2l", wc["m2l_pairs"])):
        out[k] = dict(bound="mfma", flops=pairs * nk * 8 * kt * kt)
    out["m2p"] = dict(bound="valu", flops=wc["m2p_evals"] * (8 * nk * kt + 8 * kt + 40))
    out["m2l_sum"] = dict(bound="hbm", bytes=wc["m2l_pairs"] * (2 * kt * s_v + 4))
    out["l2p"] = dict(bound="valu", flops=wc["near_rows"] * (8 * nk * kt + 8 * kt + 40))
    return out


def roofline_of(phase, model, ms, prec, name):
    m = model[phase]
    if ms <= 0:
        return dict(kernel=name, bound=m["bound"], achieved=0.0, peak=None, unit=None, frac=0.0,
                    algorithmic_per_launch=None, launch_ms=0.0)
    if m["bound"] == "hbm":
        ach = m["bytes"] / (ms * 1e-3) / 1e9
        return dict(kernel=name, bound="hbm", achieved=ach, peak=HBM_PEAK_GBS,
                     unit="GB/s", frac=ach / HBM_PEAK_GBS, algorithmic_per_launch=m["bytes"], launch_ms=ms)
    peak = FP32_PEAK_TFS if prec == "fp32" else FP64_PEAK_TFS
    ach = m["flops"] / (ms * 1e-3) / 1e12
    return dict(kernel=name, bound="mfma" if m["bound"] == "mfma" else "valu",
                achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak, algorithmic_per_launch=m["flops"],
                launch_ms=ms)


def traffic_for(kernel, workload):
    """HBM bytes per launch of this kernel from the committed rocprofv3
    counter summary of the workload (profiles/traffic_*.json: the workload
    and, per kernel, its bytes), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for f in sorted(os.listdir(d)):
        if f.startswith("traffic_") and f.endswith(".json"):
            try:
                t = json.load(open(os.path.join(d, f)))
            except (OSError, ValueError):
                continue
            if t.get("workload") == workload and kernel in t.get("kernels", {}):
                return t["kernels"][kernel].get("traffic_bytes")
    return None


def make_problem(config, leaf=None):
    """Configuration 2 or 3; leaf: max_nodes_per_leaf (None: the reference's
    default of 20)."""
    import dataclasses
    from pibem_amd import mesh as M
    pr = M.config_sphere(64, 6) if config == 2 else M.config_spheroid(128, 8)
    return dataclasses.replace(pr, max_nodes_per_leaf=int(leaf)) if leaf else pr


def tune_leaves(config):
    """The leaf-size autotuner (pibem_amd.tune_leaf_size: whole solves timed,
    a walk over the ladder 8..160 from the reference's 20 until the best is
    bracketed) for the mixed and the FP64 solve of a configuration."""
    import pibem_amd as P
    pr = make_problem(config)
    out = {}
    for name, prec in (("mixed", P.PBFMM_MIXED), ("fp64", P.PBFMM_FP64)):
        best, times, info, br = P.tune_leaf_size(pr, precision=prec, solves=2)
        log(f"config {config} leaf tuning, {name}: " + " ".join(f"{k}:{v:.2f}ms" for k, v in sorted(times.items()))
            + f" -> {best}" + ("" if br else " (not bracketed)"))
        out[name] = dict(best=best, bracketed=br, ms_per_solve={str(k): v for k, v in sorted(times.items())},
                         steps={str(k): info[k][0] for k in sorted(info)},
                         setup_s={str(k): info[k][2] for k in sorted(info)})
    return out


def setup_handle(pr, stream):
    import pibem_amd as P
    t0 = time.perf_counter()
    f = P.BEMFMA(trunc_order=pr.p, num_octree_levels=pr.levels, max_nodes_per_leaf=pr.max_nodes_per_leaf,
                 quad_order=pr.quad_order, sing_quad_order=pr.sing_quad_order, precision=P.PBFMM_MIXED,
                 stream=stream.h, sing_rule=pr.sing_rule)
    f.init_fma(pr.mesh, pr.dirichlet)
    f.generate_octree_blocking()
    t1 = time.perf_counter()
    f.direct_integrals()
    t2 = time.perf_counter()
    f.multipole_integrals()
    bp = P.BEMProblem(f)
    bp.compute_alpha()
    t3 = time.perf_counter()
    return f, bp, dict(octree_s=t1 - t0, direct_s=t2 - t1, multipole_alpha_s=t3 - t2)


def time_solves(f, bp, pr, stream, steps, warmup):
    import pibem_amd as P
    from pibem_amd import hip
    n = pr.mesh.n_dofs
    rhs = hip.DeviceArray.from_numpy(bp.compute_rhs(pr.tmp_rhs))
    x = hip.DeviceArray(n)
    zero = hip.DeviceArray(n)
    xs = pr.mesh.support_xyz
    neu = pr.dirichlet == 0
    phi_ex = pr.exact.phi(xs)
    dn_ex = (pr.exact.grad(xs) * pr.mesh.normals).sum(1)
    res = {}
    for name, prec in (("mixed", P.PBFMM_MIXED), ("fp64", P.PBFMM_FP64)):
        def solve():
            x.copy_from(zero)
            return f.gmres_dev(rhs.value, x.value, 1e-8, 100, 2000, prec)
        for _ in range(warmup):
            solve()
        stream.synchronize()
        t = time.perf_counter()
        its = rr = None
        for _ in range(steps):
            its, rr = solve()
        stream.synchronize()
        dt = (time.perf_counter() - t) / steps
        sol = x.numpy()
        err_phi = float(np.linalg.norm((sol - phi_ex)[neu]) / np.linalg.norm(phi_ex[neu]))
        err_dn = float(np.linalg.norm((sol - dn_ex)[~neu]) / np.linalg.norm(dn_ex[~neu])) if (~neu).any() else 0.0
        res[name] = dict(ms=dt * 1e3, iters=its, relres=rr, err_phi_neumann=err_phi, err_dphi_dirichlet=err_dn)
        log(f"{name}: {dt * 1e3:.2f} ms per solve, {its} operator applications, relres {rr:.3e}, "
            f"error vs analytic phi {err_phi:.2e} dphi/dn {err_dn:.2e}")
    return res


def time_phases(f, call, reps):
    f.set_phase_timing(True)
    acc = None
    for _ in range(reps):
        call()
        t = f.last_timings()
        acc = t if acc is None else {k: acc[k] + t[k] for k in t}
    f.set_phase_timing(False)
    return {k: v / reps for k, v in acc.items()}


def time_operator(f, bp, n, stream, reps=50, phase_reps=10):
    """BEMProblem::vmult on device vectors (the operator GMRES applies, whose
    FMM carries one combined expansion per block): rate and per-phase device
    times, seeded U(-1,1) input (seed 1234)."""
    import pibem_amd as P
    from pibem_amd import hip
    src = hip.DeviceArray.from_numpy(np.random.default_rng(1234).uniform(-1, 1, n))
    dst = hip.DeviceArray(n)
    e0, e1 = hip.Event(), hip.Event()
    rate, phases = {}, {}
    for name, prec in (("fp32", P.PBFMM_MIXED), ("fp64", P.PBFMM_FP64)):
        for _ in range(5):
            bp.vmult_dev(src.value, dst.value, prec)
        e0.record(stream)
        for _ in range(reps):
            bp.vmult_dev(src.value, dst.value, prec)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        rate[name] = dict(ms=ms, per_s=1e3 / ms)
        phases[name] = time_phases(f, lambda: bp.vmult_dev(src.value, dst.value, prec), phase_reps)
        log(f"operator {name}: {ms:.3f} ms; phases " + " ".join(f"{k}={v:.3f}" for k, v in phases[name].items()))
    return rate, phases


def time_matvecs(f, n, stream, reps=50, phase_reps=10):
    """FMM matvec with separate N.phi and D.dphi outputs
    (multipole_matr_vect_products on device vectors): rate (hipEvents on the
    library stream) and per-phase device times; seeded U(-1,1) inputs (seed
    1234)."""
    import pibem_amd as P
    from pibem_amd import hip
    rng = np.random.default_rng(1234)
    a = hip.DeviceArray.from_numpy(rng.uniform(-1, 1, n))
    b = hip.DeviceArray.from_numpy(rng.uniform(-1, 1, n))
    oN, oD = hip.DeviceArray(n), hip.DeviceArray(n)
    e0, e1 = hip.Event(), hip.Event()
    mv, phases, outs = {}, {}, {}
    for name, prec in (("fp32", P.PBFMM_MIXED), ("fp64", P.PBFMM_FP64)):
        for _ in range(5):
            f.matvec_dev(a.value, b.value, oN.value, oD.value, prec)
        e0.record(stream)
        for _ in range(reps):
            f.matvec_dev(a.value, b.value, oN.value, oD.value, prec)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        mv[name] = dict(ms=ms, per_s=1e3 / ms)
        outs[name] = (oN.numpy(), oD.numpy())
        phases[name] = time_phases(f, lambda: f.matvec_dev(a.value, b.value, oN.value, oD.value, prec), phase_reps)
        log(f"matvec {name}: {ms:.3f} ms; phases " + " ".join(f"{k}={v:.3f}" for k, v in phases[name].items()))
    rel = {k: float(np.linalg.norm(outs["fp32"][i] - outs["fp64"][i]) / np.linalg.norm(outs["fp64"][i]))
           for i, k in enumerate(("N", "D"))}
    return mv, phases, rel


WORKLOAD = {
    2: "config 2: unit sphere, cube-projected 6x64x64 Q1 cells, 24,578 DoFs, FMM order 6, "
       "phi = 1/|x - (2,1,1)|, Duffy-split singular rule, GMRES(100) to true rel-res 1e-8, leaf size (max nodes per "
       "leaf box) chosen by the project's autotuner from timed solves",
    3: "config 3: prolate spheroid (4,1,1), 6x128x128 Q2 cells (98,304) on an equiangular cube map "
       "(c -> tan(pi c/4) before projection), 393,218 DoFs, FMM order 8, phi = x + 8 seeded point "
       "sources (seed 3), Duffy-split singular rule, GMRES(100) to true rel-res 1e-8, mixed precision (FP32 "
       "operator; FP64 singular quadrature, Arnoldi, residual and outer loop), leaf size (max nodes per leaf box) "
       "chosen by the project's autotuner from timed solves",
}


def config_summary(c, workload, leaves=None):
    """The per-configuration object: solves, rates, phases, rooflines.  With
    the leaf tuner's results (leaves), the FP64 solve also at its own tuned
    leaf size, and mixed against FP64 each at its best."""
    r = c["roofline"]
    mixed_ms = c["solves"]["mixed"]["ms"]
    fp64_best = None
    if leaves:
        lf = leaves["fp64"]
        fp64_best = dict(leaf=lf["best"], ms=lf["ms_per_solve"][str(lf["best"])])
    return {
        "workload": workload,
        "n_dofs": c["wc"]["n_dofs"],
        "max_nodes_per_leaf": c["pr"].max_nodes_per_leaf,
        "mixed_ms_per_solve": mixed_ms,
        "fp64_ms_per_solve": c["solves"]["fp64"]["ms"],
        "mixed_speedup_vs_fp64": c["solves"]["fp64"]["ms"] / mixed_ms,
        "fp64_at_its_tuned_leaf": fp64_best,
        "mixed_speedup_vs_fp64_each_tuned": fp64_best["ms"] / mixed_ms if fp64_best else None,
        "leaf_tuning": leaves,
        "solves": c["solves"],
        "operator_applications_per_s": {k: v["per_s"] for k, v in c["operator"].items()},
        "fmm_matvecs_per_s": {k: v["per_s"] for k, v in c["matvec"].items()},
        "spmv_gbs": {k: r[k]["spmv"]["achieved"] for k in r},
        "m2l_tflops": {k: r[k]["m2l"]["achieved"] for k in r},
        "fp32_vs_fp64_matvec": c["fp32_vs_fp64_matvec"],
        "operator_phase_ms": c["operator_phases"],
        "matvec_phase_ms": c["phases"],
        "phase_roofline": {k: {ph_: {kk: v[kk] for kk in ("kernel", "achieved", "unit", "frac")}
                               for ph_, v in r[k].items()} for k in r},
        "setup_s": c["setup"],
        "work_counts": c["wc"],
        "tuned_choices": c["tuning"],
    }


def headline(c, args, leaves=None):
    """The JSON line's contract fields for configuration c (mixed solve),
    with the roofline of the dominant kernel of its FP32 operator."""
    ph = c["operator_phases"]["fp32"]
    dom = max(PHASES, key=lambda k: ph[k])
    roof = dict(c["roofline"]["fp32"][dom])
    roof["traffic"] = traffic_for(roof["kernel"], f"config 3, leaf {c['pr'].max_nodes_per_leaf}")
    hl = dict(bound=roof["bound"], achieved=roof["achieved"], peak=roof["peak"], unit=roof["unit"],
              frac=roof["frac"], traffic=roof["traffic"], kernel=roof["kernel"],
              algorithmic_per_launch=roof["algorithmic_per_launch"], launch_ms=roof["launch_ms"])
    wc = c["wc"]
    out = {
        "metric": METRIC,
        "value": 1e3 / c["solves"]["mixed"]["ms"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c["solves"]["mixed"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD[3], "n_dofs": wc["n_dofs"], "trunc_order": wc["trunc_order"],
                   "max_nodes_per_leaf": c["pr"].max_nodes_per_leaf, "parallelism": "single GPU"},
        "roofline": hl,
    }
    summ = config_summary(c, WORKLOAD[3], leaves)
    for k in ("workload", "n_dofs", "max_nodes_per_leaf"):
        summ.pop(k)
    out.update(summ)
    return out


def measure_config(config, steps, warmup, leaf=None):
    from pibem_amd import hip
    pr = make_problem(config, leaf)
    n = pr.mesh.n_dofs
    log(f"config {config}: N={n} cells={pr.mesh.n_cells} p={pr.p} leaf {pr.max_nodes_per_leaf}")
    stream = hip.Stream()
    f, bp, setup = setup_handle(pr, stream)
    log("setup: " + " ".join(f"{k}={v:.2f}" for k, v in setup.items()))
    wc = f.work_counts()
    solves = time_solves(f, bp, pr, stream, steps, warmup)
    tuning = f.tuning_state()
    g = f.graph_timings()
    tuning["graph_step_ms"] = {"fp64_direct": g[0], "fp64_graph": g[1], "fp32_direct": g[2], "fp32_graph": g[3]}
    log("tuned choices: " + " ".join(f"{k}={v}" for k, v in tuning.items()))
    op_rate, op_phases = time_operator(f, bp, n, stream)
    mv, phases, rel = time_matvecs(f, n, stream)
    roof = {prec: {ph: roofline_of(ph, work_model(wc, 4 if prec == "fp32" else 8, 1), op_phases[prec][ph], prec,
                                   kernel_name(ph, prec, pr.p, 1, tuning)) for ph in PHASES}
            for prec in op_phases}
    roof2 = {prec: {ph: roofline_of(ph, work_model(wc, 4 if prec == "fp32" else 8, 2), phases[prec][ph], prec,
                                    kernel_name(ph, prec, pr.p, 2, tuning)) for ph in PHASES} for prec in phases}
    return dict(pr=pr, f=f, stream=stream, wc=wc, setup=setup, solves=solves, tuning=tuning, operator=op_rate,
                operator_phases=op_phases, matvec=mv, phases=phases, fp32_vs_fp64_matvec=rel, roofline=roof,
                roofline_matvec=roof2)


def measure_multi_rhs(steps):
    """Configuration 5: the Wigley double body, 32 pure-Neumann right-hand
    sides solved in lockstep through the multi-RHS entry point (each its own
    Arnoldi process, every operator application serving all 32), mixed and
    FP64."""
    import pibem_amd as P
    from pibem_amd import hip
    from pibem_amd import mesh as M
    pr = M.config_wigley()
    n, R = pr.mesh.n_dofs, pr.rhs_set.shape[0]
    log(f"config 5: N={n} cells={pr.mesh.n_cells} p={pr.p} rhs={R}")
    stream = hip.Stream()
    f, bp, setup = setup_handle(pr, stream)
    cons = M.constraints(pr)
    bp.set_constraints(cons)
    rhs = np.stack([bp.compute_rhs(g) for g in pr.rhs_set])
    rhs[:, cons[4] == 1] = cons[3][cons[4] == 1]
    drhs, dx = hip.DeviceArray.from_numpy(rhs.ravel()), hip.DeviceArray(n * R)
    zero = np.zeros(n * R)
    out = dict(workload="config 5: closed Wigley double body L=1, B=0.1, T=0.0625, 2 x 400 x 100 Q1 cells, "
                        "81,002 DoFs with stem/keel double nodes, FMM order 8, 32 pure-Neumann right-hand sides "
                        "-U_k.n (seed 5) solved in lockstep, each to true rel-res 1e-8",
               n_dofs=n, n_rhs=R, setup_s=setup)
    for name, prec in (("mixed", P.PBFMM_MIXED), ("fp64", P.PBFMM_FP64)):
        dx.upload(zero)
        f.gmres_dev_multi(drhs.value, dx.value, R, n, 1e-8, 100, 2000, prec)  # warm-up
        times = []
        for _ in range(max(1, steps)):
            dx.upload(zero)
            stream.synchronize()
            t = time.perf_counter()
            its, rr = f.gmres_dev_multi(drhs.value, dx.value, R, n, 1e-8, 100, 2000, prec)
            stream.synchronize()
            times.append(time.perf_counter() - t)
        X = dx.numpy().reshape(R, n)
        err = max(float(np.linalg.norm((X[r] - X[r].mean()) - (e - e.mean())) / np.linalg.norm(e - e.mean()))
                  for r, e in enumerate(pr.exact_set[k].phi(pr.mesh.support_xyz) for k in range(R)))
        ms = min(times) * 1e3
        out[name] = dict(ms_all_rhs=ms, ms_per_rhs=ms / R, arnoldi_steps_all_rhs=int(its.sum()),
                         max_arnoldi_steps=int(its.max()), max_relres=float(rr.max()), max_error_vs_analytic=err)
        log(f"config 5 {name}: {ms:.1f} ms for {R} right-hand sides ({ms / R:.2f} ms each), Arnoldi steps "
            f"{int(its.sum())} in all, at most {int(its.max())} for one; max relres {rr.max():.2e}, max error vs "
            f"-U.x + C {err:.2e}")
    out["mixed_speedup_vs_fp64"] = out["fp64"]["ms_all_rhs"] / out["mixed"]["ms_all_rhs"]
    return out


def _batch_solves(probs, leaf, timed=True, precisions=("mixed", "fp64")):
    """Set up probs as one batch at the given leaf size and solve every
    problem to its own true relative residual 1e-8; returns set-up seconds
    and, per precision, ms, operator applications, max relres and the max
    error against each problem's exact solution."""
    import dataclasses
    import pibem_amd as P
    from pibem_amd import hip
    from pibem_amd import mesh as M
    probs = [dataclasses.replace(p, max_nodes_per_leaf=int(leaf)) for p in probs]
    t = time.perf_counter()
    b = P.setup_batch(probs, P.PBFMM_MIXED)
    bb = P.BEMProblem(b)
    bb.compute_alpha()
    cons = []
    for pr in probs:
        if pr.mesh.dbl_ptr is not None and int(np.diff(pr.mesh.dbl_ptr).max()) > 1:
            cons.append(M.constraints(pr))
        else:
            n = pr.mesh.n_dofs
            cons.append((np.zeros(n + 1, np.int64), np.zeros(0, np.int64), np.zeros(0), np.zeros(n),
                         np.zeros(n, np.uint8)))
    mc = M.merge_constraints(cons, b.offsets[:-1])
    bb.set_constraints(mc)
    rhs_h = bb.compute_rhs(np.concatenate([pr.tmp_rhs for pr in probs]))
    rhs_h[mc[4] == 1] = mc[3][mc[4] == 1]
    hip.synchronize()
    out = dict(setup_s=time.perf_counter() - t)
    rhs, x = hip.DeviceArray.from_numpy(rhs_h), hip.DeviceArray(b.n)
    for name in precisions:
        prec = P.PBFMM_MIXED if name == "mixed" else P.PBFMM_FP64
        # one untimed solve first: the first launch of each kernel loads its
        # code object, which would otherwise land in the timed solve
        x.upload(np.zeros(b.n))
        b.gmres_batch_dev(rhs.value, x.value, 1e-8, 100, 2000, prec)
        x.upload(np.zeros(b.n))
        hip.synchronize()
        t = time.perf_counter()
        it, rr = b.gmres_batch_dev(rhs.value, x.value, 1e-8, 100, 2000, prec)
        hip.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        X = x.numpy()
        err = 0.0
        for s, pr in enumerate(probs):
            xs = X[b.offsets[s]:b.offsets[s + 1]]
            ex = pr.exact.phi(pr.mesh.support_xyz)
            neu = pr.dirichlet == 0
            err = max(err, float(np.linalg.norm((xs - ex)[neu]) / np.linalg.norm(ex[neu])))
        out[name + "_ms"] = ms
        out[name + "_ops"] = it
        out[name + "_max_relres"] = float(rr.max())
        out[name + "_max_error"] = err
    b.close()
    return out


def measure_batch(tune=True):
    """Configuration 4: 64 independent problems of uneven size (seed 4) as
    one batch (pbfmm_batch_create: a forest of octrees, one launch per phase
    for all problems), solved with every problem's own relative residual
    <= 1e-8 (mixed and FP64), each against its own exact solution.  With
    tune, the leaf-size autotuner's walk (pibem_amd.leaf_walk) times mixed
    batch solves of every eighth problem at each leaf size it visits, and the
    whole batch is solved at the reference's leaf size 20 and at the one
    chosen."""
    import pibem_amd as P
    from pibem_amd import mesh as M
    t = time.perf_counter()
    probs = M.config_batch()
    t_mesh = time.perf_counter() - t
    ntot = sum(p.mesh.n_dofs for p in probs)
    log(f"config 4: {len(probs)} problems, {ntot} DoFs in total (meshes {t_mesh:.1f} s)")
    out = dict(problems=len(probs), n_dofs_total=ntot)
    leaf = 20
    if tune:
        sub = probs[::8]
        walk = {}

        def cost(lf):
            r = _batch_solves(sub, lf, precisions=("mixed",))
            walk[lf] = r["mixed_ms"]
            return r["mixed_ms"]

        leaf, br = P.leaf_walk(cost)
        log("config 4 leaf tuning on every eighth problem, mixed: "
            + " ".join(f"{k}:{v:.2f}ms" for k, v in sorted(walk.items())) + f" -> {leaf}"
            + ("" if br else " (not bracketed)"))
        out["leaf_tuning"] = dict(sample=f"{len(sub)} problems (every eighth), mixed batch solves", best=leaf,
                                  bracketed=br, ms_per_solve={str(k): v for k, v in sorted(walk.items())})
    for lf in sorted({20, leaf}):
        r = _batch_solves(probs, lf)
        if lf == leaf:
            out.update(r)
            out["max_nodes_per_leaf"] = lf
        if lf == 20:
            out["at_leaf_20"] = r
        log(f"config 4, leaf {lf}: set-up {r['setup_s']:.1f} s; solves mixed {r['mixed_ms']:.0f} ms "
            f"({r['mixed_ops']} operator applications), fp64 {r['fp64_ms']:.0f} ms ({r['fp64_ops']}); max relres "
            f"{r['mixed_max_relres']:.2e}, max error vs analytic phi {r['mixed_max_error']:.2e}")
    return dict(workload="config 4: 64 problems (seed 4) of 1.5k-25k Q1 DoFs: spheres, spheroids of aspect 1-6, "
                         "double-node cubes, each with its own 8-source exterior solution, FMM order 6, as one batch "
                         "(one launch per phase for all problems), every problem to its own true rel-res 1e-8, leaf "
                         "size chosen by the project's autotuner",
                **out)


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return dict(model=model, nproc=os.cpu_count() or 1, available=avail)


def cpu_baseline(pr, ops, label):
    """The oracle's FP64 vmult (the reference algorithm restated in
    C++/OpenMP: the same octree and lists, naive O(p^4) translations,
    per-(leaf, cell) unit multipoles) on the host cores.  One call is the
    sample (at configuration 3 about 25-50 s of CPU time on 8-16 cores); the
    per-solve figure scales it by the operator applications of the FP64 GPU
    solve, since a whole CPU solve would take ~30 min.  Threads: every CPU
    this process may use, capped by OMP_NUM_THREADS when set (the GPU host
    allots each GPU a share of its cores)."""
    import oracle
    cpu = host_cpu()
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(cpu["available"], cap) if cap > 0 else cpu["available"]
    oracle.set_threads(threads)
    t = time.perf_counter()
    o = oracle.OracleFMM(pr).setup()
    o.compute_alpha()
    setup_s = time.perf_counter() - t
    v = np.random.default_rng(7).uniform(-1, 1, pr.mesh.n_dofs)
    t = time.perf_counter()
    o.vmult(v)
    per = time.perf_counter() - t
    cpu_solve = per * ops
    log(f"cpu oracle ({cpu['model']}, {threads} of {cpu['nproc']} CPUs): set-up {setup_s:.1f} s, {per:.2f} s per "
        f"vmult -> {cpu_solve:.0f} s per solve ({ops} applications)")
    return dict(value=1.0 / cpu_solve, unit="solves/s", cores=threads, kind="port",
                sample=f"one oracle vmult at {label} (FP64 FMM operator, the reference algorithm restated in "
                       f"C++/OpenMP) timed: {per:.2f} s; scaled by the {ops} operator applications of the FP64 GPU "
                       f"solve (extrapolated: a whole CPU solve is not run)",
                s_per_solve=cpu_solve, s_per_vmult=per, setup_s=setup_s, cpu_model=cpu["model"],
                nproc=cpu["nproc"], cpus_available=cpu["available"])


def mode_matvec(args):
    import pibem_amd as P
    from pibem_amd import hip
    pr = make_problem(args.config, args.leaf or None)
    stream = hip.Stream()
    f, _, _ = setup_handle(pr, stream)
    n = pr.mesh.n_dofs
    rng = np.random.default_rng(1234)
    a = hip.DeviceArray.from_numpy(rng.uniform(-1, 1, n))
    b = hip.DeviceArray.from_numpy(rng.uniform(-1, 1, n))
    oN, oD = hip.DeviceArray(n), hip.DeviceArray(n)
    prec = P.PBFMM_MIXED if args.precision == "fp32" else P.PBFMM_FP64
    for _ in range(args.steps):
        f.matvec_dev(a.value, b.value, oN.value, oD.value, prec)
    bp = P.BEMProblem(f)
    # one fused application first: it times the launch plans (as the solves
    # do), so that the phase-timed runs below use the M2L form it chose; with
    # --plan the plan is set instead and no timing runs, so that every launch
    # of the run is a kernel alone on the device (per-kernel durations)
    if args.plan >= 0:
        f.set_side_fusion(prec, args.plan)
    bp.vmult_dev(a.value, oN.value, prec)
    # phase-timed, as the operator phases bench.py reports: every phase in its
    # own launch (the solve's fused launches are profiled with --mode solve)
    f.set_phase_timing(True)
    for _ in range(args.steps):
        bp.vmult_dev(a.value, oN.value, prec)
    f.set_phase_timing(False)
    stream.synchronize()
    log(f"{args.steps} {args.precision} FMM matvecs and {args.steps} operator applications at config "
        f"{args.config}, leaf {pr.max_nodes_per_leaf} done")


def mode_solve(args):
    import pibem_amd as P
    from pibem_amd import hip
    pr = make_problem(args.config, args.leaf or None)
    stream = hip.Stream()
    f, bp, _ = setup_handle(pr, stream)
    n = pr.mesh.n_dofs
    rhs = hip.DeviceArray.from_numpy(bp.compute_rhs(pr.tmp_rhs))
    x = hip.DeviceArray(n)
    zero = hip.DeviceArray(n)
    prec = P.PBFMM_MIXED if args.precision == "fp32" else P.PBFMM_FP64
    for _ in range(args.steps):
        x.copy_from(zero)
        its, rr = f.gmres_dev(rhs.value, x.value, 1e-8, 100, 2000, prec)
    stream.synchronize()
    log(f"{args.steps} {args.precision} solves at config {args.config}, leaf {pr.max_nodes_per_leaf} done: {its} "
        f"operator applications, relres {rr:.3e}")


def mode_copy(args):
    from pibem_amd import hip
    nb = args.mib * 1024 * 1024 // 8
    a = hip.DeviceArray.from_numpy(np.ones(nb))
    b = hip.DeviceArray(nb)
    for _ in range(args.steps):
        b.copy_from(a)
    hip.synchronize()
    log(f"{args.steps} device copies of {args.mib} MiB done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-leaf-tuning", action="store_true")
    ap.add_argument("--extra", default="2,4,5", help="further configurations to measure (comma list of 2, 4, 5)")
    ap.add_argument("--mode", choices=("full", "matvec", "solve", "copy"), default="full")
    ap.add_argument("--config", type=int, choices=(2, 3), default=3)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--leaf", type=int, default=0, help="matvec/solve modes: max nodes per leaf (0: 20)")
    ap.add_argument("--plan", type=int, default=-1, help="matvec mode: operator launch plan (-1: timed)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported; this build runs the BEM-FMM path on a "
              f"single MI355X", file=sys.stderr)
        sys.exit(2)
    from pibem_amd import hip
    if hip.device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.mode == "matvec":
        return mode_matvec(args)
    if args.mode == "copy":
        return mode_copy(args)
    if args.mode == "solve":
        return mode_solve(args)

    leaves3 = None if args.no_leaf_tuning else tune_leaves(3)
    c3 = measure_config(3, args.steps, args.warmup, leaves3["mixed"]["best"] if leaves3 else None)
    out = headline(c3, args, leaves3)
    cpu = None
    if not args.no_cpu:
        # the reference's own leaf size (20): the oracle restates the
        # reference, which has no leaf tuner; its FP64 step count there
        ops = int(leaves3["fp64"]["steps"]["20"]) if leaves3 else c3["solves"]["fp64"]["iters"]
        cpu = cpu_baseline(make_problem(3), ops + 2, "config 3 (leaf size 20, the reference's default)")
    out["cpu_baseline"] = cpu
    out["speedup_vs_cpu"] = out["value"] / cpu["value"] if cpu else None
    del c3
    extra = {int(x) for x in args.extra.split(",") if x.strip()}
    if 2 in extra:
        leaves2 = None if args.no_leaf_tuning else tune_leaves(2)
        c2 = measure_config(2, args.steps, args.warmup, leaves2["mixed"]["best"] if leaves2 else None)
        out["config2"] = config_summary(c2, WORKLOAD[2], leaves2)
        del c2
    if 5 in extra:
        out["config5"] = measure_multi_rhs(args.steps)
    if 4 in extra:
        out["config4"] = measure_batch(tune=not args.no_leaf_tuning)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
