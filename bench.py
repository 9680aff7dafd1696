"""Benchmark of the MI355X hot path on the 2047^2-cell ellipse workload.

python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (resident in HBM before timing): unit square, 2047 x 2047 cells,
free-slip walls, gz = -9.81, 256 seeded ellipses (full axes 5-60 cells,
log10 viscosity U[-3, 3] against 0, density 3000 +- 300), 3 x 3 jittered
markers per cell (37.7M markers, cell-sorted once before timing).

One step = the hot path of one Stokes solve:
  P2G of per-marker viscosity (geometric) and density (arithmetic) onto cell
  centres and vertices (2 fused calls), body force and residual F at u = p = 0,
  then the Stokes solve J dx = F to ||r||/||r0|| <= 1e-8 (FP64).
value = Stokes solves per second (whole GPU).  Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mdoodz7.0_amd"))

HBM_PEAK_GBS = 8000.0
# The solver configuration measured: the vertex-patch Chebyshev smoother fused
# into one kernel per step (smoother 3), penalty 700 and degree 12 on every
# level -- the fastest of the configurations swept on this workload.  With the
# FP32 V-cycle the sweep's optimum penalty is lower (300): the mixed-precision
# legs run their own optimum.
SOLVER = dict(smoother=3, penalty=700.0, cheb_degree=12, cheb_degree_coarse=12)
SOLVER_MIXED = dict(SOLVER, penalty=300.0)
# HBM bytes per launch of the fine-level smoother step (k_pcheb, FP64) on
# 2047^2 cells, rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a
# copy kernel of known size): profiles/hbm_traffic_pcheb_2047.txt
SMOOTHER_TRAFFIC_2047 = 415.1e6
METRIC = "Stokes solve time @1e-8 resid, 2047², plain vs learned PC; P2G Mmarkers/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=2047)
    ap.add_argument("--nper", type=int, default=3)
    ap.add_argument("--ellipses", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the untimed mixed-precision and learned solves reported beside "
                         "the FP64 one")
    ap.add_argument("--kernels", type=int, default=0, metavar="REPS",
                    help="profiling mode: only the P2G calls, the fine-level smoother and a "
                         "copy kernel (calibration of the HBM byte counters), REPS times each")
    ap.add_argument("--trace-solve", type=int, default=0, metavar="ITS",
                    help="profiling mode: P2G then ITS outer iterations of the solve with the "
                         "inner CG launched kernel by kernel (no HIP graphs), for kernel tracers")
    ap.add_argument("--rheology", type=int, default=0, metavar="REPS",
                    help="profiling mode: only the grid rheology update (NonNewtonianViscosityGrid "
                         "on the centres of the --cells grid, two visco-elasto-plastic phases), "
                         "REPS times, inputs resident in HBM")
    ap.add_argument("--markers", type=int, default=0, metavar="REPS",
                    help="profiling mode: only the marker kernels on the --cells workload's markers "
                         "(RK4 advection, centre and vertex interpolation, pressure and stress "
                         "updates), REPS times each, arrays resident in HBM")
    ap.add_argument("--pcnet", type=int, default=0, metavar="REPS",
                    help="profiling mode: only the learned network's forward pass and gradient "
                         "(BF16 MFMA convolutions) on a 1024 x 1024 image, REPS times each")
    ap.add_argument("--trace-mixed", action="store_true",
                    help="with --trace-solve: the FP32 V-cycle (MDG_PREC_MIXED) instead of FP64")
    ap.add_argument("--p2g-cases", type=int, default=0, metavar="REPS",
                    help="measurement mode: the P2G cases one field, two fused fields, phase "
                         "fractions (2 and 9 phases), the 9-cell stencil, and a P2G right after "
                         "the markers moved (re-sort and weight pass included), REPS times each")
    ap.add_argument("--solver", default="{}", metavar="JSON",
                    help="extra mdg_solver_opts fields for --trace-solve (e.g. '{\"smoother\": 3}')")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: the library runs on one GPU only (--gpus 1); multi-GPU sharding is "
              "designed, not built", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import mdgpu
    from mdgpu import learned
    from mdgpu import workloads as W

    if a.rheology:
        rheology_only(a.cells, a.rheology)
        return
    if a.markers:
        markers_only(a, a.markers)
        return
    if a.pcnet:
        pcnet_only(a.pcnet)
        return
    t0 = time.time()
    wl = W.ellipse_workload(a.cells, a.nper, a.seed, n_ell=a.ellipses)
    n = wl["x"].size
    Nx = wl["Nx"]
    nc, nv = (Nx - 1) ** 2, Nx * Nx
    log("workload: %d^2 cells, %d markers, built in %.1fs" % (a.cells, n, time.time() - t0))
    ctx = mdgpu.Context(Nx, Nx, 0.0, 1.0, 0.0, 1.0)
    ctx.set_bc(*learned.free_slip_bc(Nx))
    dx_ = ctx.device(n, host=wl["x"])
    dz_ = ctx.device(n, host=wl["z"])
    dph = ctx.device(n, np.int32, wl["phase"])
    t0 = time.perf_counter()
    ctx.markers_set(dx_, dz_, dph)
    t_sort = time.perf_counter() - t0
    eta_s = ctx.device(n)
    rho_s = ctx.device(n)
    ctx.markers_gather(wl["eta_table"][wl["phase"]], eta_s)
    ctx.markers_gather(wl["rho_table"][wl["phase"]], rho_s)
    weights_ms = []
    en, rn = ctx.device(nc, host=np.zeros(nc)), ctx.device(nc, host=np.zeros(nc))
    es, rs = ctx.device(nv, host=np.zeros(nv)), ctx.device(nv, host=np.zeros(nv))
    zc, zv = ctx.device(nc, host=np.zeros(nc)), ctx.device(nv, host=np.zeros(nv))
    rgx = ctx.device(Nx * (Nx + 1), host=np.zeros(Nx * (Nx + 1)))
    rgz = ctx.device(Nx * (Nx + 1), host=np.zeros(Nx * (Nx + 1)))
    neq, nm, ncont = ctx.eqn_count()
    F = ctx.device(neq, host=np.zeros(neq))
    sol = ctx.device(neq, host=np.zeros(neq))
    srt = 1 | mdgpu.MDG_SRC_SORTED
    f_cent = [dict(src=eta_s, dst=en, flag=srt, avg=2), dict(src=rho_s, dst=rn, flag=srt, avg=0)]
    f_vert = [dict(src=eta_s, dst=es, flag=srt, avg=2), dict(src=rho_s, dst=rs, flag=srt, avg=0)]
    D11 = ctx.device(nc)
    stats = {}

    def step():
        ctx.p2mastah_fused(1, 1, f_cent)
        k1 = ctx.last_kernel_ms()
        ctx.p2mastah_fused(0, 1, f_vert)
        k2 = ctx.last_kernel_ms()
        ctx.eval_rhs(0.0, -9.81, rn, rs, rgx, rgz)
        ctx.stokes_residual(zc, zc, zv, zc, zc, rgx, rgz, F=F)
        ctx.scale_into(en, 2.0, D11)
        o = ctx.killer_solver_dev(D11, D11, es, F, nm, sol, **SOLVER)
        stats.update(p2g_ms=[k1, k2], its=o.its, inner=o.inner_its, solve_ms=o.time_ms,
                     rel=o.rel_res, levels=o.levels)

    if a.kernels:
        kernels_only(ctx, a.kernels, step, f_cent, f_vert, D11, en, es, F, nm, sol, eta_s, rho_s, n)
        return
    if a.p2g_cases:
        p2g_cases(ctx, a.p2g_cases, wl, dx_, dz_, dph, eta_s, rho_s, en, rn, es, rs, n)
        return
    if a.trace_solve:
        ctx.p2mastah_fused(1, 1, f_cent)
        ctx.p2mastah_fused(0, 1, f_vert)
        ctx.eval_rhs(0.0, -9.81, rn, rs, rgx, rgz)
        ctx.stokes_residual(zc, zc, zv, zc, zc, rgx, rgz, F=F)
        ctx.scale_into(en, 2.0, D11)
        try:
            prec = mdgpu.MDG_PREC_MIXED if a.trace_mixed else mdgpu.MDG_PREC_FP64
            o = ctx.killer_solver_dev(D11, D11, es, F, nm, sol, max_its=a.trace_solve, graphs=1,
                                      precision=prec, **json.loads(a.solver))
        except mdgpu.MdgError as e:
            if e.code != -6:
                raise
            o = None
        print(json.dumps({"trace_solve_its": a.trace_solve,
                          "inner_its": o.inner_its if o else None}), flush=True)
        return
    # the first call per node set after markers_set caches the per-marker
    # weights (reported as weights_ms, outside the step)
    for f_set, cent in ((f_cent, 1), (f_vert, 0)):
        ctx.p2mastah_fused(cent, 1, f_set)
        weights_ms.append(ctx.last_weights_ms())
    for _ in range(a.warmup):
        step()
        log("warm-up step: solve %.1f s" % (stats["solve_ms"] / 1e3))
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        log("timed step %d of %d: solve %.1f s" % (i + 1, a.steps, stats["solve_ms"] / 1e3))
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / a.steps
    log("timed steps: %.1f ms per step" % ms_step)
    # the dominant kernel's duration inside a solve: the same FP64 solve once
    # more, launched kernel by kernel with HIP events around every fine-level
    # smoother step (graphs = 2); algorithmic bytes per launch from the library
    sol_t = ctx.device(neq, host=np.zeros(neq))
    o_t = ctx.killer_solver_dev(D11, D11, es, F, nm, sol_t, graphs=2, **SOLVER)
    k_ms_b2b, k_bytes = ctx.time_smoother(50)
    k_ms = o_t.smoother_ms
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    log("fine-level smoother in the solve: %d launches, %.4f ms each (back to back %.4f ms)"
        % (o_t.smoother_launches, k_ms, k_ms_b2b))
    # P2G kernels again, outside the timed steps: median of 10 launches each
    med = []
    for cent, fset in ((1, f_cent), (0, f_vert)):
        t = []
        for _ in range(10):
            ctx.p2mastah_fused(cent, 1, fset)
            t.append(ctx.last_kernel_ms())
        med.append(sorted(t)[len(t) // 2])
    p2g_ms = med
    # cached weight + 2 values per marker, 2 values per node, 4 uint32 quarter-cell offsets per cell
    p2g_bytes = 24.0 * n + 16.0 * nc + 16.0 * nc
    p2g_gbs = p2g_bytes / (p2g_ms[0] * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": 1e3 / ms_step, "unit": "Stokes solves/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Stokes solve to 1e-8 relative residual, %d^2 cells, %d seeded "
                               "ellipses (1e6 viscosity contrast), free slip, with P2G of "
                               "viscosity and density from %d markers" % (a.cells, a.ellipses, n),
                   "solver": "FGCR(100) + viscosity-weighted augmented Lagrangian (gamma 700) "
                             "+ CG preconditioned by a multigrid V-cycle with vertex-patch "
                             "Chebyshev smoothing (degree 12 on every level), FP64, plain "
                             "(no learned PC)", "solver_opts": SOLVER},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": SMOOTHER_TRAFFIC_2047 if a.cells == 2047 else None,
                     "kernel": "k_pcheb<double, false, 64, 16, 512> (fine-level fused vertex-patch "
                               "Chebyshev smoother step)",
                     "kernel_ms": k_ms, "alg_bytes": k_bytes,
                     "kernel_ms_source": "mean of the %d fine-level smoother launches of one FP64 "
                                         "solve (HIP events on the solver stream)" % o_t.smoother_launches,
                     "kernel_ms_back_to_back": k_ms_b2b},
        "solve": {"outer_its": stats["its"], "inner_its": stats["inner"],
                  "solve_ms": stats["solve_ms"], "rel_res": stats["rel"], "levels": stats["levels"]},
        "p2g": {"mmarkers_per_s": n / (p2g_ms[0] * 1e-3) / 1e6, "kernel_ms_centres": p2g_ms[0],
                "kernel_ms_vertices": p2g_ms[1], "achieved_gbs": p2g_gbs,
                "frac": p2g_gbs / HBM_PEAK_GBS, "fields": 2, "markers": n, "sort_s": t_sort,
                "weights_ms_centres": weights_ms[0], "weights_ms_vertices": weights_ms[1]},
    }
    if not a.no_variants:
        # same problem with the FP32 multigrid V-cycle (FP64 CG and FGCR
        # around it), after the timed steps; its error against the FP64 solve
        x64 = sol.numpy().copy()
        sol2 = ctx.device(neq, host=np.zeros(neq))
        o2 = ctx.killer_solver_dev(D11, D11, es, F, nm, sol2, precision=mdgpu.MDG_PREC_MIXED, **SOLVER_MIXED)
        x32 = sol2.numpy()
        log("mixed-precision solve: %.1f s" % (o2.time_ms / 1e3))
        res["mixed"] = {"precision": "FP32 multigrid V-cycle inside FP64 CG and FP64 FGCR",
                        "solver_opts": SOLVER_MIXED,
                        "solve_ms": o2.time_ms, "outer_its": o2.its, "inner_its": o2.inner_its,
                        "rel_res": o2.rel_res,
                        "rel_l2_vs_fp64": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
                        "speedup_vs_fp64": stats["solve_ms"] / o2.time_ms}
        res["learned"] = learned_variant(a, ctx, D11, es, F, nm, neq, x64, stats["solve_ms"])
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(wl, a, n)
    print(json.dumps(res), flush=True)


def learned_variant(a, ctx, D11, es, F, nm, neq, x64, fp64_ms):
    """The learned outer directions on the same 2047^2 problem: the network
    (U-Net, BF16 MFMA) is trained on the GPU on residual/error pairs that the
    plain solver records on 16 seeded 255^2 ellipse fields (the network is
    resolution-independent), then the 2047^2 problem is solved with precond 3
    (network correction, multigrid on the rest) in FP64 and with the FP32
    V-cycle.  Untimed by the step loop; training time reported."""
    import numpy as np
    import mdgpu
    from mdgpu import learned
    nt = 255
    tctx = mdgpu.Context(nt + 1, nt + 1, 0.0, 1.0, 0.0, 1.0)
    tctx.set_bc(*learned.free_slip_bc(nt + 1))
    n_ell = max(4, int(round(a.ellipses * (nt / a.cells) ** 2)))
    probs = [learned.EllipseProblem(tctx, nt, a.seed + 100 + i, n_ell) for i in range(16)]
    t0 = time.perf_counter()
    _, npairs = learned.collect_outer_pairs(tctx, probs, 3000, rel_tol=1e-8)
    log("learned: %d training pairs captured" % npairs)
    losses = learned.train_residual(tctx, 1000, batch=8, lr=1e-3, seed=3)
    log("learned: 1000 training steps, loss %.3f" % losses[-20:].mean())
    losses = np.concatenate([losses, tctx.pcnet_train_residual(2000, batch=8, lr=1e-3, seed=4)])
    log("learned: 3000 training steps, loss %.3f" % losses[-20:].mean())
    t_train = time.perf_counter() - t0
    ctx.pcnet_set_params(tctx.pcnet_get_params())
    out = {"training": {"pairs": npairs, "grid": "%d^2" % nt, "steps": 3000, "batch": 8,
                        "loss_first": float(losses[:20].mean()),
                        "loss_last": float(losses[-50:].mean()), "seconds": t_train}}
    for name, prec in (("fp64", mdgpu.MDG_PREC_FP64), ("mixed", mdgpu.MDG_PREC_MIXED)):
        sol3 = ctx.device(neq, host=np.zeros(neq))
        o3 = ctx.killer_solver_dev(D11, D11, es, F, nm, sol3, precond=3, precision=prec,
                                   **(SOLVER_MIXED if prec == mdgpu.MDG_PREC_MIXED else SOLVER))
        x3 = sol3.numpy()
        log("learned %s solve: %.1f s" % (name, o3.time_ms / 1e3))
        out[name] = {"solve_ms": o3.time_ms, "outer_its": o3.its, "inner_its": o3.inner_its,
                     "rel_res": o3.rel_res, "network_residual_ratio": o3.guess_gain,
                     "rel_l2_vs_fp64": float(np.linalg.norm(x3 - x64) / np.linalg.norm(x64)),
                     "speedup_vs_plain_fp64": fp64_ms / o3.time_ms}
    return out


def kernels_only(ctx, reps, step, f_cent, f_vert, D11, en, es, F, nm, sol, eta_s, rho_s, n):
    """The kernels whose rocprofv3 summaries are kept: P2G to centres and
    vertices, the fine-level smoother (after a one-iteration solve builds
    the hierarchy) and a plain copy y = 1.0 * x over the marker array (n
    doubles read, n written) that calibrates FETCH_SIZE / WRITE_SIZE."""
    import mdgpu
    out = {}
    t = []
    for _ in range(reps):
        ctx.p2mastah_fused(1, 1, f_cent)
        t.append(ctx.last_kernel_ms())
    out["p2g_centres_ms"] = sorted(t)[len(t) // 2]
    t = []
    for _ in range(reps):
        ctx.p2mastah_fused(0, 1, f_vert)
        t.append(ctx.last_kernel_ms())
    out["p2g_vertices_ms"] = sorted(t)[len(t) // 2]
    # the same two fields both arithmetically averaged (no per-marker log)
    srt = f_cent[0]["flag"]
    f_arith = [dict(src=eta_s, dst=f_cent[0]["dst"], flag=srt, avg=0),
               dict(src=rho_s, dst=f_cent[1]["dst"], flag=srt, avg=0)]
    t = []
    for _ in range(reps):
        ctx.p2mastah_fused(1, 1, f_arith)
        t.append(ctx.last_kernel_ms())
    out["p2g_centres_arith_ms"] = sorted(t)[len(t) // 2]
    ctx.scale_into(en, 2.0, D11)
    try:
        ctx.killer_solver_dev(D11, D11, es, F, nm, sol, max_its=1)
    except mdgpu.MdgError as e:
        if e.code != -6:
            raise
    out["smoother_ms"], out["smoother_bytes"] = ctx.time_smoother(reps)
    for _ in range(reps):
        ctx.scale_into(eta_s, 1.0, rho_s)
    out["copy_bytes"] = 16.0 * n
    out["markers"] = n
    print(json.dumps(out), flush=True)


def p2g_cases(ctx, reps, wl, dx_, dz_, dph, eta_s, rho_s, en, rn, es, rs, n):
    """Kernel times (HIP events, median of reps) of the P2G cases: one field
    and two fused fields onto centres and vertices, phase fractions with 2
    and 9 phases, the 9-cell stencil, and the cost of P2G after the markers
    moved (the re-sort and the per-node-set weight pass, which a step whose
    markers were advected pays once)."""
    import numpy as np
    import mdgpu

    def med(f):
        t = []
        for _ in range(reps):
            f()
            t.append(ctx.last_kernel_ms())
        return sorted(t)[len(t) // 2]

    srt = 1 | mdgpu.MDG_SRC_SORTED
    out = {"markers": n}
    for cen, dst, name in ((1, en, "centres"), (0, es, "vertices")):
        out["one_field_%s_ms" % name] = med(lambda: ctx.p2mastah_fused(cen, 1, [dict(src=rho_s, dst=dst, flag=srt)]))
        d2 = rn if cen == 1 else rs
        out["two_fields_%s_ms" % name] = med(lambda: ctx.p2mastah_fused(
            cen, 1, [dict(src=eta_s, dst=dst, flag=srt, avg=2), dict(src=rho_s, dst=d2, flag=srt, avg=0)]))
        out["stencil9_%s_ms" % name] = med(lambda: ctx.p2mastah_fused(cen, 9, [dict(src=rho_s, dst=dst, flag=srt)]))
    # phase fractions: the workload's 2 phases, then 9 phases assigned by marker index
    nc = ctx.node_count(1)
    for nph in (2, 9):
        if nph == 9:
            ph9 = np.where(wl["phase"] >= 0, np.arange(n) % 9, -1).astype(np.int32)
            ctx.markers_set(dx_, dz_, ctx.device(n, np.int32, ph9))
        perc = [ctx.device(nc) for _ in range(nph)]
        out["phase_fractions_%d_ms" % nph] = med(lambda: ctx.p2mastah(np.zeros(nph), prop=1, centroid=1,
                                                                      nb_phases=nph, phase_perc=perc))
    # markers moved: re-sort, then the first P2G of each node set computes
    # and caches the per-marker weights
    t_sort, t_w, t_p = [], [], []
    for _ in range(max(1, reps // 2)):
        t0 = time.perf_counter()
        ctx.markers_set(dx_, dz_, dph)
        t_sort.append((time.perf_counter() - t0) * 1e3)
        ctx.p2mastah_fused(1, 1, [dict(src=rho_s, dst=en, flag=srt)])
        t_p.append(ctx.last_kernel_ms())
        t_w.append(ctx.last_weights_ms())
    out["moved_resort_wall_ms"] = sorted(t_sort)[len(t_sort) // 2]
    out["moved_weights_ms"] = sorted(t_w)[len(t_w) // 2]
    out["moved_first_p2g_ms"] = sorted(t_p)[len(t_p) // 2]
    for k in list(out):
        if k.endswith("_ms") and not k.startswith("moved_resort"):
            out[k.replace("_ms", "_gmarkers_per_s")] = n / (out[k] * 1e-3) / 1e9
    print(json.dumps(out), flush=True)


def pcnet_only(reps, H=1024, W=1024):
    """The learned network alone: forward pass and loss gradient on an H x W
    image (random parameters and inputs), REPS times each.  Prints the wall
    time per call and the algorithmic convolution FLOPs (2 x pixels x 9 x
    Cin x Cout per layer) of one forward pass; the kernel times come from a
    kernel trace of the same command."""
    import numpy as np
    import mdgpu
    ctx = mdgpu.Context(H + 1, W + 1, 0.0, 1.0, 0.0, 1.0)
    rng = np.random.default_rng(5)
    ctx.pcnet_set_params(0.1 * rng.standard_normal(ctx.pcnet_param_count()).astype(np.float32))
    x = rng.standard_normal((H, W, 4)).astype(np.float32)
    t = rng.standard_normal((H, W, 4)).astype(np.float32)
    mask = np.ones((H, W, 4), np.float32)
    layers = [(4, 32, 0), (32, 32, 1), (32, 64, 2), (64, 64, 3), (64, 32, 2), (32, 32, 1), (32, 3, 0)]
    flops = sum(2.0 * (H >> l) * (W >> l) * 9 * ci * co for ci, co, l in layers)
    out = {"H": H, "W": W, "conv_gflop_forward": flops / 1e9}
    for name, f in (("forward", lambda: ctx.pcnet_forward(x)), ("grad", lambda: ctx.pcnet_grad(x, t, mask))):
        f()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        out[name + "_wall_ms"] = 1e3 * (time.perf_counter() - t0) / reps
    print(json.dumps(out), flush=True)


def rheology_only(cells, reps):
    """The grid rheology kernel alone: ViscosityConcise per centre and phase
    (power law n = 3 with a per-node prefactor, Maxwell elasticity,
    Drucker-Prager yield; 10% of the centres in the weaker phase), synthetic
    strain rates, pressures and old stresses around the yield point, every
    array on the device. Prints the wall time per call (the kernel time
    comes from a kernel trace of the same command)."""
    import numpy as np
    import mdgpu
    from mdgpu import setups as S
    nc = cells * cells
    from mdgpu import learned
    ctx = mdgpu.Context(cells + 1, cells + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(*learned.free_slip_bc(cells + 1))
    rng = np.random.default_rng(5)
    weak = rng.uniform(0.0, 1.0, nc) < 0.1
    exx = 1.0 + 0.2 * rng.standard_normal(nc)
    s0 = 0.2 + 0.05 * rng.standard_normal(nc)
    host = dict(perc=np.concatenate([(~weak).astype(float), weak.astype(float)]), G=np.ones(nc),
                T=np.ones(nc), P=0.5 + 0.1 * rng.standard_normal(nc), exx=exx, ezz=-exx,
                exz=0.1 * rng.standard_normal(nc), sxx0=s0, szz0=-s0, sxz0=np.zeros(nc),
                pref=np.exp(0.5 * rng.standard_normal(nc)))
    arrays = {k: ctx.device(v.size, host=v) for k, v in host.items()}
    arrays.update({k: ctx.device(nc) for k in mdgpu.RHEO_OUT})
    phases = [mdgpu.PhaseProp(**d) for d in S.vep_phases(C=1.5)]
    prm = mdgpu.RheoParams(average=0, step=1, elastic=1, dt=1.0, R=1.0, zeroC=0.0, min_eta=1e-4,
                           max_eta=1e4, eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
    ctx.viscosity_grid(phases, prm, 1, **arrays)
    t = []
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.viscosity_grid(phases, prm, 1, **arrays)
        t.append(ctx.last_kernel_ms())
    wall = (time.perf_counter() - t0) / reps
    k_ms = sorted(t)[len(t) // 2]
    # algorithmic bytes per centre: 2 phase fractions + 10 inputs read, 7 outputs written, 1 B tag
    nbytes = (12 + 7) * 8.0 * nc + nc
    yielding = int((arrays["eII_pl"].numpy() > 0).sum())
    print(json.dumps({"rheology_centres": nc, "phases": 2, "kernel_ms": k_ms,
                      "wall_ms_per_call": 1e3 * wall, "mnodes_per_s": nc / (k_ms * 1e-3) / 1e6,
                      "alg_bytes": nbytes, "achieved_gbs": nbytes / (k_ms * 1e-3) / 1e9,
                      "yielding_centres": yielding}), flush=True)


def markers_only(a, reps):
    """The marker kernels alone on the workload's markers (9 per cell):
    RK4 advection in a cellular flow, centre and vertex interpolation, the
    pressure and stress updates; every array on the device. Prints the
    median kernel time of each (HIP events) and the marker throughput."""
    import numpy as np
    import mdgpu
    from mdgpu import learned
    from mdgpu import workloads as W
    n = a.cells
    wl = W.ellipse_workload(n, a.nper, a.seed, n_ell=4)
    x, z = wl["x"], wl["z"]
    m = x.size
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, 0.0, 1.0)
    ctx.set_bc(*learned.free_slip_bc(n + 1))
    co = ctx.coords()
    xg, zg, xvz, zvx = co["xg"], co["zg"], co["xvz"], co["zvx"]
    u = (np.sin(np.pi * xg)[None, :] * np.pi * np.cos(np.pi * zvx)[:, None]).ravel()
    v = (-np.pi * np.cos(np.pi * xvz)[None, :] * np.sin(np.pi * zg)[:, None]).ravel()
    rng = np.random.default_rng(1)
    nc, nv = n * n, (n + 1) ** 2
    dev = lambda h: ctx.device(h.size, h.dtype, h)
    du, dv = dev(u), dev(v)
    dx_, dz_, dph = dev(x), dev(z), dev(np.zeros(m, np.int32))
    dvx, dvz, dout = (dev(np.zeros(m)) for _ in range(3))
    fc, fv = dev(rng.standard_normal(nc)), dev(rng.standard_normal(nv))
    grid = [dev(rng.standard_normal(k)) for k in (nc, nc, nv, nc, nc, nv)]
    mk = [dev(np.zeros(m)) for _ in range(4)]
    dt = 0.5 * (1.0 / n) / np.pi
    runs = {
        "advect_rk4": lambda: ctx.advect_rk4(du, dv, dt, dx_, dz_, dph, vx=dvx, vz=dvz),
        "g2p_centroids": lambda: ctx.g2p_centroids(fc, dx_, dz_, dph, out=dout),
        "g2p_vertices": lambda: ctx.g2p_vertices(fv, dx_, dz_, dph, out=dout),
        "update_particle_pressure": lambda: ctx.update_particle_pressure(grid[0], grid[3], dx_, dz_, dph, mk[0]),
        "update_particle_stress": lambda: ctx.update_particle_stress(*grid, dx_, dz_, dph, *mk[1:]),
    }
    out = {"markers": m, "cells": n}
    for name, f in runs.items():
        t = []
        for _ in range(reps):
            if name == "advect_rk4":
                dx_.upload(x); dz_.upload(z)
            f()
            t.append(ctx.last_kernel_ms())
        k = sorted(t)[len(t) // 2]
        out[name + "_ms"] = k
        out[name + "_gmarkers_per_s"] = m / (k * 1e-3) / 1e9
    print(json.dumps(out), flush=True)


def cpu_baseline(wl, a, n):
    """Oracle (CPU restatement of the reference path) on the host's cores,
    on windows of the very workload the GPU solves: the bottom-left w x w
    cells of the 2047^2 ellipse field (same markers, same ellipses, same cell
    size), for w = 127, 191, 255, each solved by KillerSolver (PH + GCR(6),
    exact banded Cholesky preconditioner) after the oracle P2Mastah.  The
    measured times are reported as they are; the 2047^2 figure is their
    power-law fit t = c w^p evaluated at 2047, reported with its exponent.
    The oracle P2Mastah rate is measured on the full marker set."""
    import numpy as np
    import oracle
    from oracle.driver import evaluate_residual, initial_fields
    nt = os.cpu_count() or 1
    h = 1.0 / a.cells
    em_all = wl["eta_table"][wl["phase"]]
    rm_all = wl["rho_table"][wl["phase"]]
    measured = []
    for w in (127, 191, 255):
        L = w * h
        sel = (wl["x"] < L) & (wl["z"] < L)
        x, z, ph = wl["x"][sel], wl["z"][sel], wl["phase"][sel]
        g = oracle.Grid(w + 1, w + 1, 0.0, L, 0.0, L)
        g.pure_shear(0.0)
        t0 = time.perf_counter()
        en = oracle.p2mastah(g, x, z, ph, em_all[sel], flag=1, avg=2, centroid=1, mode=0, nthreads=nt)
        es = oracle.p2mastah(g, x, z, ph, em_all[sel], flag=1, avg=2, centroid=0, mode=0, nthreads=nt)
        rn = oracle.p2mastah(g, x, z, ph, rm_all[sel], flag=1, avg=0, centroid=1, mode=0, nthreads=nt)
        rs = oracle.p2mastah(g, x, z, ph, rm_all[sel], flag=1, avg=0, centroid=0, mode=0, nthreads=nt)
        num = g.numbering()
        rx, rz = oracle.eval_rhs(g, 0.0, -9.81, rn, rs)
        u, v, p = initial_fields(g)
        F, _ = evaluate_residual(g, num, u, v, p, en, es, rx, rz)
        A, B, Cm = oracle.assemble_picard(g, num, en, es)
        nm = num[3]
        du, dp, st, rc = oracle.killer_solve(g, num, A, B, Cm, F[:nm], F[nm:], penalty=1e5,
                                              max_its_PH=30, rel_tol_KSP=1e-10, max_its_KSP=500,
                                              lin_abs_div=1e-12, lin_rel_div=1e-8, lin_abs_mom=1e-12,
                                              lin_rel_mom=1e-8)
        t = time.perf_counter() - t0
        measured.append({"cells": w, "s": t, "its_PH": st.its_PH, "its_KSP": st.its_KSP_tot, "status": rc})
        log("cpu baseline window %d^2: %.2f s (PH %d, KSP %d)" % (w, t, st.its_PH, st.its_KSP_tot))
    ws = np.log([m["cells"] for m in measured])
    ts = np.log([m["s"] for m in measured])
    p, c0 = np.polyfit(ws, ts, 1)
    t_proj = float(np.exp(c0 + p * np.log(a.cells)))
    # P2G rate on the full-size markers (one field to centres, interp_mode 0)
    big = oracle.Grid(wl["Nx"], wl["Nx"], 0.0, 1.0, 0.0, 1.0)
    t0 = time.perf_counter()
    oracle.p2mastah(big, wl["x"], wl["z"], wl["phase"], em_all, flag=1, avg=2, centroid=1, mode=0, nthreads=nt)
    t_p2g = time.perf_counter() - t0
    return {"value": 1.0 / t_proj, "unit": "Stokes solves/s", "cores": nt, "kind": "port",
            "sample": "oracle P2Mastah + KillerSolver (PH + GCR(6), exact banded Cholesky PC) on the "
                      "bottom-left 127^2, 191^2 and 255^2 windows of this 2047^2 workload (%s s measured on "
                      "%d threads of %s); value = 1 / (fit t = c w^%.2f at w = %d) = 1 / %.0f s; P2G alone: "
                      "%.1f Mmarkers/s on the full %d markers"
                      % ("/".join("%.2f" % m["s"] for m in measured), nt, cpu_model(), p, a.cells, t_proj,
                         n / t_p2g / 1e6, n),
            "measured": measured, "fit_exponent": float(p), "projected_s_2047": t_proj,
            "p2g_mmarkers_per_s": n / t_p2g / 1e6}


if __name__ == "__main__":
    main()
