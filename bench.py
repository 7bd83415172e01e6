"""Benchmark: ms per Poisson solve to 1e-6 relative residual on one MI355X.

Workload (default `--config c3`): the 5-level graded cut-cell octree around a
torus (R=0.15, r=0.05, z axis, centre (0.5,0.5,0.5)) with finest spacing 1/1024
(levels 3..7, finest band |d|<0.01 doubling per coarser level), 4x4 supersampled
face fractions, Neumann solid and walls, Dirichlet 0 at y=1, RHS = divergence of
N(0,1) face velocities (seed 2).  Plain MGPCG: PCG preconditioned by one FAS
V-cycle of 2+2 damped-Jacobi sweeps (w=2/3) per level, FP64 direct solve on
level 0, FP32 vectors, FP64 dot products; stop at 1e-6 relative residual.

One step = one solve from x = 0 with the right-hand side already resident in
HBM (restored on the device before each solve).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cirrus-amg_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

CONFIGS = {
    "c1": dict(name="C1 uniform 64^3 (512 tiles of 8^3), Dirichlet box, U[-1,1] seed 0 zero-mean",
               target=("uniform", 3, 3), assembly="plain", rhs=("uniform", 0)),
    "c2": dict(name="C2 3-level graded octree, finest 256^3, sphere r=0.2 Neumann solid, Dirichlet box, div N(0,1) seed 1",
               target=("bands", 3, 5, "sphere", (0.5, 0.5, 0.5, 0.2), {5: 0.05, 4: 0.15}), assembly="solid",
               rhs=("divergence", 1)),
    "c3": dict(name="C3 5-level graded cut-cell octree, finest 1024^3, torus R=0.15 r=0.05, 4x4 supersampled faces, "
                    "Dirichlet y=1, Neumann elsewhere, div N(0,1) seed 2",
               target=("bands", 3, 7, "torus", (0.5, 0.5, 0.5, 0.15, 0.05), {7: 0.01, 6: 0.02, 5: 0.04, 4: 0.08}),
               assembly="supersampled", rhs=("divergence", 2)),
}


def build(P, cfg):
    t = cfg["target"]
    if t[0] == "uniform":
        target = P.make_target(P.TARGET_UNIFORM, t[1], t[2])
    else:
        sk = P.SDF_SPHERE if t[3] == "sphere" else P.SDF_TORUS
        target = P.make_target(P.TARGET_SDF_BANDS, t[1], t[2], sk, t[4], t[5])
    g = P.HADeviceGrid.test_grid(target)
    if cfg["assembly"] == "plain":
        g.prepare_types_and_coeffs(11, 0.5)
    elif cfg["assembly"] == "solid":
        g.assemble_cutcell(P.SDF_SPHERE, t[4], -1, 11, 0.5)
    else:
        g.set_wall_types(t[1], True)
        g.assemble_cutcell(P.SDF_TORUS, t[4], 1, 11, 0.5)
    kind, seed = cfg["rhs"]
    if kind == "uniform":
        g.rhs_uniform(1, seed, True)
    else:
        g.rhs_divergence(1, 5, seed)
    return g


def canonical_bytes_per_iter(lv, ty, max_level, coef_bytes=16, coded=0, wide=0):
    """B_iter = 81 N_leaf + sum_l n_l (21+29+29.5+8.5+2*29) (tile-dense, FP32 vectors, four FP32
    coefficients = 16 B and the cell type 1 B per pass).  coef_bytes = 8: the cycle reads FP16
    coefficient copies, so each of the five cycle passes that read coefficients (first sweep,
    second sweep, residual + restriction, two post-sweeps) moves 8 B less per cell:
    sum_l n_l (13+21+21.5+8.5+2*21).  The CG part (81 B) keeps the FP32 operator.
    coded (bit mask of levels, FP32 form): on a coded level the four passes with a halo and the
    operator read one 4-B word (type and three face codes) in place of 17 B, with LEAF diagonals
    derived from the faces (NONLEAF tiles still read their 4-B diagonal): 13 B less per LEAF cell
    and pass, 9 B less per NONLEAF cell and pass.  wide (bit mask of the coded levels with
    ten-bit codes): the same for the four cycle passes, while the operator reads the FP32
    channels there."""
    n_leaf = int((ty == 1).sum()) * 512
    per_level = 21 + 29 + 29.5 + 8.5 + 2 * 29 - 5 * (16 - coef_bytes)
    levels = sum(int(((lv == l) & ((ty == 1) | (ty == 2))).sum()) * 512 * per_level for l in range(1, max_level + 1))
    coarsest = 512 * 8 * 512 + 512 * 8   # dense FP64 inverse of level 0 + its vectors
    saved = 0.0
    for l in range(max_level + 1):
        if coded >> l & 1:
            n_lf = int(((lv == l) & (ty == 1)).sum()) * 512
            n_nl = int(((lv == l) & (ty == 2)).sum()) * 512
            if l >= 1:
                saved += 4 * (13.0 * n_lf + 9.0 * n_nl)
            if not wide >> l & 1:
                saved += 13.0 * n_lf
    return 81.0 * n_leaf + levels + coarsest - saved


def oracle_grid(cfg):
    """The same workload built by the oracle (C restatement, OpenMP)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = cfg["target"]
    if t[0] == "uniform":
        tg = O.make_target(O.TARGET_UNIFORM, t[1], t[2])
    else:
        sk = O.SDF_SPHERE if t[3] == "sphere" else O.SDF_TORUS
        tg = O.make_target(O.TARGET_SDF_BANDS, t[1], t[2], sk, t[4], t[5])
    g = O.Grid(tg)
    g.set_leaf_types_preset()
    if cfg["assembly"] == "plain":
        g.coarsen_types_and_coeffs(11, 0.5)
    elif cfg["assembly"] == "solid":
        g.cutcell_coeffs(O.SDF_SPHERE, t[4], -1, 11, 0.5)
    else:
        g.set_wall_types(t[1], True)
        g.cutcell_coeffs(O.SDF_TORUS, t[4], 1, 11, 0.5)
    kind, seed = cfg["rhs"]
    if kind == "uniform":
        g.rhs_uniform(1, seed, True)
    else:
        g.rhs_divergence(1, 5, seed)
    return O, g


def cpu_baseline(cfg, max_iters):
    """The same MGPCG solve (PCG + FAS V(2,2) damped Jacobi, FP64 direct level 0,
    FP32 vectors, FP64 dots) run by the oracle's OpenMP C restatement on the host
    cores, from x = 0 to 1e-6, timed around the solve only (assembly excluded)."""
    O, g = oracle_grid(cfg)
    threads = O.lib().og_num_threads()
    cyc = O.Cycle(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_iters=0, coarsest_direct=1,
                  R_restrict=1.0, P=1.0, tmp_ch=6)
    t0 = time.perf_counter()
    it, rel = g.pcg(cyc, 11, max_iters, 1e-6)
    dt = time.perf_counter() - t0
    return {"value": dt * 1e3, "unit": "ms per solve", "cores": threads, "kind": "port",
            "threads_runtime": "OpenMP (libgomp)",
            "sample": "one full solve of the same workload from x = 0 to 1e-6 by the oracle's C restatement "
                      "(%d PCG iterations, rel res %.3e), solve only" % (it, rel),
            "iterations": it, "rel_res": rel}


def load_traffic(name):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by the factor a plain copy kernel of
    known size gives); None when no summary is present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_dispatch:"):
                return float(line.split(":", 1)[1])
    return None


KERNELS = {0: ("apply (y = A p on LEAF tiles + p.y partials)", 25.0),
           1: ("damped-Jacobi sweep, finest level", 29.0),
           2: ("residual + restrict, finest level", 29.5),
           3: ("x += a p, r -= a Ap + r.r partials", 28.0),
           4: ("FAS prolong + correct, finest level", 8.5),
           5: ("p = z + b p", 12.0),
           6: ("r.z partials", 8.0),
           7: ("first Jacobi sweep from x = 0, finest level", 13.0),
           8: ("plain copy", 8.0),
           9: ("learned smoother inference, finest level (BF16 MFMA)", 8.0),
           10: ("flat copy 1 GiB -> 1 GiB", 8.0),
           11: ("damped-Jacobi sweep, finest level, FP16 coefficient copies", 21.0),
           # coded coefficients, LEAF diagonals derived: x, b, x_out and one word (type + 3 codes)
           12: ("damped-Jacobi sweep, finest level, coded coefficients", 16.0),
           13: ("apply (y = A p on LEAF tiles + p.y partials), coded coefficients", 12.0),
           14: ("residual + restrict, finest level, coded coefficients", 13.0)}
# learned smoother: 2 * 27 taps * (1*16 + 16*16 + 16*1) channel pairs * 1000 volume
# positions per tile ("same" padding on the halo'd 10^3 volume)
NET_FLOP_PER_TILE = 2.0 * 27 * (16 + 256 + 16) * 1000
MFMA_BF16_PEAK_TFLOPS = 2500.0
WEIGHTS = os.path.join(ROOT, "cirrus-amg_amd", "data", "smoother_c4.f32")


def stream_copy_ceiling(solver, reps=10):
    """On-box HBM copy ceiling: the library's flat float4 copy of 1 GiB into a
    second 1 GiB buffer (four times the 256 MB last-level cache), timed with
    hipEvents on the solver's stream."""
    ms, _ = solver.bench_kernel(10, reps)
    nbytes = 2 * (1 << 30)
    return {"GBps": nbytes / (ms * 1e-3) / 1e9, "bytes_moved": nbytes, "ms": ms,
            "kernel": "copy_flat, 1 GiB -> 1 GiB FP32"}


def profile_kernel(cfg, kernel, reps):
    import numpy as np
    import cirrus_amg as P
    g = build(P, cfg)
    solver = P.AMGSolver(g, 11, 0.5, 1.0, 1.0)
    if kernel == 9:
        solver.set_net(np.fromfile(WEIGHTS, "<f4"))
    ms, tiles = solver.bench_kernel(kernel, reps)
    name, bpc = KERNELS[kernel]
    out = {"kernel": kernel, "name": name, "reps": reps, "ms": ms, "tiles": tiles,
           "bytes_per_launch": bpc * 512 * tiles, "GBps": bpc * 512 * tiles / (ms * 1e-3) / 1e9}
    if kernel == 9:
        out["TFLOPs"] = NET_FLOP_PER_TILE * tiles / (ms * 1e-3) / 1e12
        out["mfma_util"] = out["TFLOPs"] / MFMA_BF16_PEAK_TFLOPS
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-refinement leg")
    ap.add_argument("--profile-kernel", type=int, default=None, choices=sorted(KERNELS),
                    help="only time one kernel (for rocprofv3 counter runs) and exit")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--no-legs", action="store_true",
                    help="time the headline solve and its dominant kernel only (no secondary legs)")
    ap.add_argument("--learned", type=int, default=1, help="also time learned-FCG when trained weights exist")
    ap.add_argument("--cycle-coef", choices=("fp16", "fp32"), default="fp32",
                    help="coefficients the preconditioning cycle reads: the FP32 channels (read as exact per-level "
                         "codes where a level has at most 255 distinct face values: the same values), or FP16 "
                         "copies of c / h (the PCG operator is FP32 either way); the other form is timed as a "
                         "secondary leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs the single-GPU solve only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import cirrus_amg as P

    cfg = CONFIGS[args.config]
    if args.profile_kernel is not None:
        profile_kernel(cfg, args.profile_kernel, args.reps)
        return
    t_build = time.time()
    g = build(P, cfg)
    t_build = time.time() - t_build
    lv, co, ty = g.tiles()
    types = g.cell_types()
    dof = int((types[ty == 1] == 2).sum())
    g.copy_channel(1, 5)   # keep b resident

    q16 = 1 if args.cycle_coef == "fp16" else 0
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                          max_iters=200, relative_tolerance=1e-6, level_iters=2, sync_stride=1, cycle_coef_fp16=q16,
                          no_true_residual=1)
    solver = P.AMGSolver(g, 11, 0.5, 1.0, 1.0)

    def step(o=opts):
        g.copy_channel(5, 1)
        it, rel = solver.solve(max_iters=200, relative_tolerance=1e-6, level_iters=2, coarsest_iters=0, opts=o)
        return it, rel

    def checked(o):
        """the same options with the FP64 true-residual recompute on (a diagnostic of this
        library that the reference's solve does not compute; run untimed)"""
        c = P.SolveOpts.from_buffer_copy(o)
        c.no_true_residual = 0
        return c

    for _ in range(args.warmup):
        step()
    P.device_synchronize()
    t0 = time.perf_counter()
    iters, rels, dev_ms = [], [], []
    for _ in range(args.steps):
        it, rel = step()
        iters.append(it)
        rels.append(rel)
        dev_ms.append(solver.last_stats.ms)
    P.device_synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    step(checked(opts))   # untimed: the FP64 true residual of the same solve
    true_rel = solver.last_stats.true_rel_res
    x32 = g.channel(0).astype(np.float64) if not args.no_fp64 else None
    m_in = (types == 2) & (ty == 1)[:, None]   # LEAF INTERIOR cells

    # levels whose coefficients the kernels read as exact codes (forms bit 6 clear)
    coded, wide = solver.code_levels()
    fine_coded = bool(coded >> g.max_level & 1)
    # dominant kernel: damped-Jacobi sweep on the finest level (3 per PCG iteration plus the
    # first sweep from zero); FP32 channels: x_in, b, 3 face coefficients, diagonal (4 B each),
    # type (1 B), x_out = 29 B per cell; FP16 copies: coefficients and diagonal 2 B = 21 B; coded
    # (LEAF diagonals derived): x_in, b, x_out and one 4-B word (type and three codes) = 16 B
    kid = 11 if q16 else (12 if fine_coded else 1)
    k_ms, k_tiles = solver.bench_kernel(kid, 50)
    alg_bytes = KERNELS[kid][1] * 512 * k_tiles
    k_gbs = alg_bytes / (k_ms * 1e-3) / 1e9
    traffic = load_traffic({11: "jacobi_fp16coef_finest_%s_pmc.txt", 12: "jacobi_coded_finest_%s_pmc.txt",
                            1: "jacobi_finest_%s_pmc.txt"}[kid] % args.config)
    forms = {}
    for fk in (1, 11, 12):
        if fk == 12 and not fine_coded:
            continue
        f_ms, f_tiles = (k_ms, k_tiles) if fk == kid else solver.bench_kernel(fk, 50)
        f_bytes = KERNELS[fk][1] * 512 * f_tiles
        forms[KERNELS[fk][0]] = {"ms": f_ms, "GBps": f_bytes / (f_ms * 1e-3) / 1e9, "bytes_per_launch": f_bytes}
    apply_ms, apply_tiles = solver.bench_kernel(0, 50)
    apply_kernel = {"fp32_channels": {"GBps": 25.0 * 512 * apply_tiles / (apply_ms * 1e-3) / 1e9, "ms": apply_ms,
                                      "tiles": apply_tiles}}
    if fine_coded:
        a_ms, a_tiles = solver.bench_kernel(13, 50)
        apply_kernel["coded"] = {"GBps": KERNELS[13][1] * 512 * a_tiles / (a_ms * 1e-3) / 1e9, "ms": a_ms,
                                 "tiles": a_tiles}

    # the headline form's own bytes, and the other cycle form's
    b_iter = canonical_bytes_per_iter(lv, ty, g.max_level, 8 if q16 else 16, 0 if q16 else coded, 0 if q16 else wide)
    b_iter_other = canonical_bytes_per_iter(lv, ty, g.max_level, 16 if q16 else 8, coded if q16 else 0,
                                            wide if q16 else 0)
    solve_ms = float(np.mean(dev_ms))
    it_mean = float(np.mean(iters))
    solve_gbs = b_iter * it_mean / (solve_ms * 1e-3) / 1e9

    result = {
        "metric": "ms per Poisson solve to 1e-6 rel-residual (plain MGPCG)",
        "value": ms_per_step,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": ("f32 (f16 cycle coefficients, f64 dots, f64 level-0 direct solve)" if q16
                  else "f32 (f64 dots, f64 level-0 direct solve; FP32 coefficients read as exact per-level codes)"),
        "data": "synthetic (seeded)",
        "config": {"workload": cfg["name"], "tiles": int(g.N), "leaf_tiles": int((ty == 1).sum()), "dof": dof,
                   "levels": int(g.max_level) + 1,
                   "solver": "PCG (FP32 operator) + FAS V(2,2) damped Jacobi w=2/3, direct level 0"
                             + (", cycle on FP16 coefficient copies c/h" if q16 else
                                ", FP32 cycle; coefficients of levels %s read as exact 8-bit codes, of levels %s "
                                "as exact 10-bit codes"
                                % ([l for l in range(g.max_level + 1) if (coded & ~wide) >> l & 1],
                                   [l for l in range(g.max_level + 1) if wide >> l & 1])),
                   "parallelism": "single GPU"},
        "iterations": iters,
        "rel_res": rels,
        "true_rel_res_fp64": true_rel,
        "device_ms_per_solve": solve_ms,
        "dof_per_s": dof / (ms_per_step * 1e-3),
        "solve_GBps_canonical": solve_gbs,
        "solve_frac_of_hbm": solve_gbs / HBM_PEAK_GBS,
        "canonical_bytes_per_iter": b_iter,
        "canonical_bytes_note": ("cycle coefficients at 2 B (FP16 copies): sum_l n_l (13+21+21.5+8.5+2*21) + 81 N_leaf"
                                 if q16 else "coefficients at 4 B: sum_l n_l (21+29+29.5+8.5+2*29) + 81 N_leaf, less "
                                 "13 B per LEAF cell and 9 B per NONLEAF cell for each halo pass and the operator on "
                                 "coded levels (one 4-B word of type and codes; the operator reads FP32 channels on "
                                 "10-bit-coded levels)"),
        "coded_levels": [l for l in range(g.max_level + 1) if coded >> l & 1],
        "ten_bit_coded_levels": [l for l in range(g.max_level + 1) if wide >> l & 1],
        "roofline": {"bound": "hbm", "achieved": k_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k_gbs / HBM_PEAK_GBS,
                     "traffic": traffic, "kernel": "jacobi sweep, finest level", "kernel_ms": k_ms,
                     "bytes_per_launch": alg_bytes, "tiles_per_launch": k_tiles,
                     # the same values in the FP32 channels' bytes (29 B a cell): what the sweep delivers
                     "fp32_channel_bytes_GBps": KERNELS[1][1] * 512 * k_tiles / (k_ms * 1e-3) / 1e9,
                     "traffic_GBps": traffic / (k_ms * 1e-3) / 1e9 if traffic else None},
        "jacobi_forms": forms,
        "apply_kernel": apply_kernel,
        "build_s": t_build,
    }
    ceil = stream_copy_ceiling(solver)
    result["hbm_copy_ceiling"] = ceil
    result["roofline"]["frac_of_copy_ceiling"] = k_gbs / ceil["GBps"]
    result["solve_frac_of_copy_ceiling"] = solve_gbs / ceil["GBps"]
    if args.no_legs:
        print(json.dumps(result))
        return
    if not args.no_fp64:
        # mixed-precision form: FP32 MGPCG inner solves inside FP64 iterative
        # refinement to 1e-10; its solution is the FP64 solution of the same
        # discrete operator against which the FP32 1e-6 solve's error is stated
        g.copy_channel(5, 1)
        o64 = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                             precision=P.PREC_FP64_REFINE, outer_tolerance=1e-10, inner_tolerance=1e-4,
                             cycle_coef_fp16=q16)
        t64 = time.perf_counter()
        it64, rel64 = solver.solve(max_iters=200, relative_tolerance=1e-6, level_iters=2, coarsest_iters=0, opts=o64)
        P.device_synchronize()
        t64 = (time.perf_counter() - t64) * 1e3
        st = solver.last_stats
        x64 = solver.x64()
        m = (types == 2) & (ty == 1)[:, None]
        err = float(np.linalg.norm((x32 - x64)[m]) / np.linalg.norm(x64[m]))
        result["fp64_refinement"] = {"ms": t64, "outer_iters": int(st.outer_iters), "inner_iters_total": int(it64),
                                     "true_rel_res_fp64": float(st.true_rel_res), "tolerance": 1e-10}
        result["rel_l2_err_fp32_solve_vs_fp64"] = err
    # the same solve with the other cycle-coefficient form (the PCG operator is FP32 in both,
    # so the tolerance and the solution are the same)
    o16 = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                         max_iters=200, relative_tolerance=1e-6, level_iters=2, sync_stride=1, cycle_coef_fp16=1 - q16,
                         no_true_residual=1)
    ms16, its16, dev16 = [], [], []
    for rep in range(args.warmup + args.steps):
        g.copy_channel(5, 1)
        P.device_synchronize()
        t0 = time.perf_counter()
        it, rel = solver.solve(max_iters=200, relative_tolerance=1e-6, level_iters=2, coarsest_iters=0, opts=o16)
        P.device_synchronize()
        if rep >= args.warmup:
            ms16.append((time.perf_counter() - t0) * 1e3)
            its16.append(it)
            dev16.append(solver.last_stats.ms)
    step(checked(o16))   # untimed: the FP64 true residual of the same solve
    leg16 = {"ms": float(np.mean(ms16)), "device_ms": float(np.mean(dev16)), "iterations": its16, "rel_res": rel,
             "true_rel_res_fp64": float(solver.last_stats.true_rel_res),
             "ms_ratio_to_headline": float(np.mean(ms16)) / ms_per_step,
             "canonical_bytes_per_iter": b_iter_other,
             "solve_frac_of_hbm": b_iter_other * float(np.mean(its16)) / (float(np.mean(dev16)) * 1e-3) / 1e9
             / HBM_PEAK_GBS}
    if x32 is not None and "fp64_refinement" in result:
        x16 = g.channel(0).astype(np.float64)
        leg16["rel_l2_err_vs_fp64"] = float(np.linalg.norm((x16 - x64)[m]) / np.linalg.norm(x64[m]))
    result["fp32_cycle_coefficients" if q16 else "fp16_cycle_coefficients"] = leg16
    if x32 is not None and "fp64_refinement" in result:
        # the fastest stopping tolerance of the headline form whose solution is within 1e-5
        # (relative L2 over LEAF INTERIOR cells) of the FP64-refined solution
        acc = []
        for tol in (7e-7, 5e-7, 3e-7, 2e-7, 1e-7):
            oa = P.SolveOpts.from_buffer_copy(opts)
            oa.relative_tolerance = tol
            msa = []
            for rep in range(1 + args.steps):
                g.copy_channel(5, 1)
                P.device_synchronize()
                t0 = time.perf_counter()
                it, rel = solver.solve(max_iters=200, relative_tolerance=tol, level_iters=2, coarsest_iters=0, opts=oa)
                P.device_synchronize()
                if rep:
                    msa.append((time.perf_counter() - t0) * 1e3)
            xa = g.channel(0).astype(np.float64)
            ea = float(np.linalg.norm((xa - x64)[m]) / np.linalg.norm(x64[m]))
            acc.append({"relative_tolerance": tol, "ms": float(np.mean(msa)), "iterations": it, "rel_res": rel,
                        "rel_l2_err_vs_fp64": ea})
            if ea <= 1e-5:
                break
        result["accuracy_target_leg"] = {"target_rel_l2_err_vs_fp64": 1e-5, "tried": acc,
                                         "meets_target": acc[-1]["rel_l2_err_vs_fp64"] <= 1e-5,
                                         "ms": acc[-1]["ms"], "relative_tolerance": acc[-1]["relative_tolerance"]}
    # the reference's drop-in settings (FluidEuler::project, cirrus_cutcell/FluidEuler.cu:702-706;
    # src/AMGSolver.h:25-30): FAS W-cycle, red-black SOR w=1.5, 2+2 sweeps, 10 coarsest
    # iterations on level 0, at most 100 PCG iterations, 1e-6
    od = P.default_opts(no_true_residual=1)
    msd, itd = [], []
    for rep in range(args.warmup + args.steps):
        g.copy_channel(5, 1)
        P.device_synchronize()
        t0 = time.perf_counter()
        it, rel = solver.solve(max_iters=100, relative_tolerance=1e-6, level_iters=2, coarsest_iters=10, opts=od)
        P.device_synchronize()
        if rep >= args.warmup:
            msd.append((time.perf_counter() - t0) * 1e3)
            itd.append(it)
    result["fluideuler_default_w_cycle"] = {
        "ms": float(np.mean(msd)), "device_ms": float(solver.last_stats.ms), "iterations": itd, "rel_res": rel,
        "solver": "PCG + FAS W-cycle (mu=2), red-black SOR w=1.5, 2+2 sweeps, 10 RB-GS iterations on level 0, "
                  "FP32 cycle coefficients"}
    weights = WEIGHTS
    if args.learned and os.path.exists(weights):
        # learned-FCG: flexible CG whose V-cycle takes the learned tile smoother
        # (trained on the moving-sphere scene) as the finest pre-smoothing step
        solver.set_net(np.fromfile(weights, "<f4"))
        ol = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                            max_iters=200, relative_tolerance=1e-6, level_iters=2, sync_stride=1, flexible=1, learned=1,
                            cycle_coef_fp16=q16)
        of = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                            max_iters=200, relative_tolerance=1e-6, level_iters=2, sync_stride=1, flexible=1,
                            cycle_coef_fp16=q16)
        legs = {}
        for name, o in (("learned_fcg", ol), ("plain_fcg", of)):
            ms, its = [], []
            for rep in range(1 + args.steps):
                g.copy_channel(5, 1)
                P.device_synchronize()
                t0 = time.perf_counter()
                it, rel = solver.solve(max_iters=200, relative_tolerance=1e-6, level_iters=2, coarsest_iters=0, opts=o)
                P.device_synchronize()
                if rep:
                    ms.append((time.perf_counter() - t0) * 1e3)
                    its.append(it)
            leg = {"ms": float(np.mean(ms)), "iterations": its, "rel_res": rel,
                   "true_rel_res_fp64": float(solver.last_stats.true_rel_res)}
            if o.learned:
                judged, kept = solver.learned_counts()
                leg["tiles_kept_net_correction"] = kept / max(judged, 1)
            if x32 is not None and "fp64_refinement" in result:
                xl = g.channel(0).astype(np.float64)
                leg["rel_l2_err_vs_fp64"] = float(np.linalg.norm((xl - x64)[m]) / np.linalg.norm(x64[m]))
            legs[name] = leg
        legs["speedup_learned_vs_plain_mgpcg"] = ms_per_step / legs["learned_fcg"]["ms"]
        # the inference kernel alone, back to back (overwrites the work channels)
        ims, itiles = solver.bench_kernel(9, 20)
        tf = NET_FLOP_PER_TILE * itiles / (ims * 1e-3) / 1e12
        legs["inference_kernel"] = {"ms": ims, "tiles": itiles, "TFLOPs": tf, "peak_TFLOPs": MFMA_BF16_PEAK_TFLOPS,
                                    "mfma_util": tf / MFMA_BF16_PEAK_TFLOPS,
                                    "flop_per_tile": NET_FLOP_PER_TILE}
        result["learned"] = legs
    if not q16 and coded:
        # the headline solve with the codes switched off (forms bit 6): the same values from the
        # FP32 channels; the solution differs only through the grouping of FP64 partial sums
        # where the two forms launch different thread counts
        mask0, _ = P.get_kernel_forms()
        P.set_kernel_forms(mask0 | 64, 0)
        try:
            msn, itn, devn = [], [], []
            for rep in range(args.warmup + args.steps):
                it, rel = step()
                if rep >= args.warmup:
                    itn.append(it)
                    devn.append(solver.last_stats.ms)
            P.device_synchronize()
            xn = g.channel(0).astype(np.float64)
            t0 = time.perf_counter()
            for rep in range(args.steps):
                step()
            P.device_synchronize()
            msn = (time.perf_counter() - t0) * 1e3 / args.steps
        finally:
            P.set_kernel_forms(mask0, 0)
        leg = {"ms": msn, "device_ms": float(np.mean(devn)), "iterations": itn, "rel_res": rel,
               "ms_ratio_to_headline": msn / ms_per_step,
               "canonical_bytes_per_iter": canonical_bytes_per_iter(lv, ty, g.max_level, 16, 0)}
        if x32 is not None:
            leg["rel_l2_diff_vs_headline"] = float(np.linalg.norm((xn - x32)[m_in]) / np.linalg.norm(x32[m_in]))
        result["fp32_channels_without_codes"] = leg
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(cfg, 200)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
