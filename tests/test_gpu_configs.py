"""The benchmark configurations at their stated sizes, on the GPU, against the oracle.

* C2: 3-level graded octree, finest 256^3, Neumann sphere, Dirichlet box,
  b = div of N(0,1) face velocities (seed 1);
* C3: 5-level graded cut-cell octree, finest 1024^3, torus, 4x4 supersampled faces,
  Dirichlet y = 1 (seed 2);
* C4 (reduced): the learned smoother's data generation and training on a few
  frames of the moving-sphere scene, the loss falling;
* C5 (reduced): three warm-started frames of the translating torus.

For C2 and C3 the library's host assembly equals the oracle's bit for bit, the
GPU solve (PCG + FAS V(2,2) damped Jacobi, FP64 direct level 0, cycle on FP32 or
FP16 coefficient copies) reaches 1e-6 in the oracle's MGPCG iteration count +-1,
and its error is measured against an FP64 solution computed on the CPU: FP64
iterative refinement whose residual is the oracle's FP64 CSR operator and whose
inner solves are the oracle's MGPCG (true relative residual <= 1e-11).
"""
import os
import sys
import time

import numpy as np
import pytest

import cirrus_amg as P
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workload definitions)

CYCLE = dict(mu=1, smoother=1, omega=2.0 / 3.0, level_iters=2, coarsest_iters=0, coarsest_direct=1, R_restrict=1.0,
             P=1.0, tmp_ch=6)


def _cpu_fp64_solution(go):
    """FP64 solution of the oracle's CSR operator by iterative refinement around the oracle's MGPCG;
    returns (flat cell indices of the unknowns, x64, oracle MGPCG iterations from x = 0)"""
    A = go.assemble_csr(H.C0)
    b_full = go.channel(1).copy()
    b = b_full.ravel()[A.cell].astype(np.float64)
    cyc = O.Cycle(**CYCLE)
    x = np.zeros(A.n)
    first_iters = None
    for _ in range(6):
        r = b - A.matvec(x)
        if np.linalg.norm(r) <= 1e-11 * np.linalg.norm(b):
            break
        sc = float(np.abs(r).max())
        rhs = np.zeros(b_full.size, np.float32)
        rhs[A.cell] = (r / sc).astype(np.float32)
        go.channel(1)[:] = rhs.reshape(b_full.shape)
        it, _ = go.pcg(cyc, H.C0, 200, 1e-6)
        if first_iters is None:
            first_iters = it
        x += go.channel(0).ravel()[A.cell].astype(np.float64) * sc
    go.channel(1)[:] = b_full
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-11 * np.linalg.norm(b)
    return A.cell, x, first_iters


def _config(name):
    cfg = bench.CONFIGS[name]
    t0 = time.time()
    gp = bench.build(P, cfg)
    _, go = bench.oracle_grid(cfg)
    print("\n%s: %d tiles, built in %.1f s" % (name, gp.N, time.time() - t0), flush=True)
    return gp, go


def _solve(gp, fp16, tol=1e-6):
    o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                       cycle_coef_fp16=fp16)
    gp.copy_channel(5, 1)
    s = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    it, rel = s.solve(opts=o, max_iters=200, relative_tolerance=tol, coarsest_iters=0)
    assert rel < tol
    return it, s.last_stats.true_rel_res, gp.channel(0).copy()


def _check_config(name, err_1e6, tol_acc):
    gp, go = _config(name)
    # topology and host assembly: bit for bit
    for a, b in zip(go.tiles(), gp.tiles()):
        assert np.array_equal(a, b)
    assert np.array_equal(go.slots(), gp.slots())
    assert np.array_equal(go.cell_types(), gp.cell_types())
    for a in range(4):
        assert np.array_equal(H.bits(go.channel(H.C0 + a)), H.bits(gp.channel(H.C0 + a))), a
    assert np.array_equal(H.bits(go.channel(1)), H.bits(gp.channel(1)))
    gp.copy_channel(1, 5)
    t0 = time.time()
    cells, x64, oracle_iters = _cpu_fp64_solution(go)
    print("%s: CPU FP64 reference in %.1f s, oracle MGPCG %d iterations" % (name, time.time() - t0, oracle_iters),
          flush=True)
    errs = {}
    for fp16 in (0, 1):
        it, true_rel, x = _solve(gp, fp16)
        err = np.linalg.norm(x.ravel()[cells] - x64) / np.linalg.norm(x64)
        errs[fp16] = err
        print("%s: fp16 cycle %d: %d iterations, true rel res %.3e, rel L2 err vs FP64 %.3e" % (name, fp16, it, true_rel,
                                                                                                 err), flush=True)
        assert abs(it - oracle_iters) <= 1
        assert true_rel < 1.1e-6
        assert err <= err_1e6
    # the accuracy target: a tighter stop gives <= 1e-5 against FP64
    it, _, x = _solve(gp, 1, tol_acc)
    err = np.linalg.norm(x.ravel()[cells] - x64) / np.linalg.norm(x64)
    print("%s: fp16 cycle at %.0e: %d iterations, rel L2 err %.3e" % (name, tol_acc, it, err), flush=True)
    assert err <= 1e-5
    # the GPU's own FP64 refinement reaches the CPU's FP64 solution
    o64 = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                         precision=P.PREC_FP64_REFINE, outer_tolerance=1e-10, inner_tolerance=1e-4, cycle_coef_fp16=1)
    gp.copy_channel(5, 1)
    s = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    s.solve(opts=o64, max_iters=200, relative_tolerance=1e-6, coarsest_iters=0)
    assert s.last_stats.true_rel_res <= 1e-10
    xg = s.x64().ravel()[cells]
    assert np.linalg.norm(xg - x64) <= 1e-9 * np.linalg.norm(x64)
    return errs


def test_c2_full_size():
    _check_config("c2", 1e-5, 5e-7)


def test_c3_full_size():
    _check_config("c3", 2e-5, 3e-7)


def test_c4_reduced_training_loss_falls():
    """three frames of the moving-sphere data generation, then Adam on the collected tiles"""
    sys.path.insert(0, os.path.join(ROOT, "cirrus-amg_amd"))
    import train_smoother as T
    tr = P.Trainer(200_000, seed=3)
    added = [T.collect_frame(tr, f, 3) for f in (0, 10, 20)]
    print("\nC4 reduced: tiles per frame", added, "total", tr.size, flush=True)
    assert all(a > 0 for a in added) and tr.size == sum(added)
    n = min(tr.size, 4096)
    first, _ = tr.loss_grad(0, n, bf16=True)
    losses = [tr.step(1024, lr=1e-3, bf16=True) for _ in range(60)]
    last, _ = tr.loss_grad(0, n, bf16=True)
    print("C4 reduced: loss on the first %d tiles %.4e -> %.4e (step losses %.3e .. %.3e)" % (n, first, last, losses[0],
                                                                                             np.mean(losses[-10:])),
          flush=True)
    assert np.isfinite(losses).all()
    assert last < 0.8 * first
    assert np.mean(losses[-10:]) < np.mean(losses[:10])


def test_c5_reduced_warm_started_sequence():
    """three frames of the translating torus at full size: warm starts carried across octrees, plain
    MGPCG and learned FCG reach 1e-6 (true FP64 residual) within 1e-4 of each frame's FP64 solution"""
    sys.path.insert(0, os.path.join(ROOT, "cirrus-amg_amd"))
    import c5_sequence as S
    w = np.fromfile(os.path.join(ROOT, "cirrus-amg_amd", "data", "smoother_c4.f32"), "<f4")
    prev = None
    rows = []
    for k in range(3):
        g = S.frame_grid(k)
        lv, co, ty = g.tiles()
        mask = (g.cell_types().reshape(g.N, 512) == S.INTERIOR) & (ty & S.LEAF).astype(bool)[:, None]
        solver = P.AMGSolver(g, S.C0, S.R_MATRIX, 1.0, 1.0)
        solver.set_net(w)
        o64 = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                             precision=P.PREC_FP64_REFINE, outer_tolerance=1e-10, inner_tolerance=1e-4)
        g.copy_channel(S.CH_B, 1)
        solver.solve(max_iters=400, relative_tolerance=1e-6, coarsest_iters=0, opts=o64)
        assert solver.last_stats.true_rel_res <= 1e-10
        ref = solver.x64()
        x0p = x0l = None
        if prev is not None:
            og, xp, xl = prev
            x0p = S.transfer_leaf_values(og, xp, g)
            x0l = S.transfer_leaf_values(og, xl, g)
            og.close()
        cold, _ = S.run_solve(solver, g, None, S._opts(0), ref, mask)
        plain, xp = S.run_solve(solver, g, x0p, S._opts(0), ref, mask)
        learned, xl = S.run_solve(solver, g, x0l, S._opts(1), ref, mask)
        row = {"frame": k, "cold_iters": cold["iters"], "plain_iters": plain["iters"], "learned_iters": learned["iters"],
               "plain_err": plain["rel_l2_err_vs_fp64"], "learned_err": learned["rel_l2_err_vs_fp64"]}
        print("\nC5 reduced:", row, flush=True)
        rows.append(row)
        for r in (cold, plain, learned):
            assert r["rel_res"] < 1e-6 and r["true_rel_res_fp64"] < 1.1e-6
            assert r["rel_l2_err_vs_fp64"] < 1e-4
        # (every frame's right-hand side has its own seed, so the carried solution is no closer
        # to the new one than x = 0 is: the warm start has to converge, not to save iterations)
        solver.close()
        prev = (g, xp, xl)
    prev[0].close()
