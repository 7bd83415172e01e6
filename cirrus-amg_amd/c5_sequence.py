"""Learned flexible CG vs plain MGPCG over a sequence of warm-started pressure solves.

Scene: the 5-level graded cut-cell octree of the torus workload (R = 0.15, r = 0.05,
axis z, finest 1024^3 within |d| < 0.01, band doubling per coarser level, 4x4
supersampled face fractions, Dirichlet at y = 1, Neumann elsewhere), with the
torus translating along x at 0.5 units/s (dt = 0.01) from (0.5, 0.5, 0.5).  Frame
k rebuilds the octree, assembles the cut-cell operator and takes its right-hand
side from the divergence of i.i.d. N(0,1) face velocities with seed 4 + k.  Each
method warm-starts from its own previous solution, carried to the new octree by
transfer_leaf_values, and stops at 1e-6 relative residual.  The FP64-refined
solve of every frame (true relative residual <= 1e-10) is the reference for the
error of both.

    python cirrus-amg_amd/c5_sequence.py --frames 50 --out gpurun_out/c5.json

prints one JSON line per frame and a summary line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import cirrus_amg as P  # noqa: E402

C0, R_MATRIX = 11, 0.5
CH_B = 8
LEAF, NONLEAF, INTERIOR = 1, 2, 2
BANDS = {7: 0.01, 6: 0.02, 5: 0.04, 4: 0.08}


def frame_grid(k, dt=0.01, speed=0.5):
    c = (0.5 + speed * dt * k, 0.5, 0.5, 0.15, 0.05)
    target = P.make_target(P.TARGET_SDF_BANDS, 3, 7, P.SDF_TORUS, c, BANDS)
    g = P.HADeviceGrid.test_grid(target)
    g.set_wall_types(3, True)
    g.assemble_cutcell(P.SDF_TORUS, c, 1, C0, R_MATRIX)
    g.rhs_divergence(1, 5, 4 + k)
    g.copy_channel(1, CH_B)
    return g


def _block_mean(v):
    """(8,8,8) child values -> (4,4,4) means of aligned 2^3 blocks"""
    return v.reshape(4, 2, 4, 2, 4, 2).mean(axis=(1, 3, 5))


def transfer_leaf_values(old_grid, old_x, new_grid):
    """Values of a cell field on the LEAF tiles of new_grid from its values on the
    LEAF tiles of old_grid: a tile present in both (LEAF, or NONLEAF holding the
    means of its children) is copied; a tile that old_grid has coarser takes the
    value of the enclosing old cell.  Cells that are not INTERIOR in new_grid get 0."""
    lv, co, ty = old_grid.tiles()
    vals = np.zeros((old_grid.N, 8, 8, 8), np.float64)
    vals[:] = old_x.reshape(-1, 8, 8, 8)
    key = {}
    for t in range(old_grid.N):
        if ty[t] & (LEAF | NONLEAF):
            key[(int(lv[t]), int(co[t, 0]), int(co[t, 1]), int(co[t, 2]))] = t
    # NONLEAF tiles: means of their children, finest levels first
    for t in np.argsort(-lv, kind="stable"):
        if not ty[t] & NONLEAF:
            continue
        l, T = int(lv[t]), co[t]
        for q in range(8):
            o = np.array([q & 1, (q >> 1) & 1, (q >> 2) & 1])
            ch = key.get((l + 1, *(int(v) for v in 2 * T + o)))
            if ch is not None:
                vals[t, 4 * o[0]:4 * o[0] + 4, 4 * o[1]:4 * o[1] + 4, 4 * o[2]:4 * o[2] + 4] = _block_mean(vals[ch])
    nlv, nco, nty = new_grid.tiles()
    out = np.zeros((new_grid.N, 8, 8, 8), np.float64)
    cell = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1)   # [i][j][k][axis]
    for t in np.where(nty & LEAF)[0]:
        l, T = int(nlv[t]), nco[t]
        for s in range(l + 1):
            A = T >> s
            src = key.get((l - s, int(A[0]), int(A[1]), int(A[2])))
            if src is None:
                continue
            if s == 0:
                out[t] = vals[src]
            else:
                g = ((8 * T + cell) >> s) - 8 * A
                out[t] = vals[src][g[..., 0], g[..., 1], g[..., 2]]
            break
    types = new_grid.cell_types().reshape(new_grid.N, 512)
    out = out.reshape(new_grid.N, 512)
    out[types != INTERIOR] = 0
    return out.astype(np.float32)


def _opts(learned):
    if learned:
        return P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                              flexible=1, learned=learned)
    return P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1)


def run_solve(solver, g, x0, opts, ref, mask):
    g.copy_channel(CH_B, 1)
    if x0 is not None:
        g.set_channel(0, x0)
        opts.warm_start = 1
    else:
        opts.warm_start = 0
    P.device_synchronize()
    t0 = time.perf_counter()
    it, rel = solver.solve(max_iters=200, relative_tolerance=1e-6, level_iters=2, coarsest_iters=0, opts=opts)
    P.device_synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    x = g.channel(0)
    err = float(np.linalg.norm((x.astype(np.float64) - ref)[mask]) / max(np.linalg.norm(ref[mask]), 1e-300))
    out = {"ms": wall, "device_ms": float(solver.last_stats.ms), "iters": int(it), "rel_res": float(rel),
           "true_rel_res_fp64": float(solver.last_stats.true_rel_res), "rel_l2_err_vs_fp64": err}
    if opts.learned:
        judged, kept = solver.learned_counts()
        out["tiles_kept_net_correction"] = kept / max(judged, 1)
    return out, x


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=50)
    ap.add_argument("--learned", type=int, default=1, help="learned mode (1: replaces the finest pre-smoothing, "
                    "2: runs before it)")
    ap.add_argument("--weights", default=os.path.join(HERE, "data", "smoother_c4.f32"))
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    w = np.fromfile(args.weights, "<f4")
    prev = None
    rows = []
    t_start = time.time()
    for k in range(args.frames):
        g = frame_grid(k)
        lv, co, ty = g.tiles()
        mask = (g.cell_types().reshape(g.N, 512) == INTERIOR) & (ty & LEAF).astype(bool)[:, None]
        solver = P.AMGSolver(g, C0, R_MATRIX, 1.0, 1.0)
        solver.set_net(w)
        o64 = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                             precision=P.PREC_FP64_REFINE, outer_tolerance=1e-10, inner_tolerance=1e-4)
        g.copy_channel(CH_B, 1)
        solver.solve(max_iters=400, relative_tolerance=1e-6, coarsest_iters=0, opts=o64)
        ref = solver.x64()
        x0p = x0l = None
        if prev is not None:
            og, xp, xl = prev
            x0p = transfer_leaf_values(og, xp, g)
            x0l = transfer_leaf_values(og, xl, g)
            og.close()
        plain, xp = run_solve(solver, g, x0p, _opts(0), ref, mask)
        learned, xl = run_solve(solver, g, x0l, _opts(args.learned), ref, mask)
        row = {"frame": k, "leaf_tiles": int((ty & LEAF).astype(bool).sum()), "plain_mgpcg": plain,
               "learned_fcg": learned, "s": round(time.time() - t_start, 1)}
        rows.append(row)
        print(json.dumps(row), flush=True)
        solver.close()
        prev = (g, xp, xl)
    if prev is not None:
        prev[0].close()

    def stat(name, f):
        v = [r[name][f] for r in rows]
        return float(np.mean(v))

    summary = {"frames": len(rows), "learned_mode": args.learned,
               "plain_mgpcg": {f: stat("plain_mgpcg", f) for f in ("ms", "device_ms", "iters", "rel_l2_err_vs_fp64")},
               "learned_fcg": {f: stat("learned_fcg", f) for f in ("ms", "device_ms", "iters", "rel_l2_err_vs_fp64")}}
    summary["speedup_learned_vs_plain"] = summary["plain_mgpcg"]["ms"] / summary["learned_fcg"]["ms"]
    summary["max_err_plain"] = max(r["plain_mgpcg"]["rel_l2_err_vs_fp64"] for r in rows)
    summary["max_err_learned"] = max(r["learned_fcg"]["rel_l2_err_vs_fp64"] for r in rows)
    print(json.dumps({"summary": summary}), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"frames": rows, "summary": summary}, f, indent=1)


if __name__ == "__main__":
    main()
