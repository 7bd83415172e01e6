"""The learned tile smoother on the GPU: MFMA operand layout, training gradients
and Adam against the FP64 oracle, BF16 MFMA inference against the oracle's
forward pass, and the learned flexible-CG solve."""
import numpy as np
import pytest

import cirrus_amg as P
import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu


def _bf16(x):
    """round float32 to bfloat16 (nearest, ties to even); returns float32 values"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("K", [16, 32])
def test_mfma_operand_layout(K):
    rng = np.random.default_rng(1)
    A = rng.integers(-8, 9, (16, K)).astype(np.float32)
    B = rng.integers(-8, 9, (K, 16)).astype(np.float32)
    bits = lambda m: (m.view(np.uint32) >> 16).astype(np.uint16)
    C = P.mfma_probe(bits(A), bits(B))
    assert np.array_equal(C, A @ B)


def _tiles(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 1000)).astype(np.float32)
    e = rng.normal(0, 1, (n, 1000)).astype(np.float32)
    return x, e


def test_training_gradient_matches_oracle():
    x, e = _tiles(6, 2)
    tr = P.Trainer(16, seed=5)
    tr.add_tiles(x, e)
    w = tr.weights()
    loss, g = tr.loss_grad(0, 6, bf16=False)
    oloss, og = O.net_loss_grad(w.astype(np.float64), x, e)
    assert abs(loss - oloss) <= 1e-4 * oloss
    assert np.linalg.norm(g - og) <= 1e-4 * np.linalg.norm(og)
    # the BF16 forward pass stays close to the FP32 one
    lb, gb = tr.loss_grad(0, 6, bf16=True)
    assert abs(lb - oloss) <= 5e-2 * oloss
    assert np.linalg.norm(gb - og) <= 1e-1 * np.linalg.norm(og)


def test_adam_step_matches_oracle():
    x, e = _tiles(4, 3)
    tr = P.Trainer(8, seed=6)
    tr.add_tiles(x, e)
    w0 = tr.weights().astype(np.float64)
    _, g = tr.loss_grad(0, 4, bf16=False)
    tr.step(4, lr=1e-3, bf16=False)     # the batch is a permutation of the same 4 tiles
    w1 = tr.weights().astype(np.float64)
    m, v = np.zeros_like(w0), np.zeros_like(w0)
    wo = w0.copy()
    O.adam_step(wo, m, v, g.astype(np.float64), 1, lr=1e-3)
    assert np.abs(w1 - wo).max() <= 1e-6


def _gather_inputs(go, tiles, r, dch):
    """the inference input volumes (r / D on INTERIOR cells of the tile and its
    same-level LEAF neighbours) restated with numpy"""
    lv, co, ty = go.tiles()
    T = go.cell_types()
    nbr = go.neighbors()
    R, D = go.channel(r), go.channel(dch)
    feat = np.where((T == 2) & (D != 0), R / np.where(D != 0, D, 1), 0).astype(np.float32)
    out = np.zeros((len(tiles), 10, 10, 10), np.float32)
    for b, t in enumerate(tiles):
        out[b, 1:9, 1:9, 1:9] = feat[t].reshape(8, 8, 8)
        for a in range(3):
            for side, q in ((-1, a), (1, a + 3)):
                nt = nbr[t, q]
                if nt < 0 or not (ty[nt] & 1):
                    continue
                src = feat[nt].reshape(8, 8, 8)
                sl_dst = [slice(1, 9)] * 3
                sl_src = [slice(0, 8)] * 3
                sl_dst[a] = 0 if side < 0 else 9
                sl_src[a] = 7 if side < 0 else 0
                out[b][tuple(sl_dst)] = src[tuple(sl_src)]
    return out.reshape(len(tiles), 1000)


@pytest.mark.parametrize("case", ["sphere_empty24", "c3_small"])
def test_mfma_inference_matches_oracle_forward(case):
    go, gp = H.build_pair(case)
    rng = np.random.default_rng(4)
    r = rng.uniform(-1, 1, (go.N, 512)).astype(np.float32)
    go.channel(1)[:] = r
    gp.set_channel(1, r)
    w = rng.normal(0, 0.15, P.NET_WEIGHTS).astype(np.float32)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    solver.set_net(w)
    solver.net_infer(1, H.C0 + 3, 6)
    y = gp.channel(6)
    lv, co, ty = go.tiles()
    tiles = np.where((lv == go.max_level) & (ty == 1))[0][:24]
    X = _gather_inputs(go, tiles, 1, H.C0 + 3)
    T = go.cell_types()
    worst = 0.0
    for b, t in enumerate(tiles):
        yo = O.net_forward(_bf16(w).astype(np.float64), _bf16(X[b]).astype(np.float64)).reshape(10, 10, 10)[1:9, 1:9, 1:9]
        yo = np.where(T[t] == 2, yo.reshape(512), 0)
        scale = max(np.abs(yo).max(), 1e-30)
        worst = max(worst, np.abs(y[t] - yo).max() / scale)
    assert worst <= 3e-2, worst


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("case", ["sphere_empty24", "c3_small"])
def test_learned_fcg_solve_converges(case, mode):
    """With any weights (here untrained) the learned flexible-CG solve reaches the
    tolerance: tiles whose residual the net does not lower fall back to Jacobi."""
    go, gp = H.build_pair(case)
    go.rhs_uniform(1, 77, True)
    gp.set_channel(1, go.channel(1))
    A = go.assemble_csr(H.C0)
    b = go.channel(1).ravel()[A.cell].astype(np.float64)
    x64, _, _ = A.bicgstab(b, 1e-13, 20000)
    solver = P.AMGSolver(gp, H.C0, H.R_MATRIX, 1.0, 1.0)
    solver.set_net(np.random.default_rng(8).normal(0, 0.1, P.NET_WEIGHTS).astype(np.float32))
    opts = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=2.0 / 3.0, coarsest_direct=1,
                          flexible=1, learned=mode)
    it, rel = solver.solve(max_iters=200, relative_tolerance=1e-6, coarsest_iters=0, opts=opts)
    assert rel < 1e-6, (it, rel)
    judged, kept = solver.learned_counts()
    assert judged > 0 and 0 <= kept <= judged
    x = gp.channel(0).ravel()[A.cell]
    assert np.linalg.norm(x - x64) <= 1e-5 * np.linalg.norm(x64)
