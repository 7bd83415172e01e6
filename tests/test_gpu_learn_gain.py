"""What the trained tile smoother gains over the method without it.

Held-out data: frame 150 of the moving-sphere scene (training used frames 0..99),
finest 256^3, the solver's own MGPCG iterates (cirrus-amg_amd/train_smoother.py).

* Per tile, on the (r/D, e) pairs of that frame, the training loss
  ||L(e_hat - e)||^2 (L the 7-point Laplacian on the tile's interior) of three
  smoothers of the same architecture: the trained net, the plain first damped-Jacobi
  step from zero (e_hat = w r/D, written as a net: two ReLU channels carrying +x and
  -x) and a random net.  The trained net must beat the Jacobi step it replaces.
* In flexible CG preconditioned by the V-cycle, the residual reduction per
  iteration over 10 iterations with the learned pre-smoothing (trained and random
  weights) against the plain cycle, and the share of tiles that keep the net's
  correction.  The trained cycle may not be slower to converge than the plain one;
  random weights are rejected on almost every tile (the fallback keeps the plain
  sweep).
"""
import os
import sys

import numpy as np
import pytest

import cirrus_amg as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cirrus-amg_amd"))
import train_smoother as T  # noqa: E402

WEIGHTS = os.path.join(ROOT, "cirrus-amg_amd", "data", "smoother_c4.f32")
OMEGA = 2.0 / 3.0
HELD_OUT = 150


def _jacobi_net():
    """e_hat = w x as a net of the smoother's architecture (1 -> 16 -> 16 -> 1, 3^3 taps, ReLU)"""
    w = np.zeros(P.NET_WEIGHTS, np.float32)
    c = 13                                   # centre tap
    w[0 * 27 + c], w[1 * 27 + c] = 1.0, -1.0                       # layer 1: relu(x), relu(-x)
    w[432 + (0 * 16 + 0) * 27 + c] = 1.0                           # layer 2: carry both channels
    w[432 + (1 * 16 + 1) * 27 + c] = 1.0
    w[7344 + 0 * 27 + c], w[7344 + 1 * 27 + c] = OMEGA, -OMEGA     # layer 3: w (relu(x) - relu(-x))
    return w


def _random_net(seed=5):
    return np.random.default_rng(seed).normal(0, 0.1, P.NET_WEIGHTS).astype(np.float32)


def test_trained_net_beats_the_jacobi_step_it_replaces():
    trained = np.fromfile(WEIGHTS, "<f4")
    tr = P.Trainer(100_000, seed=3)
    n = T.collect_frame(tr, HELD_OUT, 3)
    assert n > 1000
    losses = {}
    for name, w in (("trained", trained), ("jacobi", _jacobi_net()), ("random", _random_net())):
        tr.set_weights(w)
        losses[name], _ = tr.loss_grad(0, tr.size, bf16=False)
    print("\nheld-out frame %d, %d tiles: loss trained %.4e, Jacobi step %.4e, random %.4e"
          % (HELD_OUT, tr.size, losses["trained"], losses["jacobi"], losses["random"]), flush=True)
    assert losses["trained"] < 0.9 * losses["jacobi"]
    assert losses["trained"] < losses["random"]


def test_learned_fcg_convergence_rate():
    """residual reduction per flexible-CG iteration over a fixed number of iterations"""
    g = T.frame_grid(HELD_OUT)
    solver = P.AMGSolver(g, T.C0, T.R_MATRIX, 1.0, 1.0)
    its = 10
    rows = {}
    for name, w, learned in (("plain", None, 0), ("trained", np.fromfile(WEIGHTS, "<f4"), 1),
                             ("random", _random_net(), 1)):
        if w is not None:
            solver.set_net(w)
        o = P.default_opts(cycle=P.CYCLE_FAS_V, smoother=P.SMOOTH_JACOBI, omega=OMEGA, coarsest_direct=1, flexible=1,
                           learned=learned)
        g.copy_channel(T.CH_B, 1)
        it, rel = solver.solve(opts=o, max_iters=its, relative_tolerance=0.0, coarsest_iters=0)
        assert it == its
        rate = rel ** (1.0 / its)
        kept = None
        if w is not None:
            judged, k = solver.learned_counts()
            kept = k / max(judged, 1)
        rows[name] = (rate, kept)
        print("\n%s: residual reduction per FCG iteration %.4f%s" % (name, rate, "" if kept is None else
                                                                   ", tiles keeping the net's correction %.2f" % kept),
              flush=True)
    # the trained smoother does not slow the iteration, and its corrections are kept where a
    # random net's are not
    assert rows["trained"][0] <= 1.05 * rows["plain"][0]
    assert rows["trained"][1] > 0.5 and rows["random"][1] < 0.2
