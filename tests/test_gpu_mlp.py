"""The scheduler's cost-model MLP on the GPU.

* device forward and backward equal the CPU restatement (tests/mlp_ref.py,
  itself pinned by finite differences) within float32 tolerance;
* device training (Adam) fits a smooth target;
* ablation: a model trained on the decomposition's own pass logs, attached
  as the scheduler, gives the oracle's bits, and its time is reported next
  to the work-share rule's.
"""
import sys

import numpy as np
import pytest

import dcx
import mlp_ref
import oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def test_forward_and_gradient_match_cpu():
    rng = np.random.default_rng(7)
    say("mlp forward/backward vs CPU")
    with dcx.Context(0) as c:
        m = dcx.MLP(c, seed=3)
        p = rng.normal(0, 0.25, m.n_params).astype(np.float32)
        m.params = p
        assert np.array_equal(m.params, p)
        X = rng.uniform(0, 1, (300, 8)).astype(np.float32)
        T = rng.normal(0, 1, 300).astype(np.float32)
        y = m.predict(X)
        y_ref, _ = mlp_ref.forward(p, X)
        assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-4)
        g, loss = m.gradient(X, T)
        g_ref, loss_ref = mlp_ref.gradient(p, X, T)
        scale = np.abs(g_ref).max()
        assert np.abs(g - g_ref).max() <= 2e-4 * scale
        assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref)
        m.free()


def test_training_fits_smooth_target():
    rng = np.random.default_rng(11)
    X = rng.uniform(0, 1, (512, 8)).astype(np.float32)
    T = (np.sin(3 * X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 3]).astype(np.float32)
    say("mlp training")
    with dcx.Context(0) as c:
        m = dcx.MLP(c, seed=5)
        y0 = m.predict(X)
        l0 = 0.5 * float(((y0 - T) ** 2).mean())
        l1 = m.train(X, T, epochs=1500, lr=3e-3)
        y1 = m.predict(X)
        l2 = 0.5 * float(((y1 - T) ** 2).mean())
        m.free()
    assert l2 < l0 / 10 and l2 < 0.02, (l0, l1, l2)


def test_learned_scheduler_bit_exact_and_ablation():
    src, dst = dcx.gen_rmat(16, 16 << 16, 13)
    V = 1 << 16
    ref = oracle.decompose(oracle.csr_from_edges(V, src, dst))
    Xs, Ts = [], []
    heur = []
    for tm in (2, 4, 8, 16, 32):
        say(f"rmat16 heuristic team_max={tm}")
        # chains off: the model schedules peeled team rows, and every team
        # row of this graph would otherwise be derived from row 0
        with dcx.Context(0, team_max=tm, chain_rows=1) as c:
            g = c.upload_edges(V, src, dst)
            for _ in range(2):
                g.decompose(dcx.VARIANT_BATCHED)
            if tm == 32:
                heur.append(c.timing().rows_kernel_ms)
            for wl in (True, False):
                X, T = g.pass_features(with_log=wl)
                Xs.append(X)
                Ts.append(T)
            g.free()
    X = np.concatenate(Xs)
    T = np.concatenate(Ts)
    say(f"train on {len(T)} samples; schedule with the model")
    with dcx.Context(0, chain_rows=1) as c:
        m = dcx.MLP(c, seed=1)
        m.train(X, T, epochs=2000, lr=3e-3)
        pred = m.predict(X)
        corr = float(np.corrcoef(pred, T)[0, 1])
        g = c.upload_edges(V, src, dst)
        g.decompose(dcx.VARIANT_BATCHED)  # logs for the with-log features
        c.set_scheduler(m)
        times = []
        for _ in range(2):
            K, lmax = g.decompose(dcx.VARIANT_BATCHED)
            assert c.timing().rows_scheduler == 1
            times.append(c.timing().rows_kernel_ms)
        assert K == ref[0] and lmax == ref[2]
        assert np.array_equal(g.kmaxes(), ref[1]) and np.array_equal(g.rows(), ref[3])
        c.set_scheduler(None)
        g.decompose(dcx.VARIANT_BATCHED)
        assert c.timing().rows_scheduler == 0
        g.free()
        m.free()
    print(f"[ablation] rmat16 rows kernel: work-share rule {min(heur):.2f} ms, "
          f"learned widths {min(times):.2f} ms; fit corr {corr:.3f}", flush=True)
    assert corr > 0.6
