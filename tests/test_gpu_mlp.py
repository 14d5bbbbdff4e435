"""The warm-start MLP's BF16 matrix-core kernels against the FP64 numpy
reference (forward, gradients), training, and the engine hook."""
import numpy as np
import pytest

import mlp_ref as M
from shapes_amd import engine, scenes

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 14)).astype(np.float32)
    T = (np.sin(X[:, :2]) + 0.3 * X[:, 2:4]).astype(np.float32)
    return X, T


def test_forward_matches_reference():
    p = M.init_params(5)
    X, _ = _data(300, 1)
    net = engine.Mlp(p)
    y = net.forward(X)
    ref, _, _ = M.forward(p, X)
    err = np.abs(y - ref).max()
    scale = np.abs(ref).max()
    print(f"forward max abs err {err:.3e} (scale {scale:.3e})")
    # FP32 accumulation order differs from numpy's FP64 sums; the BF16
    # rounding of hidden activations can flip on those differences
    assert err <= 2e-2 * scale + 1e-3
    net.close()


def test_gradient_matches_reference():
    p = M.init_params(6)
    X, T = _data(256, 2)
    net = engine.Mlp(p)
    g, loss = net.gradient(X, T)
    gr, lr = M.gradient(p, X, T)
    assert abs(loss - lr) <= 1e-2 * lr
    k = 0
    for name, shape in M.SHAPES:
        n = int(np.prod(shape))
        a, b = g[k:k + n].astype(np.float64), gr[k:k + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print(f"{name}: rel err {rel:.3e} cos {cos:.6f}")
        assert rel < 5e-2 and cos > 0.998, name
        k += n
    net.close()


def test_training_reduces_loss():
    X, T = _data(8192, 3)
    net = engine.Mlp(M.init_params(7))
    net.set_data(X, T)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.permutation(len(X)) for _ in range(20)])
    batch = 512
    idx = idx[: (len(idx) // batch) * batch]
    loss = net.train(idx, batch, 2e-3, want_loss=True)
    print(f"loss {loss[:3]} ... {loss[-3:]}")
    assert loss[-20:].mean() < 0.25 * loss[:5].mean()
    # the trained weights reproduce the reference forward pass
    y = net.forward(X[:256])
    ref, _, _ = M.forward(net.params(), X[:256])
    assert np.abs(y - ref).max() <= 2e-2 * np.abs(ref).max() + 1e-3
    net.close()


def test_engine_logs_and_uses_learned_start():
    sc = scenes.pyramid(rows=6)
    w = sc.worlds[0]
    b = engine.make_batch([w], sc.config)
    b.set_logging(100000)
    b.step(40)
    f, t = b.read_log()
    assert len(f) > 20 and np.isfinite(f).all() and np.isfinite(t).all()
    b.close()
    # a model that predicts the logged targets exactly is not required; any
    # finite model must produce projected (feasible) warm starts
    lb = engine.make_batch([w], sc.config, warm_start=engine.WARM_LEARNED)
    lb.set_mlp(M.init_params(1))
    lb.step(40)
    s = lb.stats()
    assert s["misses"] >= 0
    c = lb.contacts()
    mu = (w.mu[c["keys"][:, 0]] + w.mu[c["keys"][:, 1]]) / 2
    assert (c["lam"][:, 0] >= 0).all()
    assert (np.abs(c["lam"][:, 1]) <= mu * c["lam"][:, 0] + 1e-12).all()
    lb.close()


def test_learned_start_counts_misses():
    """Every new contact of the first contact step gets a learned start."""
    sc = scenes.pyramid(rows=4)
    w = sc.worlds[0]
    lb = engine.make_batch([w], sc.config, warm_start=engine.WARM_LEARNED)
    lb.set_mlp(M.init_params(2))
    lb.step(1)
    s = lb.stats()
    c = lb.contacts()
    assert s["misses"] == (c["warm"] != 1).sum() > 0
    assert s["learned_starts"] == (c["warm"] == 2).sum()
    assert s["learned_starts"] > 0
    lb.close()


def _pile_quality(model, iterations, steps=600):
    sc = scenes.polygon_pile(iterations=iterations)
    ws = engine.WARM_LEARNED if model is not None else engine.WARM_ZERO
    b = engine.make_batch(sc.worlds, sc.config, solver=engine.SOLVER_COLORED_F32, warm_start=ws)
    if model is not None:
        b.set_mlp(model)
    b.step(steps)
    s = b.stats()
    b.close()
    return s


def test_trained_model_beats_zero_start():
    """The committed network (shapes_amd/warmstart_mlp.npy, trained by the
    library's own kernels on 5M contacts of polygon piles with other seeds)
    against the zero start of new contacts on the 8192-polygon pile at 4 and
    2 sweeps: lower run means of the per-step maximum normal residual and
    penetration.  The bound for any predictor of new contacts' impulses is
    measured by tests/test_warmstart_bound.py."""
    from shapes_amd import warmstart
    model = np.load(warmstart.DEFAULT_MODEL)
    for it in (4, 2):
        z = _pile_quality(None, it)
        m = _pile_quality(model, it)
        print(f"{it} sweeps: zero start mean normal residual {z['mean_normal_residual']:.4f} "
              f"penetration {z['mean_max_depth']:.4f}; learned {m['mean_normal_residual']:.4f} "
              f"{m['mean_max_depth']:.4f}")
        assert m["mean_normal_residual"] < 0.97 * z["mean_normal_residual"]
        assert m["mean_max_depth"] < z["mean_max_depth"]
