"""The rebuild-interval network on the GPU: forward pass and gradients of the
library's own kernels against the CPU restatement (oracle/mlp_ref.py),
training, and what the trained network gains in the guarded stepping while
every result stays bit-identical."""
import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _rand(seed=0, n=300):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.15, size=M.NPARAMS).astype(np.float32)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    return p, X, t


def test_forward_and_gradient_match_cpu_reference():
    """FP32 kernels against the float64 reference: forward within 1e-5,
    every gradient component within 1e-4 of the gradient's scale."""
    from rockable_amd.engine import Mlp
    p, X, t = _rand()
    mlp = Mlp(seed=3)
    mlp.set_params(p)
    assert np.array_equal(mlp.params(), p)
    y = mlp.forward(X)
    yr, _ = M.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(y - yr)) < 1e-5 * max(1.0, np.max(np.abs(yr)))
    loss, g = mlp.loss_grad(X, t)
    lr, gr = M.loss_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lr) < 1e-5 * max(1.0, lr)
    assert np.max(np.abs(g - gr)) < 1e-4 * np.max(np.abs(gr))


def test_training_fits_a_smooth_function():
    """Adam with the library's kernels: on t = sin(2 x0) + x1^2 / 2 the loss
    falls by more than 20x."""
    from rockable_amd.engine import Mlp
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(2048, 16)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) + 0.5 * X[:, 1] ** 2).astype(np.float32)
    mlp = Mlp(seed=7)
    l0, _ = mlp.loss_grad(X, t)
    l1 = mlp.train(X, t, epochs=200, batch=256, lr=3e-3, seed=1)
    assert l1 < l0 / 20.0, (l0, l1)


def test_trained_network_sizes_batches_without_changing_results(tmp_path):
    """Histories of two C2-style runs train the network; on a third run
    (another seed) the guarded stepping with the network reads the guard far
    less often than the plain method (one read per iteration) and wastes
    fewer suspended iterations than a fixed batch of 256, and all three
    give bit-identical states and rebuild counts."""
    from rockable_amd import rebuild_model as RM
    from rockable_amd.engine import Engine
    Xs, ts = [], []
    for k, (seed, disp, ang) in enumerate(((11, 2e-4, 4.0), (12, 3e-4, 5.0))):
        d = tmp_path / ("train%d" % k)
        d.mkdir()
        e = Engine(RM.hull_deck(str(d), seed, disp, ang))
        e.update_nl()
        X, t, n = RM.collect(e, 4000, 60)
        assert n >= 10
        Xs.append(X)
        ts.append(t)
    mlp, loss = RM.train(np.concatenate(Xs), np.concatenate(ts), epochs=200)
    tt = np.concatenate(ts)
    assert loss < 0.5 * np.var(tt)
    d = tmp_path / "test"
    d.mkdir()
    path = RM.hull_deck(str(d), 13, 4e-4, 6.0)
    runs = {}
    for name in ("plain", "fixed256", "network"):
        e = Engine(path)
        e.update_nl()
        if name == "plain":
            e.set_option("guard_batch", 1)
        elif name == "fixed256":
            e.set_option("guard_batch", 256)
        else:
            e.set_rebuild_model(mlp)
        e.step(1500)
        runs[name] = (e.bodies(), e.counts())
    b0, c0 = runs["plain"]
    for name in ("fixed256", "network"):
        b, c = runs[name]
        for k in ("pos", "vel", "Q", "vrot", "force", "moment"):
            assert np.array_equal(b[k], b0[k]), (name, k)
        assert c["nl_updates"] == c0["nl_updates"]
    assert c0["nl_updates"] > 10
    cp, cf, cn = c0, runs["fixed256"][1], runs["network"][1]
    assert cn["guard_reads"] * 3 < cp["guard_reads"], (cn["guard_reads"], cp["guard_reads"])
    assert cn["guard_wasted"] < cf["guard_wasted"], (cn["guard_wasted"], cf["guard_wasted"])
