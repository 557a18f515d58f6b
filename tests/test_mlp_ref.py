"""The CPU restatement of the rebuild-interval network (oracle/mlp_ref.py):
its analytic gradient against central finite differences, in float64."""
import numpy as np

from oracle import mlp_ref as M


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(scale=0.2, size=M.NPARAMS)
    X = rng.normal(size=(40, 16))
    t = rng.normal(size=40)
    _, g = M.loss_grad(p, X, t)
    h = 1e-6
    for k in rng.choice(M.NPARAMS, 60, replace=False):
        e = np.zeros(M.NPARAMS)
        e[k] = h
        fd = (M.loss_grad(p + e, X, t)[0] - M.loss_grad(p - e, X, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_reference_layout_matches_the_abi_header():
    """W1[64][16] b1[64] W2[64][64] b2[64] W3[64] b3[1] = 5313 floats, as
    include/rockable_amd.h declares."""
    assert M.NPARAMS == 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1 == 5313
