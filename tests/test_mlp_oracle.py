"""The MLP oracle's analytic gradient against central finite differences (FP64)."""
import numpy as np

from oracle import mlp_oracle as M


def test_mlp_oracle_gradient_finite_difference():
    rng = np.random.default_rng(0)
    F = 5
    n_par = F * M.H + M.H + M.H * M.H + M.H + M.H + 1
    p = rng.normal(0, 0.3, n_par)
    X = rng.normal(size=(17, F))
    y = rng.normal(size=17)
    _, g = M.loss_grad(p, X, y)
    idx = rng.choice(n_par, 60, replace=False)
    idx = np.concatenate([idx, [0, F * M.H, n_par - 1, n_par - 2]])     # W1, b1, b3, W3 entries
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, X, y)[0] - M.loss_grad(pm, X, y)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])


def test_mlp_oracle_forward_shapes_and_linearity_of_output_layer():
    rng = np.random.default_rng(1)
    F = 3
    n_par = F * M.H + M.H + M.H * M.H + M.H + M.H + 1
    p = rng.normal(0, 0.3, n_par)
    X = rng.normal(size=(4, F))
    y0, (h1, h2) = M.forward(p, X)
    p2 = p.copy()
    p2[-1] += 1.5                               # b3 shifts every output by 1.5
    assert np.allclose(M.forward(p2, X)[0], y0 + 1.5)
    assert h1.shape == (4, M.H) and h2.shape == (4, M.H)
