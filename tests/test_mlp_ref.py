"""The FP64 numpy restatement of the learned-guess network (oracle/mlp_ref.py) against central finite differences:
it is the reference the HIP training kernels' gradients are checked against."""
import numpy as np


def test_mlp_reference_gradient_matches_finite_differences():
    from oracle import mlp_ref as R
    rng = np.random.default_rng(0)
    P = rng.normal(scale=0.1, size=R.N_IN * R.N_H + R.N_H + R.N_H * R.N_H + R.N_H + R.N_H + 1)
    X = rng.normal(size=(6, R.N_IN))
    T = rng.normal(size=6)
    loss, g = R.loss_grad(P, X, T)
    idx = np.concatenate([rng.integers(0, P.size, 40), [P.size - 1, R.N_IN * R.N_H + 3]])
    for i in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (R.loss_grad(Pp, X, T)[0] - R.loss_grad(Pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (abs(g[i]) + 1e-3), (i, fd, g[i])
