"""Learned Newton initial guess on the GPU.

The HIP training kernels (FP32 forward, backward, Adam) are checked against the FP64 numpy restatement
(oracle/mlp_ref.py, itself pinned by finite differences); the BF16 MFMA inference kernel against an emulation of its
operand roundings; the stepper's use of the guess by the plain guess's role as fallback (an untrained network must leave
the run bit-identical to the plain one) and by what a trained network gains on an unseen seed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf16(a):
    """Round FP32 values to bfloat16 (nearest, ties to even), returned as FP64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _net_and_data(seed, B):
    from diffusionimex_amd import mlp
    rng = np.random.default_rng(seed)
    P = mlp.init_params(seed)
    P[mlp.N_IN * mlp.N_H:mlp.N_IN * mlp.N_H + mlp.N_H] = rng.normal(scale=0.1, size=mlp.N_H)   # nonzero biases
    P[-1] = 0.05
    X = rng.normal(size=(B, mlp.N_IN)).astype(np.float32)
    Y = rng.normal(size=B).astype(np.float32)
    net = mlp.Network()
    net.set_params(P)
    return net, P, X, Y


def test_fp32_forward_matches_fp64_reference():
    from oracle import mlp_ref as R
    net, P, X, _ = _net_and_data(1, 300)
    y = net.forward(X)
    yr, _ = R.forward(P.astype(np.float64), X.astype(np.float64))
    # FP32 arithmetic over 175- and 128-term sums: a few FP32 ulps of the sum of |terms|
    assert np.max(np.abs(y - yr)) <= 2e-5 * (1 + np.max(np.abs(yr))), np.max(np.abs(y - yr))


@pytest.mark.parametrize("B", [1, 37, 512])
def test_fp32_gradients_match_fp64_reference(B):
    from oracle import mlp_ref as R
    net, P, X, Y = _net_and_data(2 + B, B)
    g, loss = net.gradients(X, Y)
    lr, gr = R.loss_grad(P.astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - lr) <= 1e-5 * (lr + 1.0)   # y - t cancels: FP32 accuracy relative to the terms
    # every block of the parameter vector (W1, b1, W2, b2, w3, b3) to FP32 accuracy relative to its own size
    from diffusionimex_amd import mlp
    cuts = np.cumsum([0, mlp.N_IN * mlp.N_H, mlp.N_H, mlp.N_H * mlp.N_H, mlp.N_H, mlp.N_H, 1])
    for a, b in zip(cuts[:-1], cuts[1:]):
        err = np.linalg.norm(g[a:b] - gr[a:b]) / (np.linalg.norm(gr[a:b]) + 1e-30)
        assert err <= 2e-5, (a, b, err)


def _gelu_tanh(x):
    """the inference kernel's GELU: the tanh form x sigma(2u), u = sqrt(2/pi) (x + 0.044715 x^3)"""
    u = np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)
    return x / (1 + np.exp(-2 * u))


def test_bf16_mfma_inference_matches_its_rounding_model():
    """Inputs, W1, b1, W2 and the first hidden layer are rounded to BF16; products accumulate in FP32 on MFMA; both GELUs
    take the tanh form (within 4.8e-4 of the erf form the network is trained with) and the second GELU and the head run
    in FP32. The kernel must agree with that model up to FP32 accumulation order, the native exp2 / reciprocal and rare
    BF16 ties of the hidden layer, and stay within BF16 accuracy of the FP64 network."""
    from oracle import mlp_ref as R
    net, P, X, _ = _net_and_data(5, 1000)
    y = net.infer_bf16(X)
    W1, b1, W2, b2, w3, b3 = R.unpack(P.astype(np.float64))
    Z1 = _bf16(X) @ _bf16(W1) + _bf16(b1)   # the bias enters the MFMA through a constant-1 input slot
    H1 = _bf16(_gelu_tanh(Z1).astype(np.float32))
    H2 = _gelu_tanh(H1 @ _bf16(W2) + b2)
    ym = H2 @ w3 + b3
    scale = np.abs(H2) @ np.abs(w3) + abs(b3)
    dev = np.abs(y - ym) / scale
    assert np.median(dev) <= 1e-6 and np.max(dev) <= 2e-3, (np.median(dev), np.max(dev))
    yr, _ = R.forward(P.astype(np.float64), X.astype(np.float64))
    rel = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    assert rel <= 2e-2, rel


def test_adam_training_fits_a_smooth_target():
    from diffusionimex_amd import mlp
    rng = np.random.default_rng(3)
    X = rng.normal(size=(16384, mlp.N_IN)).astype(np.float32)
    Y = (np.sin(X[:, 0]) + 0.5 * X[:, 3] * X[:, 7] - 0.3 * X[:, 100]).astype(np.float32)
    net = mlp.Network()
    net.set_params(mlp.init_params(4), np.concatenate([np.zeros(mlp.N_IN), np.ones(mlp.N_IN), [0, 1]]).astype(np.float32))
    net.set_data(X, Y)
    loss = net.train(25, batch=1024, lr=2e-3, seed=1)
    assert np.all(np.isfinite(loss))
    assert loss[-1] < 0.1 * loss[0], loss
    # the trained FP32 and BF16 paths agree on held-out data
    Xt = rng.normal(size=(2000, mlp.N_IN)).astype(np.float32)
    yt = (np.sin(Xt[:, 0]) + 0.5 * Xt[:, 3] * Xt[:, 7] - 0.3 * Xt[:, 100])
    y32, y16 = net.forward(Xt), net.infer_bf16(Xt)
    assert np.mean((y32 - yt) ** 2) < 0.5 * np.var(yt)
    assert np.linalg.norm(y16 - y32) <= 0.05 * np.linalg.norm(y32)


def _bubble(seed, **kw):
    from diffusionimex_amd import cases
    return cases.shock_bubble(128, 64, seed=seed, cfl=0.5, **kw)


def test_untrained_network_falls_back_to_the_plain_guess_bitwise():
    """Newton starts from the plain guess whenever the network's guess has the larger residual, so a random network
    must leave every stage, and the whole run, identical to the plain run."""
    from diffusionimex_amd import mlp, spb6
    cfg, U = _bubble(3)
    a = spb6.Stepper(cfg); a.set_state(U); assert a.run(1.0, 6) == 0
    cfgl, _ = _bubble(3, learned_guess=1)
    b = spb6.Stepper(cfgl); b.set_state(U)
    net = mlp.Network()
    net.set_params(mlp.init_params(11), np.concatenate([np.zeros(mlp.N_IN), np.ones(mlp.N_IN), [0.0, 1.0]]).astype(np.float32))
    mlp.attach(b, net)
    assert b.run(1.0, 6) == 0
    s = b.stats()
    assert s["guess_fallback"] == s["stages"] and s["guess_used"] == 0
    assert np.array_equal(a.state(), b.state())
    assert np.array_equal(a.pressure()[0], b.pressure()[0])


def test_trained_guess_gains_over_the_plain_guess_on_an_unseen_seed():
    """Train on solver dumps of three seeded shock-bubble runs, then run a fourth seed: the learned guess must start
    Newton closer to the converged pressure than the plain guess (the previous stage's P), on average by at least 3x in
    the residual norm, with the same step sequence and no more Newton iterations."""
    from diffusionimex_amd import mlp, spb6
    runs = [_bubble(s) for s in (1, 2, 4)]
    X, Y = mlp.collect(runs, steps=16, stride=1, warm=2, capacity=600_000)
    assert X.shape[0] > 20_000 and np.all(np.isfinite(X)) and np.all(np.isfinite(Y))
    net, norm, loss = mlp.train_network(X, Y, epochs=30, batch=4096, lr=1e-3, seed=0)
    assert loss[-1] < 0.5 * loss[0]
    cfg, U = _bubble(9)
    cfgl, _ = _bubble(9, learned_guess=1)
    a = spb6.Stepper(cfg); a.set_state(U); assert a.run(1.0, 2) == 0
    b = spb6.Stepper(cfgl); b.set_state(U); assert b.run(1.0, 2) == 0
    mlp.attach(b, net)
    s0a, s0b = a.stats(), b.stats()
    assert a.run(1.0, 16) == 0 and b.run(1.0, 16) == 0
    sa, sb = a.stats(), b.stats()
    stages = sb["stages"] - s0b["stages"]
    gain = (sb["guess_log10_gain"] - s0b["guess_log10_gain"]) / stages
    nt_plain = (sa["newton_it"] - s0a["newton_it"]) / (sa["stages"] - s0a["stages"])
    nt_learn = (sb["newton_it"] - s0b["newton_it"]) / stages
    print(f"learned guess: mean log10 residual gain {gain:.2f} per stage, used {sb['guess_used'] - s0b['guess_used']}"
          f"/{stages}, Newton per stage {nt_learn:.2f} (plain {nt_plain:.2f})")
    assert sa["steps"] == sb["steps"] and sb["rejects"] == sa["rejects"] == 0
    assert sb["maxres"] < 1e-13
    assert gain >= np.log10(3.0)
    assert nt_learn <= nt_plain


@pytest.mark.parametrize("case", ["bubble", "cavitation_periodic", "cavitation_ragged"])
def test_guess_residual_on_the_changed_cells_equals_the_full_evaluation_bitwise(case):
    """A stage computes the learned guess from the features of the listed tiles' stencils only, and evaluates the
    guess's residual only on the cells the guess can change (an active cell and its four neighbours, periodic partners
    included), taking the plain guess's residual elsewhere: the guess must equal the one from the features of every
    cell, and that residual and its max-norm the residual evaluated on every cell at the guess, bit for bit, on a mix
    of both kinds of cell."""
    from diffusionimex_amd import cases, mlp, spb6
    if case == "bubble":
        cfg, U = _bubble(5, learned_guess=1)
    elif case == "cavitation_periodic":
        cfg, U = cases.cavitation_2d(96, 64, seed=2, learned_guess=1, bc=(spb6.BC_PERIODIC,) * 4)
    else:   # nx not a multiple of 16: tiles straddle grid rows, the features of every cell are computed
        cfg, U = cases.cavitation_2d(90, 40, seed=4, learned_guess=1)
    s = spb6.Stepper(cfg); s.set_state(U)
    net = mlp.Network()
    net.set_params(mlp.init_params(7), np.concatenate([np.zeros(mlp.N_IN), np.ones(mlp.N_IN), [0.0, 1.0]]).astype(np.float32))
    mlp.attach(s, net)
    assert s.run(1.0, 3) == 0
    adt = 0.29 * s.stats()["dt_last"]
    assert s.explicit_stage(adt) == 0 and s.stage_prep(adt) == 0
    P = s.pressure()[0].ravel()
    Pg, Pgf, G, Gf, nrm, cnt = s.guess_residual(P)
    assert np.array_equal(Pg.view(np.int64), Pgf.view(np.int64))   # features of the listed tiles' stencils only
    n = P.size
    changed = np.count_nonzero(Pg != P)
    print(f"{case}: {changed} cells changed, {cnt} of {n} evaluated, norm {nrm[0]:.3e}")
    assert 0 < changed <= cnt <= n
    if case == "bubble":   # quiescent gas away from the shock and the bubble: many cells keep the plain residual
        assert cnt < n
    assert np.array_equal(G.view(np.int64), Gf.view(np.int64))
    assert nrm[0] == nrm[1] and nrm[0] > 0
