"""Learned Newton initial guess: the network, its training data and its use by the stepper.

The network (175 -> 128 GELU -> 128 GELU -> 1) maps the 5x5 stencil of 7 stage fields of a cell (explicit predictor
alpha1, log rho, u*, v*, mixture pressure, plain pressure guess, frozen sound speed; relative to the centre cell) to the
stage's pressure change (P_converged - P_plain) / K_w. Training pairs are dumped by the solver itself from plain-guess
runs; training and inference run on the GPU (FP32 training kernels, BF16 MFMA inference). The stepper keeps the plain
guess as fallback and judge: Newton starts from whichever guess has the smaller residual norm.
"""
import ctypes as C
import os

import numpy as np

from . import spb6

N_IN, N_H = 175, 128
NPARAM = N_IN * N_H + N_H + N_H * N_H + N_H + N_H + 1
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "guess_mlp.npz")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def init_params(seed=0):
    """Glorot-uniform weights, zero biases; the generator is seeded with SplitMix64 of `seed`."""
    from .cases import SplitMix64
    rs = np.random.Generator(np.random.PCG64(SplitMix64(seed).next()))
    P = np.zeros(NPARAM, dtype=np.float32)
    o = 0

    def glorot(fi, fo):
        lim = np.sqrt(6.0 / (fi + fo))
        return rs.uniform(-lim, lim, size=(fi, fo)).astype(np.float32)
    P[o:o + N_IN * N_H] = glorot(N_IN, N_H).ravel(); o += N_IN * N_H
    o += N_H
    P[o:o + N_H * N_H] = glorot(N_H, N_H).ravel(); o += N_H * N_H
    o += N_H
    P[o:o + N_H] = glorot(N_H, 1).ravel(); o += N_H
    return P


class Network:
    """A device-resident network (spb6_mlp)."""

    def __init__(self, device=0):
        self.L = spb6.load()
        self.h = self.L.spb6_mlp_create(int(device))
        if not self.h:
            raise spb6.SPB6Error("spb6_mlp_create failed")
        assert self.L.spb6_mlp_nparams() == NPARAM
        self.norm = None

    def close(self):
        if getattr(self, "h", None):
            self.L.spb6_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, norm=None):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == NPARAM
        if norm is not None:
            norm = np.ascontiguousarray(norm, dtype=np.float32)
            assert norm.size == 2 * N_IN + 2
            self.norm = norm
        rc = self.L.spb6_mlp_set_params(self.h, _fp(params), _fp(norm) if norm is not None else None)
        if rc:
            raise spb6.SPB6Error(f"set_params: {rc}")

    def params(self):
        P = np.empty(NPARAM, dtype=np.float32)
        self.L.spb6_mlp_get_params(self.h, _fp(P))
        return P

    def set_data(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        if self.L.spb6_mlp_set_data(self.h, _fp(X), _fp(Y), X.shape[0]):
            raise spb6.SPB6Error("set_data failed")

    def train(self, epochs, batch=8192, lr=1e-3, seed=0):
        loss = np.zeros(epochs, dtype=np.float32)
        rc = self.L.spb6_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed), _fp(loss))
        if rc:
            raise spb6.SPB6Error(f"train: {rc}")
        return loss

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty(X.shape[0], dtype=np.float32)
        self.L.spb6_mlp_forward(self.h, _fp(X), X.shape[0], _fp(y))
        return y

    def gradients(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.empty(NPARAM, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self.L.spb6_mlp_gradients(self.h, _fp(X), _fp(Y), X.shape[0], _fp(g), _fp(loss))
        return g, float(loss[0])

    def infer_bf16(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty(X.shape[0], dtype=np.float32)
        self.L.spb6_mlp_infer_bf16(self.h, _fp(X), X.shape[0], _fp(y))
        return y


def dump_enable(stepper, capacity, stride):
    if stepper.L.spb6_dump_enable(stepper.h, int(capacity), int(stride)):
        raise spb6.SPB6Error("dump_enable failed")


def dump_fetch(stepper, max_rows):
    X = np.empty((max_rows, N_IN), dtype=np.float32)
    Y = np.empty(max_rows, dtype=np.float32)
    n = stepper.L.spb6_dump_fetch(stepper.h, _fp(X), _fp(Y), int(max_rows))
    if n < 0:
        raise spb6.SPB6Error("dump_fetch failed")
    return X[:n], Y[:n]


def collect(runs, steps, stride, warm=0, capacity=4_000_000):
    """Run the plain solver on each (cfg, U) of `runs` for `steps` steps with training dumps every `stride`-th cell of
    every stage (after `warm` undumped steps); returns raw features and targets."""
    Xs, Ys = [], []
    for cfg, U in runs:
        st = spb6.Stepper(cfg)
        st.set_state(U)
        if warm:
            assert st.run(1.0, warm) == 0
        dump_enable(st, capacity, stride)
        rc = st.run(1.0, steps)
        if rc != 0:
            raise spb6.SPB6Error(f"data run failed: {spb6.STATUS.get(rc, rc)}")
        X, Y = dump_fetch(st, capacity)
        Xs.append(X)
        Ys.append(Y)
        st.close()
    return np.concatenate(Xs), np.concatenate(Ys)


def standardise(X, Y, sample=1_000_000, seed=0, chunk=1_000_000):
    """Feature and target statistics, norm = [mean(175), std(175), mean_t, std_t], from at most `sample` rows; X and Y
    are standardised in place (chunk by chunk, FP32). Returns (X, Y, norm)."""
    rs = np.random.default_rng(seed)
    idx = rs.choice(X.shape[0], size=min(sample, X.shape[0]), replace=False) if X.shape[0] > sample else slice(None)
    Xs, Ys = X[idx].astype(np.float64), Y[idx].astype(np.float64)
    mu, sd = Xs.mean(axis=0), Xs.std(axis=0)
    sd = np.where(sd > 1e-12, sd, 1.0)
    mt, st = float(Ys.mean()), float(Ys.std())
    st = st if st > 0 else 1.0
    norm = np.concatenate([mu, sd, [mt, st]]).astype(np.float32)
    m32, s32 = mu.astype(np.float32), sd.astype(np.float32)
    for a in range(0, X.shape[0], chunk):
        X[a:a + chunk] -= m32
        X[a:a + chunk] /= s32
    Y -= np.float32(mt)
    Y /= np.float32(st)
    return X, Y, norm


def train_network(X, Y, epochs=20, batch=8192, lr=1e-3, seed=0, device=0, log=None, rounds=None):
    """Standardise (in place), then train a fresh network on the GPU; returns (network, norm, loss per epoch). With
    `log`, training runs in rounds of a few epochs and log(message) is called after each."""
    Xs, Ys, norm = standardise(X, Y, seed=seed)
    net = Network(device)
    net.set_params(init_params(seed), norm)
    net.set_data(Xs, Ys)
    if log is None:
        return net, norm, net.train(epochs, batch=batch, lr=lr, seed=seed)
    rounds = rounds or max(1, epochs // 4)
    losses, done = [], 0
    while done < epochs:
        k = min(rounds, epochs - done)
        losses.append(net.train(k, batch=batch, lr=lr, seed=seed + done))
        done += k
        log(f"epochs {done}/{epochs}: loss {losses[-1][-1]:.4e}")
    return net, norm, np.concatenate(losses)


def save(net, norm, path=WEIGHTS, **meta):
    np.savez(path, params=net.params(), norm=norm, **{k: np.asarray(v) for k, v in meta.items()})


def load_network(path=WEIGHTS, device=0):
    z = np.load(path, allow_pickle=False)
    net = Network(device)
    net.set_params(z["params"], z["norm"])
    return net


def attach(stepper, net):
    """Use `net` for the stepper's learned guess (the stepper's config must have learned_guess = 1)."""
    stepper._guess_net = net   # keep the network alive as long as the stepper
    if stepper.L.spb6_set_guess_model(stepper.h, net.h):
        raise spb6.SPB6Error("set_guess_model failed")


def load_into(stepper, path=WEIGHTS):
    attach(stepper, load_network(path))
