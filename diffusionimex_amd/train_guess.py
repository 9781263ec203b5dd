"""Train the learned Newton guess on solver dumps and save the network.

The plain FP64 solver runs seeded workloads with training dumps switched on (every `stride`-th active cell of every
stage: features of the stage state at the plain guess, target (P_converged - P_plain) / K_w); the network is then trained
on the GPU by the library's FP32 forward/backward/Adam kernels and written as an .npz (params, norm, metadata).

    python -m diffusionimex_amd.train_guess --out diffusionimex_amd/weights/guess_mlp.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def workloads(a):
    from . import cases
    out = []
    for s in a.bubble_seeds:
        out.append((f"shock_bubble {a.nx}x{a.ny} seed {s}", cases.shock_bubble(a.nx, a.ny, seed=s, cfl=0.5)))
    for s in a.interface_seeds:
        out.append((f"perturbed_interface {a.nx}x{a.ny} seed {s}",
                    cases.perturbed_interface(a.nx, a.ny, seed=s, cfl=0.5)))
    return out


def main(argv=None):
    from . import mlp
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--out", default=mlp.WEIGHTS)
    p.add_argument("--nx", type=int, default=2048)
    p.add_argument("--ny", type=int, default=1024)
    p.add_argument("--bubble-seeds", type=lambda s: [int(x) for x in s.split(",") if x], default=[1, 2, 3, 4, 5, 6, 7, 8])
    p.add_argument("--interface-seeds", type=lambda s: [int(x) for x in s.split(",") if x], default=[1, 2, 3, 4])
    p.add_argument("--warm", type=int, default=12, help="undumped steps first (start-up ramp)")
    p.add_argument("--steps", type=int, default=30)
    p.add_argument("--stride", type=int, default=1)
    p.add_argument("--capacity", type=int, default=6_000_000, help="rows per run")
    p.add_argument("--epochs", type=int, default=16)
    p.add_argument("--batch", type=int, default=8192)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--seed", type=int, default=0)
    a = p.parse_args(argv)

    t0 = time.perf_counter()
    Xs, Ys, names = [], [], []
    for name, run in workloads(a):
        X, Y = mlp.collect([run], steps=a.steps, stride=a.stride, warm=a.warm, capacity=a.capacity)
        print(f"dumped {X.shape[0]} rows from {name} ({time.perf_counter() - t0:.1f} s)", flush=True)
        Xs.append(X); Ys.append(Y); names.append(name)
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    del Xs, Ys
    print(f"{X.shape[0]} rows ({X.nbytes / 2**30:.1f} GiB of features)", flush=True)
    t1 = time.perf_counter()
    net, norm, loss = mlp.train_network(X, Y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                                        log=lambda m: print(f"{m} ({time.perf_counter() - t1:.1f} s)", flush=True), rounds=2)
    t2 = time.perf_counter()
    for e in range(0, a.epochs, max(1, a.epochs // 10)):
        print(f"epoch {e}: loss {loss[e]:.4e}", flush=True)
    print(f"epoch {a.epochs - 1}: loss {loss[-1]:.4e}; trained on {X.shape[0]} rows in {t2 - t1:.1f} s", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    meta = {"rows": int(X.shape[0]), "workloads": names, "warm": a.warm, "steps": a.steps, "stride": a.stride,
            "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "seed": a.seed, "loss": [float(v) for v in loss],
            "train_seconds": t2 - t1}
    mlp.save(net, norm, a.out, meta=json.dumps(meta))
    print(f"saved {a.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
