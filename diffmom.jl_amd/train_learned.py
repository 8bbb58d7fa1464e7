#!/usr/bin/env python3
"""Train the learned FGMRES preconditioner on the GPU.

    python diffmom.jl_amd/train_learned.py [--meshes 2000] [--epochs 30] [--out PATH]

Training set: seeded random plates (sides 0.05-0.3 m, 4-30 cells per side) and
icospheres (2-4 subdivisions), vertex jitter 0.1 mean edge length, wavelength
0.1 m.  On each mesh the library's own right-preconditioned GMRES (near-field
SPAI over 2 mean edge lengths) solves a random plane wave to 1e-10; the
iterates after 1, 2, 4, 8, 16, 32 steps give the (residual, error) pairs.  The
network is trained by the library's HIP forward / backward kernels with Adam
and saved as an .npz of its FP32 parameters, with a JSON summary beside it.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "weights", "learned_precond.npz")


def load_pkg():
    name = "diffmom_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[train] {msg}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32, help="samples per Adam step")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--max-sphere-subdiv", type=int, default=4)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    dm = load_pkg()
    k = 2 * math.pi / 0.1
    t0 = time.perf_counter()
    samples = []
    chunk = 50
    nodes = 0
    for c0 in range(0, a.meshes, chunk):
        cnt = min(chunk, a.meshes - c0)
        meshes = dm.make_training_meshes(cnt, seed=a.seed + c0, max_sphere_subdiv=a.max_sphere_subdiv)
        part = dm.generate_training_samples(meshes, k, seed=a.seed + 7919 + c0)
        samples += part
        nodes += sum(len(s["z"]) for s in part)
        log(f"meshes {c0 + cnt}/{a.meshes}: {len(samples)} samples, {nodes} nodes, "
            f"{time.perf_counter() - t0:.1f} s")
    t_gen = time.perf_counter() - t0
    # hold out every 10th sample for validation
    val = samples[::10]
    train = [s for i, s in enumerate(samples) if i % 10]
    rng = np.random.default_rng(a.seed)
    order = rng.permutation(len(train))
    train = [train[i] for i in order]
    dtrain = dm.MPNNData(train)
    dval = dm.MPNNData(val)
    net = dm.MPNN(seed=a.seed)
    v0, _ = net.loss_grad(dval)
    log(f"data: {len(train)} training / {len(val)} validation samples; initial validation loss {v0:.4f}")
    t1 = time.perf_counter()
    hist, vhist = [], []
    for ep in range(a.epochs):
        lr = a.lr * 0.5 * (1 + math.cos(math.pi * ep / a.epochs))  # cosine schedule
        h = net.train(dtrain, epochs=1, batch_samples=a.batch, lr=max(lr, 1e-5), seed=a.seed + ep)
        v, _ = net.loss_grad(dval)
        hist.append(float(h[0]))
        vhist.append(float(v))
        log(f"epoch {ep + 1}/{a.epochs}: train {h[0]:.4f}  validation {v:.4f}  lr {lr:.2e}")
    t_train = time.perf_counter() - t1
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    summary = dict(meshes=a.meshes, samples_train=len(train), samples_validation=len(val),
                   nodes_train=int(dtrain.nnode), epochs=a.epochs, batch_samples=a.batch, lr=a.lr, seed=a.seed,
                   initial_validation_loss=float(v0), train_loss=hist, validation_loss=vhist,
                   data_generation_s=round(t_gen, 1), training_s=round(t_train, 1))
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(f"saved {a.out}: validation loss {v0:.4f} -> {vhist[-1]:.4f}; generation {t_gen:.0f} s, training {t_train:.0f} s")


if __name__ == "__main__":
    main()
