#!/usr/bin/env python3
"""Train the learned FGMRES preconditioner on the GPU.

    python diffmom.jl_amd/train_learned.py [--meshes 2000] [--epochs 20] [--inputs arnoldi] [--out PATH]

Training set: seeded random plates (sides 0.05-0.3 m, 4-30 cells per side) and
icospheres (2-3 subdivisions), vertex jitter 0.1 mean edge length, wavelength
0.1 m.  On each mesh a random plane wave b and the near-field SPAI S over 2
mean edge lengths give the sampled vectors: the Arnoldi basis vectors of the
S-preconditioned Krylov space of b (--inputs arnoldi, the vectors FGMRES hands
its preconditioner), or the residuals of the library's GMRES iterates
(--inputs residual).  The network is trained by the library's HIP forward /
backward kernels with Adam on the relative residual that one application of
the preconditioner M r = S (r + d(S r, r)) leaves, ||Z M r - r||^2 / ||r||^2,
with each mesh's operator A = Z S kept on the device (complex64).  The decoder
starts at zero, so the untrained network is the SPAI alone; the parameters of
the epoch with the lowest objective on the held-out meshes are kept.

Gate: the correction is then scaled by the gamma in {0, 1/4, 1/2, 1} that
needs the fewest FGMRES iterations (tol 1e-6, one plane wave) summed over the
held-out meshes plus one sphere larger than any training mesh; a gamma > 0 is
taken only if it saves at least 2 % of the SPAI's iterations.  gamma = 0 is
the SPAI alone, which the solver then applies without running the network.  Saved as an .npz (the
gated parameters under "params", the ungated ones under "params_raw") with a
JSON summary beside it.
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
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=32, help="samples per Adam step")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--max-sphere-subdiv", type=int, default=3)
    ap.add_argument("--iters", default="1,2,3,5,8,12,16,24,32,48",
                    help="GMRES steps / Arnoldi indices of the sampled vectors")
    ap.add_argument("--gate-meshes", type=int, default=16, help="held-out meshes of the gate's FGMRES runs")
    ap.add_argument("--inputs", choices=["residual", "arnoldi"], default="arnoldi",
                    help="sampled vectors: GMRES residuals or Arnoldi basis vectors")
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    dm = load_pkg()
    k = 2 * math.pi / 0.1
    t0 = time.perf_counter()
    samples = []
    chunk = 50
    nodes = 0
    val_meshes = []  # every 10th mesh is held out
    for c0 in range(0, a.meshes, chunk):
        cnt = min(chunk, a.meshes - c0)
        meshes = dm.make_training_meshes(cnt, seed=a.seed + c0, max_sphere_subdiv=a.max_sphere_subdiv)
        part = dm.generate_training_samples(meshes, k, iters=tuple(int(v) for v in a.iters.split(",")),
                                            seed=a.seed + 7919 + c0, inputs=a.inputs)
        for i, m in enumerate(meshes):
            if (c0 + i) % 10 == 0 and len(val_meshes) < a.gate_meshes:
                val_meshes.append(m)
        samples += part
        nodes += sum(len(s["z"]) for s in part)
        log(f"meshes {c0 + cnt}/{a.meshes}: {len(samples)} samples, {nodes} nodes, "
            f"{time.perf_counter() - t0:.1f} s")
    t_gen = time.perf_counter() - t0
    # hold out every 10th mesh for validation (all of its samples), keeping
    # the samples of one mesh together (they share its operator)
    groups, cur = [], []
    for s in samples:
        if cur and s["nbr"] is not cur[-1]["nbr"]:
            groups.append(cur)
            cur = []
        cur.append(s)
    groups.append(cur)
    val = [s for i, g in enumerate(groups) if i % 10 == 0 for s in g]
    rng = np.random.default_rng(a.seed)
    tg = [g for i, g in enumerate(groups) if i % 10]
    train = [s for i in rng.permutation(len(tg)) for s in tg[i]]
    dtrain = dm.MPNNData(train)
    dval = dm.MPNNData(val)
    for s in samples:  # the data objects hold their own copies of the operators
        s.pop("A")
    net = dm.MPNN(seed=a.seed)
    v0, _ = net.loss_grad(dval)
    best_v, best_p, best_ep = v0, net.params(), 0
    log(f"data: {len(train)} training / {len(val)} validation samples; "
        f"validation objective of the SPAI alone {v0:.4f}")
    t1 = time.perf_counter()
    hist, vhist = [], []
    for ep in range(a.epochs):
        lr = a.lr * 0.5 * (1 + math.cos(math.pi * ep / a.epochs))  # cosine schedule
        h = net.train(dtrain, epochs=1, batch_samples=a.batch, lr=max(lr, 1e-5), seed=a.seed + ep)
        v, _ = net.loss_grad(dval)
        hist.append(float(h[0]))
        vhist.append(float(v))
        if v < best_v:
            best_v, best_p, best_ep = v, net.params(), ep + 1
        log(f"epoch {ep + 1}/{a.epochs}: train {h[0]:.4f}  validation {v:.4f}  lr {lr:.2e}")
    t_train = time.perf_counter() - t1
    # gate by FGMRES iterations on held-out meshes
    gammas = (0.0, 0.25, 0.5, 1.0)
    its = {g: 0 for g in gammas}
    nets = {}
    for g in gammas[1:]:
        pg = best_p.copy()
        pg[-2 * 64:] *= g
        nets[g] = dm.MPNN(pg)
    grng = np.random.default_rng(a.seed + 1)
    big = dm.make_icosphere(float(grng.uniform(0.8, 1.2)) * 0.1, a.max_sphere_subdiv + 1)
    for m in val_meshes + [big]:
        rwg = dm.build_rwg(m)
        Z = dm.assemble_Z_efie(m, rwg, k, forms=dm.DM_Z64P | dm.DM_Z32P, device=True, mesh_precheck=False,
                               max_output_bytes=None)
        v = dm.assemble_v_plane_wave(m, rwg, [0, 0, -k], 1.0, [1, 0, 0])
        cut = 2.0 * float(np.mean(rwg.len))
        for g in gammas:
            P = dm.build_nearfield_preconditioner(None, m, rwg, cut) if g == 0 else \
                dm.LearnedPreconditioner(nets[g], cut)
            _, st = dm.solve_gmres(Z, v, preconditioner=P, precond_side="right", tol=1e-6, maxiter=1000)
            its[g] += st.niter
        Z.free()
    best_g = min(gammas[1:], key=lambda g: (its[g], g))
    gamma = best_g if its[best_g] <= 0.98 * its[0.0] else 0.0
    log("gate: FGMRES iterations on %d held-out meshes + 1 larger sphere %s -> gamma %g"
        % (len(val_meshes), ", ".join(f"{g:g}: {its[g]}" for g in gammas), gamma))
    gated = best_p.copy()
    gated[-2 * 64:] *= gamma
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.savez(a.out, params=gated.astype(np.float32), params_raw=best_p.astype(np.float32))
    summary = dict(meshes=a.meshes, objective="||Z M r - r||^2 / ||r||^2", max_sphere_subdiv=a.max_sphere_subdiv,
                   samples_train=len(train), samples_validation=len(val),
                   nodes_train=int(dtrain.nnode), epochs=a.epochs, batch_samples=a.batch, lr=a.lr, seed=a.seed,
                   inputs=a.inputs, iters=a.iters,
                   initial_validation_loss=float(v0), best_validation_loss=float(best_v), best_epoch=best_ep,
                   gate_meshes=len(val_meshes), gate_iterations={str(g): its[g] for g in gammas}, gamma=gamma,
                   train_loss=hist, validation_loss=vhist,
                   data_generation_s=round(t_gen, 1), training_s=round(t_train, 1))
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(f"saved {a.out} (epoch {best_ep}, gamma {gamma:g}): validation objective {v0:.4f} -> {best_v:.4f}; "
        f"generation {t_gen:.0f} s, training {t_train:.0f} s")


if __name__ == "__main__":
    main()
