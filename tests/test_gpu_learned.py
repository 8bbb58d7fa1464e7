"""GPU kernels of the learned preconditioner against the CPU oracle
(oracle/mpnn.py): graph, the network's correction, the training objective
(the relative residual ||Z M r - r||^2 / ||r||^2 one application of
M r = S (r + d) leaves) and its gradient; the operator A = Z S the objective
uses; training on residuals of the library's own GMRES; the trained network
inside FGMRES."""
import math
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
from oracle import mpnn as NN  # noqa: E402

pytestmark = pytest.mark.gpu
K0 = 2 * math.pi / 0.1


def _graph_sample(dm, m, rng):
    r = dm.build_rwg(m)
    nbr, sgn, geom = dm.rwg_graph(r, K0)
    N = r.nedges
    cplx = lambda: rng.standard_normal(N) + 1j * rng.standard_normal(N)  # noqa: E731
    return r, dict(z=cplx(), r=cplx(), nbr=nbr, sgn=sgn, geom=geom)


def _operator(dm, m, r):
    """(A = Z S from the library, Z, NearFieldOperator S) of a mesh."""
    Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64, device=True)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    S = dm.NearFieldOperator(P, Z)
    return dm.residual_operator(S), Z, S


def _oracle_sample(s):
    nbr = s["nbr"].astype(np.int64)
    deg = (nbr >= 0).sum(axis=1)
    return dict(z=s["z"], r=s["r"], geom=s["geom"].astype(np.float64), nbr=nbr,
                sgn=s["sgn"].astype(np.float64), deg=deg, A=np.asarray(s["A"], np.complex128))


def _params(seed):
    th = NN.init_params(seed)
    p = NN.unflatten(th)
    rng = np.random.default_rng(seed + 1)
    p["W_d"] = 0.3 * rng.standard_normal(p["W_d"].shape)
    for l in range(NN.LAYERS):
        p[f"b{l}"] = 0.1 * rng.standard_normal(NN.HIDDEN)
    return NN.flatten(p).astype(np.float32)


def test_graph_and_geometry_match_oracle(dm, O):
    for m in (dm.make_rect_plate(0.1, 0.08, 5, 4), dm.make_icosphere(0.05, 2)):
        r = dm.build_rwg(m)
        nbr, sgn, geom = dm.rwg_graph(r, K0)
        nb_o, sg_o, deg_o = NN.build_graph(r.tplus - 1, r.tminus - 1, m.nt)
        assert np.array_equal(nbr, nb_o) and np.array_equal(sgn, sg_o)
        g_o = NN.geometry(r.len, r.area_plus, r.area_minus, K0)
        assert np.allclose(geom, g_o, rtol=1e-6, atol=0)


def test_forward_objective_and_gradients_match_oracle(dm):
    """FP32 kernels with bfloat16 weights vs the FP64 oracle at the same rounded
    weights and the same complex64 operators: correction, objective and every
    parameter block's gradient."""
    rng = np.random.default_rng(3)
    meshes = [dm.make_rect_plate(0.1, 0.08, 5, 4), dm.make_icosphere(0.05, 2), dm.make_rect_plate(0.06, 0.1, 3, 6)]
    samples, rwgs = [], []
    for m in meshes:
        r, _ = _graph_sample(dm, m, rng)
        A, Z, S = _operator(dm, m, r)
        for _ in range(2):
            _, smp = _graph_sample(dm, m, rng)
            smp["A"] = A
            samples.append(smp)
        rwgs.append(r)
        del S
        Z.free()
    theta = _params(5)
    net = dm.MPNN(theta)
    data = dm.MPNNData(samples)
    loss, grad = net.loss_grad(data)
    osamples = [_oracle_sample(s) for s in samples]
    loss_o, grad_o = NN.loss_and_grad(theta, osamples, bf16=True)
    print(f"objective: device {loss:.6e} oracle {loss_o:.6e}")
    assert abs(loss - loss_o) <= 1e-5 * abs(loss_o)
    assert np.linalg.norm(grad - grad_o) <= 2e-5 * np.linalg.norm(grad_o)
    off = 0
    for name, shape in NN.param_layout():
        cnt = int(np.prod(shape))
        g, go = grad[off:off + cnt], grad_o[off:off + cnt]
        assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go) + 1e-9, name
        off += cnt
    # a sub-range of samples that starts inside the first mesh's pair
    l1, g1 = net.loss_grad(data, 1, 4)
    lo1, go1 = NN.loss_and_grad(theta, osamples[1:4], bf16=True)
    assert abs(l1 - lo1) <= 1e-5 * lo1 and np.linalg.norm(g1 - go1) <= 2e-5 * np.linalg.norm(go1)
    # the correction on one mesh
    s = samples[2]
    out = net.apply(rwgs[1], K0, s["z"], s["r"])
    o = osamples[2]
    out_o = NN.apply(theta, o["z"], o["r"], o["geom"], o["nbr"], o["sgn"], o["deg"])
    assert np.linalg.norm(out - out_o) <= 1e-5 * np.linalg.norm(out_o)
    # untrained network (W_d = 0): the objective is the SPAI's own relative residual
    l0, _ = dm.MPNN(seed=3).loss_grad(data)
    ref = np.mean([np.linalg.norm(o["r"] - o["A"] @ o["r"]) ** 2 / np.linalg.norm(o["r"]) ** 2 for o in osamples])
    assert abs(l0 - ref) <= 1e-5 * ref
    for s in samples:
        s.pop("A")
    with pytest.raises(ValueError):
        dm.MPNNData(samples)


def test_residual_operator_is_z_times_s(dm):
    """A = Z S from the library equals the product of the library's complex128
    matrix and the SPAI applied to unit vectors (caller order), to complex64
    rounding; A r = Z (S r)."""
    m = dm.make_icosphere(0.05, 2)
    r = dm.build_rwg(m)
    A, Z, S = _operator(dm, m, r)
    N = r.nedges
    Sd = np.stack([S @ np.eye(N)[:, j] for j in range(N)], axis=1)
    ref = Z.to_numpy(dm.DM_Z64) @ Sd
    assert A.dtype == np.complex64 and A.shape == (N, N)
    assert np.linalg.norm(A - ref) <= 2e-7 * np.linalg.norm(ref)
    x = np.random.default_rng(4).standard_normal(N) + 0j
    assert np.linalg.norm(A @ x - Z.matmul(S @ x)) <= 1e-6 * np.linalg.norm(A @ x)
    Z.free()


def test_training_lowers_loss_and_runs_in_fgmres(dm):
    meshes = dm.make_training_meshes(10, seed=11, max_sphere_subdiv=2)
    samples = dm.generate_training_samples(meshes, K0, iters=(2, 8, 24), seed=12)
    assert len(samples) >= 20
    data = dm.MPNNData(samples)
    net = dm.MPNN(seed=1)
    l0, _ = net.loss_grad(data)
    hist = net.train(data, epochs=40, batch_samples=8, lr=2e-3, seed=2)
    l1, _ = net.loss_grad(data)
    print(f"training objective {l0:.4f} -> {l1:.4f} (epoch means {hist[0]:.4f} .. {hist[-1]:.4f})")
    assert np.all(np.isfinite(hist)) and l1 < l0
    # an unseen mesh: FGMRES with the learned preconditioner reaches the true residual bar
    m = dm.make_icosphere(0.08, 2)
    r = dm.build_rwg(m)
    Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P, device=True)
    v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
    P = dm.LearnedPreconditioner(net, 2.0 * float(np.mean(r.len)))
    x, st = dm.solve_gmres(Z, v, preconditioner=P, precond_side="right", tol=1e-8, maxiter=300)
    Zh = Z.to_numpy(dm.DM_Z64P)
    assert st.solved and np.linalg.norm(Zh @ x - v) / np.linalg.norm(v) < 1.05e-8
    xm, sm = dm.solve_mixed(Z, v, preconditioner=P, tol=1e-6)
    assert sm.solved and np.linalg.norm(Zh @ xm - v) / np.linalg.norm(v) <= 1e-6
    with pytest.raises(ValueError):
        dm.solve_gmres(Z, v, preconditioner=P, precond_side="left")


WEIGHTS = os.path.join(ROOT, "diffmom.jl_amd", "weights", "learned_precond.npz")


def test_learned_preconditioner_ablation(dm):
    """The committed network (train_learned.py: 2000 seeded plates and spheres,
    Arnoldi vectors of the SPAI-preconditioned Krylov spaces, objective
    ||Z M r - r||^2 / ||r||^2) against the SPAI alone, on held-out meshes of a
    seed no training run draws from and on the 30720-basis sphere:
      * what it gains: the training objective on held-out vectors (the
        ungated network must lower it by >= 10 %);
      * FGMRES iterations (tol 1e-6) with the SPAI alone, the ungated network
        and the gated one that the solver uses;
      * on the sphere, iterations and time of the mixed solve for the three.
    The gate was chosen on FGMRES iterations of validation meshes; the gated
    network must never need more iterations than the SPAI alone here, and
    every solve must reach the true FP64 residual bar."""
    import json
    import time
    with np.load(WEIGHTS, allow_pickle=False) as f:
        gated, raw = f["params"], f["params_raw"]
    with open(os.path.splitext(WEIGHTS)[0] + ".json") as f:
        summary = json.load(f)
    assert summary["meshes"] >= 2000
    gamma = summary["gamma"]
    net_raw, net = dm.MPNN(raw), dm.MPNN(gated)
    meshes = dm.make_training_meshes(8, seed=777001, max_sphere_subdiv=3)
    samples = dm.generate_training_samples(meshes, K0, iters=(1, 4, 16, 32), seed=777002, inputs="arnoldi")
    data = dm.MPNNData(samples)
    l_spai, _ = dm.MPNN(seed=0).loss_grad(data)
    l_raw, _ = net_raw.loss_grad(data)
    print(f"[ablation] held-out objective ||Z M v - v||^2/||v||^2: SPAI alone {l_spai:.4f}, "
          f"network {l_raw:.4f} ({100 * (1 - l_raw / l_spai):.1f} % lower); gate gamma = {gamma:g}", flush=True)
    assert l_raw <= 0.9 * l_spai
    its = {"spai": 0, "network": 0, "gated": 0}
    for m in meshes + [dm.make_icosphere(0.1, 4)]:
        r = dm.build_rwg(m)
        Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P, device=True, max_output_bytes=None)
        v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
        cut = 2.0 * float(np.mean(r.len))
        Zh = Z.to_numpy(dm.DM_Z64P)
        row = []
        for key, P in (("spai", dm.build_nearfield_preconditioner(None, m, r, cut)),
                       ("network", dm.LearnedPreconditioner(net_raw, cut)),
                       ("gated", dm.LearnedPreconditioner(net, cut))):
            x, st = dm.solve_gmres(Z, v, preconditioner=P, precond_side="right", tol=1e-6, maxiter=1000)
            assert st.solved and np.linalg.norm(Zh @ x - v) / np.linalg.norm(v) <= 1.05e-6
            its[key] += st.niter
            row.append(f"{key} {st.niter}")
        print(f"[ablation] N={r.nedges}: " + ", ".join(row), flush=True)
        Z.free()
    print(f"[ablation] FGMRES iterations over 9 held-out meshes: {its}", flush=True)
    assert its["gated"] <= its["spai"]
    # the 30720-basis sphere, mixed solve
    m = dm.make_icosphere(0.1, 5)
    r = dm.build_rwg(m)
    cut = 2.0 * float(np.mean(r.len))
    pw = [dm.make_plane_wave([0, 0, -K0], 1.0, [1, 0, 0])]
    res = {}
    for key, P in (("spai", dm.build_nearfield_preconditioner(None, m, r, cut)),
                   ("network", dm.LearnedPreconditioner(net_raw, cut)),
                   ("gated", dm.LearnedPreconditioner(net, cut))):
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            X, st, _ = dm.solve_planewaves(m, r, K0, pw, "mixed", P, tol=1e-6, maxiter=1000, batch=-1,
                                           compress_tol=5e-8)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        assert st.solved and st.residual <= 1e-6
        res[key] = (st.niter, best)
    print("[ablation] sphere N=30720 mixed solve: " +
          ", ".join(f"{k} {n} it {1e3 * t:.1f} ms" for k, (n, t) in res.items()), flush=True)
    assert res["gated"][0] <= res["spai"][0]
