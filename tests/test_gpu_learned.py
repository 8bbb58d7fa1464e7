"""GPU kernels of the learned preconditioner against the CPU oracle
(oracle/mpnn.py): graph, forward output, loss and gradients of the
message-passing network; training on (residual, error) pairs from the
library's own GMRES; the trained network inside FGMRES."""
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
    return r, dict(z=cplx(), r=cplx(), e=cplx(), nbr=nbr, sgn=sgn, geom=geom)


def _oracle_sample(s):
    nbr = s["nbr"].astype(np.int64)
    deg = (nbr >= 0).sum(axis=1)
    return dict(z=s["z"], r=s["r"], e=s["e"], geom=s["geom"].astype(np.float64), nbr=nbr,
                sgn=s["sgn"].astype(np.float64), deg=deg)


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


def test_forward_loss_and_gradients_match_oracle(dm):
    """FP32 kernels with bfloat16 weights vs the FP64 oracle at the same rounded
    weights: output, loss and every parameter block's gradient."""
    rng = np.random.default_rng(3)
    meshes = [dm.make_rect_plate(0.1, 0.08, 5, 4), dm.make_icosphere(0.05, 2), dm.make_rect_plate(0.06, 0.1, 3, 6)]
    pairs = [_graph_sample(dm, m, rng) for m in meshes]
    theta = _params(5)
    net = dm.MPNN(theta)
    data = dm.MPNNData([s for _, s in pairs])
    loss, grad = net.loss_grad(data)
    osamples = [_oracle_sample(s) for _, s in pairs]
    loss_o, grad_o = NN.loss_and_grad(theta, osamples, bf16=True)
    assert abs(loss - loss_o) <= 1e-5 * abs(loss_o)
    assert np.linalg.norm(grad - grad_o) <= 2e-5 * np.linalg.norm(grad_o)
    off = 0
    for name, shape in NN.param_layout():
        cnt = int(np.prod(shape))
        g, go = grad[off:off + cnt], grad_o[off:off + cnt]
        assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go) + 1e-9, name
        off += cnt
    # sub-range of samples = the oracle over the same samples
    l1, g1 = net.loss_grad(data, 1, 3)
    lo1, go1 = NN.loss_and_grad(theta, osamples[1:3], bf16=True)
    assert abs(l1 - lo1) <= 1e-5 * lo1 and np.linalg.norm(g1 - go1) <= 2e-5 * np.linalg.norm(go1)
    # the preconditioner output on one mesh
    r, s = pairs[1]
    out = net.apply(r, K0, s["z"], s["r"])
    o = osamples[1]
    out_o = NN.apply(theta, o["z"], o["r"], o["geom"], o["nbr"], o["sgn"], o["deg"])
    assert np.linalg.norm(out - out_o) <= 1e-5 * np.linalg.norm(out_o)


def test_training_lowers_loss_and_runs_in_fgmres(dm):
    meshes = dm.make_training_meshes(10, seed=11, max_sphere_subdiv=2)
    samples = dm.generate_training_samples(meshes, K0, iters=(2, 8, 24), seed=12)
    assert len(samples) >= 20
    data = dm.MPNNData(samples)
    net = dm.MPNN(seed=1)
    l0, _ = net.loss_grad(data)
    hist = net.train(data, epochs=40, batch_samples=8, lr=2e-3, seed=2)
    l1, _ = net.loss_grad(data)
    print(f"training loss {l0:.4f} -> {l1:.4f} (epoch means {hist[0]:.4f} .. {hist[-1]:.4f})")
    assert np.all(np.isfinite(hist)) and l1 < 0.9 * l0
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
