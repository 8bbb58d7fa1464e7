"""Pins the CPU oracle of the learned preconditioner's network: the graph
against brute-force triangle sharing, the backward pass against central
finite differences of the forward pass and objective, the objective against
its closed forms and against ||Z M r - r|| recomputed from M r, the
equivariances against their closed forms, and Adam against its textbook
recurrence."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
from oracle import mpnn as NN  # noqa: E402
from oracle import oracle as O  # noqa: E402

K0 = 2 * np.pi / 0.1


def _mesh_graph(nx=4, ny=3):
    m = O.make_rect_plate(0.1, 0.08, nx, ny)
    r = O.build_rwg(m)
    nbr, sgn, deg = NN.build_graph(r.tplus - 1, r.tminus - 1, m.nt)
    return m, r, nbr, sgn, deg


def test_graph_matches_triangle_sharing():
    m, r, nbr, sgn, deg = _mesh_graph()
    N = len(r.tplus)
    for i in range(N):
        tri_i = {int(r.tplus[i]) - 1: 1.0, int(r.tminus[i]) - 1: -1.0}
        expect = []
        for j in range(N):
            if j == i:
                continue
            tri_j = {int(r.tplus[j]) - 1: 1.0, int(r.tminus[j]) - 1: -1.0}
            shared = set(tri_i) & set(tri_j)
            assert len(shared) <= 1
            for t in shared:
                expect.append((j, tri_i[t] * tri_j[t]))
        got = [(int(nbr[i, q]), float(sgn[i, q])) for q in range(deg[i])]
        assert got == sorted(expect)
        assert np.all(nbr[i, deg[i]:] == -1)
    # symmetric with equal signs
    for i in range(N):
        for q in range(deg[i]):
            j = nbr[i, q]
            qq = list(nbr[j, :deg[j]]).index(i)
            assert sgn[j, qq] == sgn[i, q]


def _sample(seed, nx=4, ny=3):
    m, r, nbr, sgn, deg = _mesh_graph(nx, ny)
    rng = np.random.default_rng(seed)
    N = len(r.tplus)
    cplx = lambda: rng.standard_normal(N) + 1j * rng.standard_normal(N)  # noqa: E731
    geom = NN.geometry(r.len, r.area_plus, r.area_minus, K0)
    A = np.eye(N) + 0.3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / np.sqrt(N)
    return dict(z=cplx(), r=cplx(), geom=geom, nbr=nbr, sgn=sgn, deg=deg, A=A)


def _theta(seed):
    th = NN.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    p = NN.unflatten(th)
    p["W_d"] = 0.3 * rng.standard_normal(p["W_d"].shape)
    for l in range(NN.LAYERS):
        p[f"b{l}"] = 0.1 * rng.standard_normal(NN.HIDDEN)
    return NN.flatten(p)


def test_backward_matches_finite_differences():
    samples = [_sample(1), _sample(2, 3, 3)]
    th = _theta(3)
    f = lambda t: NN.loss_and_grad(t, samples, bf16=False)[0]  # noqa: E731
    _, g = NN.loss_and_grad(th, samples, bf16=False)
    rng = np.random.default_rng(4)
    sizes = [int(np.prod(s)) for _, s in NN.param_layout()]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    # a few coordinates of every parameter block, plus two random directions
    idx = np.concatenate([rng.integers(offs[b], offs[b + 1], 3) for b in range(len(sizes))])
    h = 1e-6
    for i in idx:
        d = np.zeros_like(th)
        d[i] = h
        fd = (f(th + d) - f(th - d)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])
    for _ in range(2):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        fd = (f(th + h * d) - f(th - h * d)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd))


def test_planted_backward_mistakes_are_caught():
    """A dropped neighbour term or a sign flip in the aggregation adjoint must
    break the finite-difference agreement (so the check above has teeth)."""
    s = _sample(5)
    th = _theta(6)
    _, g = NN.loss_and_grad(th, [s], bf16=False)
    orig = NN._scatter_back
    try:
        NN._scatter_back = lambda gm, nbr, sgn, deg: -orig(gm, nbr, sgn, deg)
        _, g_bad = NN.loss_and_grad(th, [s], bf16=False)
    finally:
        NN._scatter_back = orig
    off = NN.HIDDEN * NN.F_IN
    assert np.linalg.norm(g_bad[:off] - g[:off]) > 1e-3 * np.linalg.norm(g[:off])


def test_apply_equivariances():
    s = _sample(7)
    th = _theta(8)
    base = NN.apply(th, s["z"], s["r"], s["geom"], s["nbr"], s["sgn"], s["deg"])
    alpha = 3.7 * np.exp(0.9j)
    scaled = NN.apply(th, alpha * s["z"], alpha * s["r"], s["geom"], s["nbr"], s["sgn"], s["deg"])
    assert np.linalg.norm(scaled - alpha * base) <= 1e-12 * np.linalg.norm(alpha * base)
    # W_d = 0: the network adds nothing
    p = NN.unflatten(th)
    p["W_d"][:] = 0
    out = NN.apply(NN.flatten(p), s["z"], s["r"], s["geom"], s["nbr"], s["sgn"], s["deg"])
    assert not np.any(out)
    # renumbering the bases permutes the output
    N = len(s["z"])
    perm = np.random.default_rng(9).permutation(N)
    inv = np.argsort(perm)
    nbr_p = np.where(s["nbr"][perm] >= 0, inv[np.maximum(s["nbr"][perm], 0)], -1)
    order = np.argsort(np.where(nbr_p >= 0, nbr_p, 1 << 40), axis=1)
    nbr_p = np.take_along_axis(nbr_p, order, 1)
    sgn_p = np.take_along_axis(s["sgn"][perm], order, 1)
    outp = NN.apply(th, s["z"][perm], s["r"][perm], s["geom"][perm], nbr_p, sgn_p, s["deg"][perm])
    assert np.allclose(outp, base[perm], rtol=1e-12, atol=1e-12)


def test_bf16_rounding_and_adam():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415927, 1e-30, 65504.0], np.float32)
    b = NN.bf16_round(x)
    # bfloat16 keeps 8 significant bits; ties go to even
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == 1.0078125
    assert abs(b[3] + 3.140625) < 1e-7
    assert np.all(np.abs(b - x) <= np.abs(x) * 2.0**-8)
    th = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    t1, m, v = NN.adam_step(th, g, np.zeros(2), np.zeros(2), 1, 0.1)
    # first Adam step moves each coordinate by lr * sign(g) (bias-corrected m/sqrt(v) = g/|g|)
    assert np.allclose(t1, th - 0.1 * np.sign(g), atol=1e-7)


def _efie_sample(seed, nx=4, ny=3):
    """A plate's EFIE matrix Z (the oracle's), a dense stand-in S for the
    near-field inverse (Z^-1 perturbed), A = Z S and a random residual."""
    m = O.make_rect_plate(0.1, 0.08, nx, ny)
    r = O.build_rwg(m)
    nbr, sgn, deg = NN.build_graph(r.tplus - 1, r.tminus - 1, m.nt)
    Z = O.assemble_Z_efie(r, K0, quad_order=3)
    rng = np.random.default_rng(seed)
    N = len(r.tplus)
    Zi = np.linalg.inv(Z)
    S = Zi + 0.2 * np.linalg.norm(Zi, 2) * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) / N
    res = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    geom = NN.geometry(r.len, r.area_plus, r.area_minus, K0)
    return dict(z=S @ res, r=res, geom=geom, nbr=nbr, sgn=sgn, deg=deg, A=Z @ S, Z=Z, S=S)


def test_objective_matches_finite_differences_on_efie_operator():
    samples = [_efie_sample(11), _efie_sample(12, 3, 3)]
    th = _theta(13)
    f = lambda t: NN.loss_and_grad(t, samples, bf16=False)[0]  # noqa: E731
    _, g = NN.loss_and_grad(th, samples, bf16=False)
    rng = np.random.default_rng(14)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        fd = (f(th + h * d) - f(th - h * d)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd)), (fd, g @ d)
    # the decoder block coordinate by coordinate (where the objective enters)
    off = NN.n_params() - 2 * NN.HIDDEN
    for i in rng.integers(off, NN.n_params(), 6):
        e = np.zeros_like(th)
        e[i] = h
        fd = (f(th + e) - f(th - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_objective_is_the_residual_one_application_leaves():
    """The objective equals ||Z M r - r||^2 / ||r||^2 with M r recomputed from
    its definition S r + S d; at W_d = 0 it is the SPAI's own relative
    residual; the exact correction d = A^-1 r - r gives 0.  A planted mistake
    (p = r - Z r instead of r - A r) changes the value."""
    s = _efie_sample(15)
    th = _theta(16)
    loss, _ = NN.loss_and_grad(th, [s])  # both with the bfloat16-rounded weights
    Mr = NN.precondition(th, s["S"], s["r"], s["geom"], s["nbr"], s["sgn"], s["deg"])
    ref = np.linalg.norm(s["Z"] @ Mr - s["r"]) ** 2 / np.linalg.norm(s["r"]) ** 2
    assert abs(loss - ref) <= 1e-9 * ref, (loss, ref)
    p = NN.unflatten(th)
    p["W_d"][:] = 0
    loss0, _ = NN.loss_and_grad(NN.flatten(p), [s], bf16=False)
    ref0 = np.linalg.norm(s["r"] - s["Z"] @ s["z"]) ** 2 / np.linalg.norm(s["r"]) ** 2
    assert abs(loss0 - ref0) <= 1e-12 * ref0
    _, c = NN.features(s["z"], s["r"], s["geom"])
    d = np.linalg.solve(s["A"], s["r"]) - s["r"]
    y = c * d
    l_exact, g_exact = NN.sample_residual_loss(np.column_stack([y.real, y.imag]), s["r"], c, s["A"])
    assert l_exact <= 1e-20 and np.linalg.norm(g_exact) <= 1e-8
    bad = dict(s, A=s["Z"])
    l_bad, _ = NN.loss_and_grad(NN.flatten(p), [bad], bf16=False)
    assert abs(l_bad - ref0) > 1e-3 * ref0
