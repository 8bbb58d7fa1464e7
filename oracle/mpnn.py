"""TEST INFRASTRUCTURE ONLY — CPU (numpy, FP64 arithmetic) restatement of the
learned preconditioner's message-passing network: graph construction, input
normalisation, forward pass, loss and the backward pass.  Only tests/ may
import it.  The library's HIP kernels (diffmom.jl_amd/csrc/k_mpnn.hip) must
reproduce it.

The reference has no learned preconditioner; the network is this project's
own design, so there is no reference line to follow.  Pinning: the backward
pass is checked against central finite differences of the forward pass
(tests/test_mpnn_oracle.py), the graph against brute-force triangle sharing,
and the equivariances (global complex scale of the residual, basis renumbering)
against their closed forms.

Network (per mesh, nodes = RWG bases, edges = bases sharing a triangle), for a
vector r (a residual or a Krylov basis vector) and z = S r (near-field sparse
approximate inverse):
  rho = sqrt(mean |r|^2), q = sum r |r|, c = conj(q) / (|q| rho)
  tau = sqrt(mean |c z|^2)
  u_i = [Re c z_i / tau, Im c z_i / tau, Re c r_i, Im c r_i, g1_i, g2_i]
  h^0 = W_e u
  l = 0..3:  m_i = sum_j s_ij h_j / max(deg_i, 1);  a = W_s h + W_n m + b;  h <- h + tanh(a)
  y = W_d h^4  (2 channels: Re, Im)
  d = (y_0 + i y_1) / c          (the correction, in the units of r)
  M r = z + S d = S (r + d)
Geometry: g1 = k len_i / (2 pi / 10), g2 = (A+_i + A-_i) / len_i^2.  Edge sign
s_ij = sigma_i(T) sigma_j(T) on the shared triangle T (sigma = +1 on T+, -1 on T-).
Weights are used rounded to bfloat16 (round to nearest even); activations and
accumulation are FP32 on the device (FP64 here).
Objective of one sample: the relative residual one application of M leaves,
||Z M r - r||^2 / ||r||^2 = ||A y - c p||^2 / (|c|^2 ||r||^2) with A = Z S and
p = r - A r.  Pinned by finite differences like the rest, by its closed form
at W_d = 0 (the SPAI's own relative residual ||r - A r||^2 / ||r||^2), by the
exact correction d = A^-1 r - r (objective 0), and by recomputing
||Z M r - r|| from M r directly.
"""
import numpy as np

HIDDEN = 64
LAYERS = 4
F_IN = 6
MAX_DEG = 4


def param_layout():
    """(name, shape) in the order of the flat FP32 parameter vector."""
    out = [("W_e", (HIDDEN, F_IN))]
    for l in range(LAYERS):
        out += [(f"W_s{l}", (HIDDEN, HIDDEN)), (f"W_n{l}", (HIDDEN, HIDDEN)), (f"b{l}", (HIDDEN,))]
    out.append(("W_d", (2, HIDDEN)))
    return out


def n_params():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unflatten(theta):
    p, off = {}, 0
    for name, shape in param_layout():
        cnt = int(np.prod(shape))
        p[name] = np.array(theta[off:off + cnt], dtype=np.float64).reshape(shape)
        off += cnt
    return p


def flatten(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in param_layout()])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {"W_e": rng.standard_normal((HIDDEN, F_IN)) / np.sqrt(F_IN)}
    for l in range(LAYERS):
        p[f"W_s{l}"] = 0.5 * rng.standard_normal((HIDDEN, HIDDEN)) / np.sqrt(HIDDEN)
        p[f"W_n{l}"] = 0.5 * rng.standard_normal((HIDDEN, HIDDEN)) / np.sqrt(HIDDEN)
        p[f"b{l}"] = np.zeros(HIDDEN)
    p["W_d"] = np.zeros((2, HIDDEN))
    return flatten(p).astype(np.float32)


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def build_graph(tplus, tminus, nt):
    """Bases sharing a triangle (0-based triangle ids).  Returns nbr (N x 4,
    -1 padded, neighbours in increasing basis index), sgn (N x 4) and deg."""
    tplus = np.asarray(tplus, np.int64)
    tminus = np.asarray(tminus, np.int64)
    N = len(tplus)
    on_tri = [[] for _ in range(nt)]
    for i in range(N):
        on_tri[tplus[i]].append((i, 1.0))
        on_tri[tminus[i]].append((i, -1.0))
    nbr = -np.ones((N, MAX_DEG), np.int64)
    sgn = np.zeros((N, MAX_DEG))
    deg = np.zeros(N, np.int64)
    for i in range(N):
        lst = []
        for t, si in ((tplus[i], 1.0), (tminus[i], -1.0)):
            for j, sj in on_tri[t]:
                if j != i:
                    lst.append((j, si * sj))
        lst.sort()
        if len(lst) > MAX_DEG:
            raise ValueError("a basis shares its triangles with more than 4 bases")
        for q, (j, s) in enumerate(lst):
            nbr[i, q] = j
            sgn[i, q] = s
        deg[i] = len(lst)
    return nbr, sgn, deg


def geometry(length, area_plus, area_minus, k):
    length = np.asarray(length, float)
    g1 = k * length / (2 * np.pi / 10)
    g2 = (np.asarray(area_plus, float) + np.asarray(area_minus, float)) / length**2
    return np.stack([g1, g2], axis=1)


def normaliser(r):
    """c of the docstring (0 when r == 0)."""
    r = np.asarray(r, np.complex128)
    rho = np.sqrt(np.mean(np.abs(r) ** 2))
    if not rho > 0:
        return 0.0 + 0.0j
    q = np.sum(r * np.abs(r))
    if abs(q) > 0:
        return np.conj(q) / (abs(q) * rho)
    return 1.0 / rho + 0.0j


def features(z, r, geom):
    c = normaliser(r)
    cz, cr = c * z, c * r
    tau = np.sqrt(np.mean(np.abs(cz) ** 2))
    cz = cz / tau if tau > 0 else 0 * cz
    u = np.column_stack([cz.real, cz.imag, cr.real, cr.imag, geom[:, 0], geom[:, 1]])
    return u, c


def _aggregate(h, nbr, sgn, deg):
    m = np.zeros_like(h)
    for q in range(MAX_DEG):
        j = nbr[:, q]
        ok = j >= 0
        m[ok] += sgn[ok, q, None] * h[j[ok]]
    return m / np.maximum(deg, 1)[:, None]


def _scatter_back(gm, nbr, sgn, deg):
    """Adjoint of _aggregate: gh_j += sum_{i: j in N(i)} s_ij gm_i / deg_i."""
    gh = np.zeros_like(gm)
    w = gm / np.maximum(deg, 1)[:, None]
    for q in range(MAX_DEG):
        j = nbr[:, q]
        ok = j >= 0
        np.add.at(gh, j[ok], sgn[ok, q, None] * w[ok])
    return gh


def forward(theta, u, nbr, sgn, deg, bf16=True):
    """Returns y (N x 2) and the cache for backward."""
    th = bf16_round(theta).astype(np.float64) if bf16 else np.asarray(theta, np.float64)
    p = unflatten(th)
    h = u @ p["W_e"].T
    cache = {"u": u, "h": [h], "t": [], "m": []}
    for l in range(LAYERS):
        m = _aggregate(h, nbr, sgn, deg)
        a = h @ p[f"W_s{l}"].T + m @ p[f"W_n{l}"].T + p[f"b{l}"]
        t = np.tanh(a)
        h = h + t
        cache["m"].append(m)
        cache["t"].append(t)
        cache["h"].append(h)
    y = h @ p["W_d"].T
    cache["p"] = p
    return y, cache


def sample_residual_loss(y, r, c, A):
    """Objective of one sample and d(objective)/dy:
    ||A y - c p||^2 / (|c|^2 ||r||^2), p = r - A r."""
    A = np.asarray(A, np.complex128)
    p = r - A @ r
    q = A @ (y[:, 0] + 1j * y[:, 1]) - c * p
    den = abs(c) ** 2 * np.sum(np.abs(r) ** 2)
    loss = np.sum(np.abs(q) ** 2) / den
    g = 2 * (np.conj(A).T @ q) / den
    return loss, np.column_stack([g.real, g.imag])


def backward(cache, gy, nbr, sgn, deg):
    """Gradient of the loss w.r.t. the (bf16-rounded) parameters, flat."""
    p = cache["p"]
    g = {name: np.zeros(shape) for name, shape in param_layout()}
    h4 = cache["h"][LAYERS]
    g["W_d"] = gy.T @ h4
    gh = gy @ p["W_d"]
    for l in reversed(range(LAYERS)):
        h_in, t, m = cache["h"][l], cache["t"][l], cache["m"][l]
        ga = gh * (1 - t**2)
        g[f"W_s{l}"] = ga.T @ h_in
        g[f"W_n{l}"] = ga.T @ m
        g[f"b{l}"] = ga.sum(axis=0)
        gm = ga @ p[f"W_n{l}"]
        gh = gh + ga @ p[f"W_s{l}"] + _scatter_back(gm, nbr, sgn, deg)
    g["W_e"] = gh.T @ cache["u"]
    return flatten(g)


def loss_and_grad(theta, samples, bf16=True):
    """Mean objective over samples and its gradient.  Each sample is a dict
    with z = S r, r (complex N), geom (N x 2), nbr, sgn, deg and its mesh's
    operator A = Z S (N x N)."""
    tot, grad = 0.0, np.zeros(n_params())
    for s in samples:
        u, c = features(s["z"], s["r"], s["geom"])
        y, cache = forward(theta, u, s["nbr"], s["sgn"], s["deg"], bf16)
        loss, gy = sample_residual_loss(y, s["r"], c, s["A"])
        tot += loss
        grad += backward(cache, gy, s["nbr"], s["sgn"], s["deg"])
    return tot / len(samples), grad / len(samples)


def apply(theta, z, r, geom, nbr, sgn, deg):
    """The network's correction d = (y0 + i y1) / c (0 when r == 0)."""
    u, c = features(z, r, geom)
    if c == 0:
        return np.zeros_like(np.asarray(r, np.complex128))
    y, _ = forward(theta, u, nbr, sgn, deg)
    return (y[:, 0] + 1j * y[:, 1]) / c


def precondition(theta, S, r, geom, nbr, sgn, deg):
    """M r = S r + S d(S r, r) with S a dense matrix."""
    z = S @ r
    return z + S @ apply(theta, z, r, geom, nbr, sgn, deg)


def adam_step(theta, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t = step number starting at 1); returns new theta, m, v."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1**t)
    vh = v / (1 - b2**t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
