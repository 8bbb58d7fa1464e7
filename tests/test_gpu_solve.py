"""GPU solves (dense LU, GMRES with/without preconditioners, mixed precision,
solve_scattering) against the oracle's FP64 solution and the reference's own
bars: LU residual < 1e-10 (test/runtests.jl:2308), GMRES vs direct < 1e-6
(test/runtests.jl:7426,8793,9986), sphere RCS vs Mie (test/runtests.jl:5237-5240)."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, oracle_rwg, rel_fro

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")
K0 = 2 * math.pi / 0.1


@pytest.fixture(scope="module")
def c0(dm):
    m = dm.make_rect_plate(0.1, 0.1, 6, 6)
    r = dm.build_rwg(m)
    Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
    return m, r, Z, v


def test_dense_direct(dm, c0):
    npz = np.load(os.path.join(GOLD, "c0_plate.npz"))
    m, r, Z, v = c0
    I = dm.solve_forward(Z, v, solver="direct")
    Zh = Z.to_numpy()
    assert np.linalg.norm(Zh @ I - v) / np.linalg.norm(v) < 1e-10
    assert np.linalg.norm(I - npz["I"]) / np.linalg.norm(npz["I"]) < 1e-9


def test_dense_direct_blocked_panels(dm):
    """Blocked LU over many 32-column panels with a partial last one (N =
    1976 on a 26 x 26 plate): true residual < 1e-10 and the currents of
    numpy's LU (LAPACK zgesv on the same matrix) to 1e-9."""
    m = dm.make_rect_plate(0.3, 0.3, 26, 26)
    r = dm.build_rwg(m)
    assert r.nedges % 32 != 0 and r.nedges > 40 * 32
    Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64, device=True)
    v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
    I = dm.solve_forward(Z, v, solver="direct")
    Zh = Z.to_numpy()
    assert np.linalg.norm(Zh @ I - v) / np.linalg.norm(v) < 1e-10
    In = np.linalg.solve(Zh, v)
    assert np.linalg.norm(I - In) / np.linalg.norm(In) < 1e-9


@pytest.mark.parametrize("precond", [None, "diag", "nf", "block"])
@pytest.mark.parametrize("side", ["left", "right"])
def test_gmres_matches_direct(dm, c0, precond, side):
    m, r, Z, v = c0
    Id = dm.solve_forward(Z, v, solver="direct")
    P = None
    if precond == "diag":
        P = dm.build_nearfield_preconditioner(Z, m, r, 0.0, factorization="diag")
    elif precond == "nf":
        P = dm.build_nearfield_preconditioner(Z, m, r, 0.03)
    elif precond == "block":
        P = dm.build_nearfield_preconditioner(Z, m, r, 0.0, factorization="block", block_size=32)
    x, st = dm.solve_gmres(Z, v, preconditioner=P, precond_side=side, tol=1e-8, maxiter=200)
    assert st.solved
    assert np.linalg.norm(x - Id) / np.linalg.norm(Id) < 1e-6


def test_gmres_restarted_and_nonconvergence(dm, c0):
    m, r, Z, v = c0
    Id = dm.solve_forward(Z, v, solver="direct")
    x, st = dm.solve_gmres(Z, v, tol=1e-8, maxiter=400, restart=30)
    assert st.solved and np.linalg.norm(x - Id) / np.linalg.norm(Id) < 1e-6
    with pytest.raises(dm.ConvergenceError):
        dm.solve_gmres(Z, v, tol=1e-12, maxiter=2)  # test/runtests.jl:9990-9998
    x, st = dm.solve_gmres(Z, v, tol=1e-12, maxiter=2, check_gmres_convergence=False)
    assert not st.solved and st.niter == 2


def test_mixed_precision_refinement(dm, c0):
    m, r, Z, v = c0
    Id = dm.solve_forward(Z, v, solver="direct")
    P = dm.build_nearfield_preconditioner(Z, m, r, 0.03)
    x, st = dm.solve_mixed(Z, v, P, tol=1e-10, maxiter=300)
    assert st.solved and st.residual <= 1e-10
    assert np.linalg.norm(x - Id) / np.linalg.norm(Id) < 1e-8
    assert st.outer >= 1


def test_multiple_rhs(dm, c0):
    m, r, Z, v = c0
    exc = [dm.make_plane_wave(K0 * np.array([math.sin(t), 0, math.cos(t)]), 1.0,
                              [math.cos(t), 0, -math.sin(t)]) for t in (0.1, 0.5, 1.0)]
    V = dm.assemble_multiple_excitations(m, r, exc)
    X, st = dm.solve_gmres(Z, V, tol=1e-8, maxiter=200)
    Zh = Z.to_numpy()
    assert np.linalg.norm(Zh @ X - V) / np.linalg.norm(V) < 1e-7


def test_plate_10k_gmres_true_residual(dm):
    """The 0.5 m plate solve at reduced size (30x30 instead of 60x60): GMRES to
    1e-8 with the near-field preconditioner and a theta=30 deg incidence."""
    m = dm.make_rect_plate(0.5, 0.5, 30, 30)
    r = dm.build_rwg(m)
    Z = dm.assemble_Z_efie(m, r, K0, device=True)
    th = math.radians(30)
    v = dm.assemble_v_plane_wave(m, r, K0 * np.array([math.sin(th), 0, math.cos(th)]), 1.0,
                                 [math.cos(th), 0, -math.sin(th)])
    P = dm.build_nearfield_preconditioner(Z, m, r, 0.0)
    x, st = dm.solve_gmres(Z, v, preconditioner=P, precond_side="right", tol=1e-8, maxiter=500)
    x0, st0 = dm.solve_gmres(Z, v, precond_side="right", tol=1e-8, maxiter=500)
    print(f"plate 30x30: {st.niter} iterations with the near-field preconditioner, {st0.niter} without")
    assert st.niter < st0.niter
    assert st.solved and st.residual < 1e-7
    Id = dm.solve_forward(Z, v, solver="direct")
    assert np.linalg.norm(x - Id) / np.linalg.norm(Id) < 1e-6


def test_solve_scattering_sphere_mie_gate(dm, O):
    """test/runtests.jl:5175-5242 through the GPU path, and parity of the
    currents with the oracle's FP64 dense solve."""
    m = dm.make_icosphere(0.05, 2)
    f = 2.0e9
    k = 2 * math.pi * f / dm.C0_DEFAULT
    khat, pol = np.array([0, 0, -1.0]), np.array([1.0, 0, 0])
    pw = dm.make_plane_wave(k * khat, 1.0, pol)
    res = dm.solve_scattering(m, f, pw, method="dense_direct")
    assert res.method == "dense_direct" and res.N == 480
    r = dm.build_rwg(m)
    orr = oracle_rwg(O, m, r)
    Zo = O.assemble_Z_efie(orr, k)
    vo = O.assemble_v_plane_wave(orr, k * khat, 1.0, pol)
    Io = O.solve_direct(Zo, vo)
    assert np.linalg.norm(res.I_coeffs - Io) / np.linalg.norm(Io) < 1e-9
    grid = dm.make_sph_grid(37, 18)
    G = dm.radiation_vectors(m, r, grid, k)
    E = dm.compute_farfield(G, res.I_coeffs, len(grid.w))
    sig = dm.bistatic_rcs(E)
    Eo = O.compute_farfield(O.radiation_vectors(orr, grid.rhat, k), Io)
    assert np.linalg.norm(sig - O.bistatic_rcs(Eo)) / np.linalg.norm(sig) < 1e-8
    a = 0.05
    idx = sorted([q for q in range(len(grid.w)) if abs(grid.phi[q] - grid.phi.min()) < 1e-12],
                 key=lambda q: grid.theta[q])
    mie = np.array([O.mie_bistatic_rcs_pec(k, a, khat, pol, grid.rhat[q]) for q in idx])
    d = 10 * np.log10(sig[idx]) - 10 * np.log10(mie)
    assert np.mean(np.abs(d)) < 0.5 and math.sqrt(np.mean(d ** 2)) < 0.6 and np.abs(d).max() < 1.2
    # the GMRES and mixed-precision paths agree with the direct one
    for prec in ("fp64", "mixed"):
        rg = dm.solve_scattering(m, f, pw, method="dense_gmres", gmres_tol=1e-8, precision=prec)
        assert np.linalg.norm(rg.I_coeffs - Io) / np.linalg.norm(Io) < 1e-6


def test_solve_scattering_nearfield_options_are_honoured_or_refused(dm):
    """preconditioner=:lu / :ilu are refused (the GPU path builds a sparse
    approximate inverse, not a factorisation); nf_cutoff_lambda is used as
    given, and a cutoff wider than the pattern's column cap reaches is
    reported in the result's warnings with the radius actually covered."""
    m = dm.make_icosphere(0.05, 2)
    f = 2.0e9
    k = 2 * math.pi * f / dm.C0_DEFAULT
    pw = dm.make_plane_wave([0, 0, -k], 1.0, [1.0, 0, 0])
    for pc in ("lu", "ilu"):
        with pytest.raises(ValueError):
            dm.solve_scattering(m, f, pw, method="dense_gmres", preconditioner=pc)
    res = dm.solve_scattering(m, f, pw, method="dense_gmres", gmres_tol=1e-8)  # nf_cutoff_lambda = 1
    assert any("keeps the 80 nearest" in w for w in res.warnings)
    r = dm.build_rwg(m)
    lam = dm.C0_DEFAULT / f
    small = dm.build_nearfield_preconditioner(None, m, r, 0.2 * lam)
    nnz, rad, cap, capped = small.pattern_info(r)
    assert cap == 80 and capped == 0 and 0 < rad <= 0.2 * lam and nnz < 80 * r.nedges
    assert dm.build_nearfield_preconditioner(None, m, r, lam).pattern_info(r)[3] == r.nedges
    res2 = dm.solve_scattering(m, f, pw, method="dense_gmres", gmres_tol=1e-8, nf_cutoff_lambda=0.2)
    assert not any("nearest" in w for w in res2.warnings)
    assert np.linalg.norm(res2.I_coeffs - res.I_coeffs) / np.linalg.norm(res.I_coeffs) < 1e-6


def test_nearfield_preconditioner_columns_and_gain(dm, O):
    """Each column of the near-field sparse approximate inverse solves the same
    least-squares problem numpy solves on the oracle's matrix, and the
    preconditioner cuts GMRES iterations on a sphere."""
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    Z = dm.assemble_Z_efie(m, r, K0, device=True)
    v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
    P = dm.build_nearfield_preconditioner(Z, m, r, 0.0)
    x1, st1 = dm.solve_gmres(Z, v, P, precond_side="right", tol=1e-6, maxiter=500)
    x0, st0 = dm.solve_gmres(Z, v, None, precond_side="right", tol=1e-6, maxiter=500)
    print(f"sphere N={r.nedges}: {st1.niter} iterations with the near-field preconditioner, {st0.niter} without")
    assert st1.solved and st1.niter * 2 < st0.niter
    assert np.linalg.norm(x1 - x0) / np.linalg.norm(x0) < 1e-5
    # columns of M against numpy least squares over the same pattern on the oracle's Z
    Zo = O.assemble_Z_efie(oracle_rwg(O, m, r), K0)
    cen = O.rwg_centers(oracle_rwg(O, m, r))
    cutJ = 2.5 * float(np.mean(r.len))
    op = dm.NearFieldOperator(P, Z)
    for j in (0, 17, 999, r.nedges - 1):
        e = np.zeros(r.nedges, complex)
        e[j] = 1
        col = op @ e
        d = np.linalg.norm(cen - cen[j], axis=1)
        J = np.where(d <= cutJ)[0]
        assert len(J) < 80
        I = np.where(d <= 1.6 * d[J].max())[0]
        mj, *_ = np.linalg.lstsq(Zo[np.ix_(I, J)], (I == j).astype(complex), rcond=None)
        ref = np.zeros(r.nedges, complex)
        ref[J] = mj
        assert set(np.nonzero(col)[0]) == set(J)
        assert np.linalg.norm(col - ref) / np.linalg.norm(ref) < 1e-8


@pytest.mark.parametrize("method", ["direct", "fp64", "mixed"])
def test_end_to_end_planewaves(dm, O, method):
    """One-call device solve of several plane waves (assembly on two streams in
    the mixed method) against the oracle's FP64 dense solution."""
    m = dm.make_icosphere(0.1, 2)
    r = dm.build_rwg(m)
    exc = [dm.make_plane_wave(K0 * np.array([0, 0, -1.0]), 1.0, [1.0, 0, 0]),
           dm.make_plane_wave(K0 * np.array([1.0, 0, 0]), 2.0, [0, 0.6, 0.8])]
    P = None if method == "direct" else dm.build_nearfield_preconditioner(None, m, r, 0.0)
    X, st, t_asm = dm.solve_planewaves(m, r, K0, exc, method, P, tol=1e-9, maxiter=500)
    orr = oracle_rwg(O, m, r)
    Zo = O.assemble_Z_efie(orr, K0)
    for j, e in enumerate(exc):
        vo = O.assemble_v_plane_wave(orr, e.k_vec, e.E0, e.pol)
        Io = O.solve_direct(Zo, vo)
        assert np.linalg.norm(X[:, j] - Io) / np.linalg.norm(Io) < 1e-7
    assert st.solved and t_asm > 0
    if method == "mixed":
        assert st.outer >= 1


def _random_planewaves(dm, rng, k, count):
    """Directions normalised N(0, I3) draws, polarisations random unit vectors
    projected perpendicular to k (the many-angle workload's recipe)."""
    out = []
    for _ in range(count):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        p = rng.standard_normal(3)
        p -= d * (p @ d)
        p /= np.linalg.norm(p)
        out.append(dm.make_plane_wave(k * d, 1.0, p))
    return out


@pytest.fixture(scope="module")
def sphere2(dm):
    m = dm.make_icosphere(0.1, 2)
    r = dm.build_rwg(m)
    Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    rng = np.random.default_rng(20261202)
    pws = _random_planewaves(dm, rng, K0, 12)
    V = dm.assemble_multiple_excitations(m, r, pws)
    return m, r, Z, V


@pytest.mark.parametrize("method", ["gmres", "mixed"])
@pytest.mark.parametrize("precond", [None, "nf"])
def test_batched_solve_matches_per_column(dm, sphere2, method, precond):
    """Lockstep batch (one Z X product per step, per-column Arnoldi/Givens) vs
    the one-column solver: same true residual bar per column, same currents to
    the solve tolerance, and iteration counts that agree column for column up
    to rounding (the batch reduces dot products in another order)."""
    m, r, Z, V = sphere2
    P = dm.build_nearfield_preconditioner(Z, m, r, 2.0 * float(np.mean(r.len))) if precond else None
    tol = 1e-8 if method == "gmres" else 1e-7
    f = dm.solve_gmres if method == "gmres" else dm.solve_mixed
    kw = dict(preconditioner=P, tol=tol, maxiter=300)
    if method == "gmres":
        kw["precond_side"] = "right"
    Xb, sb = f(Z, V, batch=1, **kw)
    Xs, ss = f(Z, V, batch=-1, **kw)
    Zh = Z.to_numpy()
    res = np.linalg.norm(Zh @ Xb - V, axis=0) / np.linalg.norm(V, axis=0)
    bar = tol * (1.05 if method == "gmres" else 1.0)
    assert sb.solved and res.max() <= max(bar, sb.residual * 1.01)
    assert sb.residual <= (tol if method == "mixed" else 10 * tol)
    assert np.linalg.norm(Xb - Xs) / np.linalg.norm(Xs) < 1e-5
    # the batch's complex64 products accumulate in FP32 on the matrix cores
    # (three-part bf16 split of the complex64 inputs, 4096-term chunks), the
    # one-column GEMV in FP64: the inner solves differ by rounding, so
    # mixed-precision counts may drift slightly more.  With the near-field
    # preconditioner, 12 columns exercise a partial column group of the
    # batched sparse product.
    slack = 0.03 if method == "gmres" else 0.10
    assert abs(sb.niter - ss.niter) <= max(2, slack * ss.niter)
    Id = np.linalg.solve(Zh, V)
    assert np.linalg.norm(Xb - Id) / np.linalg.norm(Id) < 1e-4


def test_batched_nearfield_many_column_groups(dm, sphere2):
    """40 right-hand sides in one lockstep batch with the near-field
    preconditioner: the batched sparse product takes 16 columns per pass over
    each row (groups of 16, 16 and 8 here), and every column must match its
    one-column solve (the product adds each row's partial sums in the
    one-column kernel's order)."""
    m, r, Z, _ = sphere2
    rng = np.random.default_rng(20261203)
    V = dm.assemble_multiple_excitations(m, r, _random_planewaves(dm, rng, K0, 40))
    P = dm.build_nearfield_preconditioner(Z, m, r, 2.0 * float(np.mean(r.len)))
    kw = dict(preconditioner=P, tol=1e-8, maxiter=300, precond_side="right")
    Xb, sb = dm.solve_gmres(Z, V, batch=1, **kw)
    Xs, ss = dm.solve_gmres(Z, V, batch=-1, **kw)
    assert sb.solved and ss.solved
    Zh = Z.to_numpy()
    res = np.linalg.norm(Zh @ Xb - V, axis=0) / np.linalg.norm(V, axis=0)
    assert res.max() <= max(1.05e-8, sb.residual * 1.01)
    rel = np.linalg.norm(Xb - Xs, axis=0) / np.linalg.norm(Xs, axis=0)
    assert rel.max() < 1e-5
    assert abs(sb.niter - ss.niter) <= max(2, 0.03 * ss.niter)


def test_batched_restart_single_column_and_nonconvergence(dm, sphere2):
    m, r, Z, V = sphere2
    Zh = Z.to_numpy()
    Id = np.linalg.solve(Zh, V[:, :9])
    X, st = dm.solve_gmres(Z, V[:, :9], precond_side="right", tol=1e-8, maxiter=600, restart=25)
    assert st.solved and np.linalg.norm(X - Id) / np.linalg.norm(Id) < 1e-6
    # a batch of one reproduces the one-column solver
    x1, s1 = dm.solve_gmres(Z, V[:, 0], precond_side="right", tol=1e-8, maxiter=300, batch=1)
    x0, s0 = dm.solve_gmres(Z, V[:, 0], precond_side="right", tol=1e-8, maxiter=300, batch=-1)
    assert abs(s1.niter - s0.niter) <= 1
    assert np.linalg.norm(x1 - x0) / np.linalg.norm(x0) < 1e-7
    with pytest.raises(dm.ConvergenceError):
        dm.solve_gmres(Z, V, precond_side="right", tol=1e-12, maxiter=3)
    X, st = dm.solve_gmres(Z, V, precond_side="right", tol=1e-12, maxiter=3, check_gmres_convergence=False)
    assert not st.solved and st.niter == 3 * V.shape[1]


def test_batched_planewaves_end_to_end(dm):
    """64 plane waves on a 1920-unknown sphere in one call: mixed precision,
    SPAI, lockstep batches; every column's true FP64 residual <= 1e-6 and the
    currents match the dense direct solve."""
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    rng = np.random.default_rng(7)
    pws = _random_planewaves(dm, rng, K0, 64)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    X, st, t_asm, Z = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, keep_matrix=True)
    assert st.solved and st.residual <= 1e-6
    Zh = Z.to_numpy()
    V = dm.assemble_multiple_excitations(m, r, pws)
    res = np.linalg.norm(Zh @ X - V, axis=0) / np.linalg.norm(V, axis=0)
    assert res.max() <= 1e-6 * 1.001
    Id = np.linalg.solve(Zh, V)
    assert np.linalg.norm(X - Id) / np.linalg.norm(Id) < 1e-4
    Xs, ss, _ = dm.solve_planewaves(m, r, K0, pws[:3], "mixed", P, tol=1e-6, maxiter=300, batch=-1)
    assert np.linalg.norm(Xs - X[:, :3]) / np.linalg.norm(X[:, :3]) < 1e-4


def test_batched_relaxed_products_reach_the_true_residual(dm):
    """Lockstep batches with relaxed products (dm_solve_opts.relax): once the
    worst active column's residual estimate allows it, the complex64 Z X drops
    to three bf16 part products (h h', h m', m h'); every column's true FP64
    residual still meets the tolerance, with about the iterations and currents
    of the unrelaxed batch."""
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    rng = np.random.default_rng(11)
    pws = _random_planewaves(dm, rng, K0, 16)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    ctx = m.ctx
    X0, s0, _ = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, batch=1)
    ctx.profile_reset()
    ctx.profile(True)
    X1, s1, _, Z = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, batch=1, relax=0.01,
                                       keep_matrix=True)
    ctx.profile(False)
    n3, _ = ctx.profile_read("zx_z32_relaxed")
    n6, _ = ctx.profile_read("zx_z32")
    ctx.profile_reset()
    print(f"relaxed batch: {s1.niter} it, {s1.outer} refinements ({n3} three-part, {n6} six-part products); "
          f"unrelaxed {s0.niter} it, {s0.outer}", flush=True)
    assert n3 > 0 and n6 > 0
    assert s1.solved and s1.residual <= 1e-6
    V = dm.assemble_multiple_excitations(m, r, pws)
    Zh = Z.to_numpy()
    res = np.linalg.norm(Zh @ X1 - V, axis=0) / np.linalg.norm(V, axis=0)
    assert res.max() <= 1e-6 * 1.001
    assert s1.niter <= 1.05 * s0.niter + 16 and s1.outer <= s0.outer + 1
    assert np.linalg.norm(X1 - X0) <= 1e-4 * np.linalg.norm(X0)
    Z.free()


@pytest.mark.parametrize("nw", [1, 12])
def test_mixed_pipeline_complex64_is_rounded_complex128(dm, O, nw):
    """The mixed solve assembles once in FP64 and stores the complex64 matrix as
    the rounding of the complex128 one (packed forms for one wave, full forms
    for lockstep batches); the complex128 matrix meets the FP64 bar against the
    oracle.  A plate (N = 338, not a multiple of the 128-wide storage tile)
    exercises the zeroed padding of packed storage."""
    m = dm.make_rect_plate(0.12, 0.1, 12, 10)
    r = dm.build_rwg(m)
    assert r.nedges == 338
    pws = _random_planewaves(dm, np.random.default_rng(3), K0, nw)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    X, st, _, Z = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, keep_matrix=True)
    assert st.solved and st.residual <= 1e-6
    f64, f32 = (dm.DM_Z64P, dm.DM_Z32P) if nw == 1 else (dm.DM_Z64, dm.DM_Z32)
    assert Z.forms == f64 | f32
    Z64, Z32 = Z.to_numpy(f64), Z.to_numpy(f32)
    assert np.array_equal(Z32, Z64.astype(np.complex64))
    assert rel_fro(Z64, O.assemble_Z_efie(oracle_rwg(O, m, r), K0, quad_order=3)) <= 1e-10
    V = dm.assemble_multiple_excitations(m, r, pws)
    res = np.linalg.norm(Z64 @ X - V, axis=0) / np.linalg.norm(V, axis=0)
    assert res.max() <= 1e-6 * 1.001
    # padding entries are zero: the symmetric product of the packed forms
    # matches the dense product
    if nw == 1:
        x = X[:, 0]
        y = Z.matmul(x, form=dm.DM_Z32P)
        assert np.linalg.norm(y - Z32.astype(np.complex128) @ x) / np.linalg.norm(y) < 1e-12


@pytest.mark.parametrize("method", ["gmres", "mixed"])
def test_solves_on_packed_forms(dm, method):
    """GMRES / mixed-precision FGMRES on packed symmetric storage match the
    same solves on full storage (same entries, products to FP64 rounding)."""
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
    Zf = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    Zp = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P, device=True)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    if method == "gmres":
        xf, sf = dm.solve_gmres(Zf, v, preconditioner=P, precond_side="right", tol=1e-8)
        xp, sp = dm.solve_gmres(Zp, v, preconditioner=P, precond_side="right", tol=1e-8)
    else:
        xf, sf = dm.solve_mixed(Zf, v, preconditioner=P, tol=1e-7)
        xp, sp = dm.solve_mixed(Zp, v, preconditioner=P, tol=1e-7)
    assert sf.solved and sp.solved
    assert abs(sf.niter - sp.niter) <= 2
    assert np.linalg.norm(xp - xf) / np.linalg.norm(xf) < 1e-5
    Z = Zf.to_numpy()
    assert np.linalg.norm(Z @ xp - v) / np.linalg.norm(v) <= (1.05e-8 if method == "gmres" else 1e-7)


def test_adjoint_solve(dm, c0):
    """solve_gmres_adjoint: Z^H x = b (src/solver/IterativeSolve.jl:409-460)."""
    m, r, Z, v = c0
    Zh = Z.to_numpy()
    x, st = dm.solve_gmres_adjoint(Z, v, tol=1e-10, maxiter=200)
    assert st.solved
    assert np.linalg.norm(Zh.conj().T @ x - v) / np.linalg.norm(v) < 1e-9
    xd = np.linalg.solve(Zh.conj().T, v)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    X, st = dm.solve_gmres_adjoint(Z, np.stack([v, 1j * v, v.conj()], 1), precond_side="right", tol=1e-10, batch=1)
    assert np.linalg.norm(Zh.conj().T @ X[:, 2] - v.conj()) / np.linalg.norm(v) < 1e-9
    xl = dm._solve(Z, v, dm.DM_METHOD_LU, adjoint=True)[0]
    assert np.linalg.norm(xl - xd) / np.linalg.norm(xd) < 1e-10


def test_mixed_solve_on_compressed_matrix(dm):
    """Mixed-precision solve with the compressed complex64 matrix for the inner
    products: true FP64 residual <= 1e-6 against the dense complex128 matrix,
    currents within 1e-5 of the uncompressed mixed solve; a loose compression
    (1e-3) still converges through extra refinement steps."""
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    pws = _random_planewaves(dm, np.random.default_rng(12), K0, 2)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    X0, s0, _ = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, batch=-1)
    X1, s1, _, Z = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, batch=-1,
                                       compress_tol=5e-8, keep_matrix=True)
    assert Z.forms & dm.DM_Z32H
    assert s1.solved and s1.residual <= 1e-6
    Z64 = Z.to_numpy(dm.DM_Z64P)
    V = dm.assemble_multiple_excitations(m, r, pws)
    res = np.linalg.norm(Z64 @ X1 - V, axis=0) / np.linalg.norm(V, axis=0)
    assert res.max() <= 1e-6 * 1.001
    assert np.linalg.norm(X1 - X0) / np.linalg.norm(X0) < 1e-5
    X2, s2, _ = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, batch=-1, compress_tol=1e-3)
    assert s2.solved and s2.residual <= 1e-6 and s2.outer >= s1.outer
    # dm_solve with a compressed matrix: the mixed method uses it
    Zc = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P, device=True)
    Zc.compress(eps=5e-8)
    x, st = dm.solve_mixed(Zc, V[:, 0], preconditioner=P, tol=1e-7)
    assert st.solved
    assert np.linalg.norm(Z64 @ x - V[:, 0]) / np.linalg.norm(V[:, 0]) <= 1e-7 * 1.001


def test_repeated_solves_reuse_matrix_buffers(dm):
    """An RWG handle keeps its last solve's matrix buffers and a context keeps
    released device blocks: repeated solves on one handle, interleaved with
    solves of other forms, with a matrix handed to the caller and with another
    handle of a different size, give the same currents (to the solve's own
    accuracy: the tile kernel's LDS atomics add a pair's four contributions in
    no fixed order, so repeated assemblies may differ in the last bit) and the
    true FP64 residual stays below the tolerance."""
    m = dm.make_icosphere(0.1, 2)
    r = dm.build_rwg(m)
    m2 = dm.make_rect_plate(0.2, 0.15, 12, 9)
    r2 = dm.build_rwg(m2)
    pws = _random_planewaves(dm, np.random.default_rng(31), K0, 1)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    P2 = dm.build_nearfield_preconditioner(None, m2, r2, 2.0 * float(np.mean(r2.len)))

    def chosen():
        X, st, _ = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, batch=-1,
                                       compress_tol=5e-8)
        assert st.solved and st.residual <= 1e-6
        return X

    def same(X, Y):
        return np.linalg.norm(X - Y) / np.linalg.norm(Y) < 1e-5

    X0 = chosen()
    assert same(chosen(), X0)  # the kept buffers
    Xf, sf, _ = dm.solve_planewaves(m, r, K0, pws, "fp64", P, tol=1e-6, maxiter=300)
    assert sf.solved
    assert same(chosen(), X0)  # after a solve that kept other forms
    Xp, sp, _ = dm.solve_planewaves(m2, r2, K0, pws, "mixed", P2, tol=1e-6, maxiter=300, batch=-1,
                                    compress_tol=5e-8)
    assert sp.solved
    Xk, sk, _, Z = dm.solve_planewaves(m, r, K0, pws, "mixed", P, tol=1e-6, maxiter=300, batch=-1,
                                       compress_tol=5e-8, keep_matrix=True)
    assert same(Xk, X0)  # the matrix goes to the caller ...
    assert same(chosen(), X0)  # ... and the next solve allocates anew
    Xb, sb, _ = dm.solve_planewaves(m, r, K0, pws * 8, "mixed", P, tol=1e-6, maxiter=300)  # lockstep batch
    assert sb.solved
    assert same(chosen(), X0)
    assert same(Xf, X0)
    del Z


def test_c1_plate_chosen_form_matches_plain_fp64(dm):
    """The 0.5 m plate (60 x 60, N = 10680, theta = 30 deg, phi = 0,
    theta-polarised): the chosen form (one FP64 assembly with its complex64
    rounding, compressed complex64 matrix, SPAI, FGMRES, FP64 refinement)
    reaches a true FP64 residual <= 1e-6 against the dense complex128 matrix
    and the currents of the plain FP64 GMRES solve to the accuracy both solves
    guarantee."""
    m = dm.make_rect_plate(0.5, 0.5, 60, 60)
    r = dm.build_rwg(m)
    assert r.nedges == 10680
    th = math.radians(30)
    pw = [dm.make_plane_wave(K0 * np.array([math.sin(th), 0, math.cos(th)]), 1.0,
                             np.array([math.cos(th), 0, -math.sin(th)]))]
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    Xc, sc, _, Z = dm.solve_planewaves(m, r, K0, pw, "mixed", P, tol=1e-6, maxiter=500, batch=-1,
                                       compress_tol=5e-8, keep_matrix=True)
    Xp, sp, _ = dm.solve_planewaves(m, r, K0, pw, "fp64", P, tol=1e-8, maxiter=500)
    assert sc.solved and sp.solved
    v = dm.assemble_multiple_excitations(m, r, pw)[:, 0]
    res = np.linalg.norm(Z.matmul(Xc[:, 0], form=dm.DM_Z64P) - v) / np.linalg.norm(v)
    print(f"C1 plate: chosen {sc.niter} it / {sc.outer} refinements, residual {res:.2e}; plain {sp.niter} it")
    assert res <= 1e-6 * 1.001 and abs(res - sc.residual) <= 1e-9
    assert np.linalg.norm(Xc - Xp) / np.linalg.norm(Xp) < 1e-4


def test_graph_replayed_steps_match_launched_steps(dm):
    """Krylov steps replayed from a captured HIP graph (the step index read
    from device memory, one graph per basis-size bucket) against the same
    solves with every kernel launched: equal iteration counts and solutions
    to 1e-13, for right (SPAI), left (SPAI) and no preconditioning, full and
    restarted GMRES, and the mixed solve (two graph captures, one per
    refinement step); each graph solve runs three times in a row with other
    device work between, every run checked."""
    m = dm.make_rect_plate(0.1, 0.1, 12, 12)
    r = dm.build_rwg(m)
    Z = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P, device=True)
    v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    for Pc, side, restart in ((P, "right", 0), (P, "left", 0), (None, "right", 0), (P, "right", 12)):
        x0, s0 = dm.solve_gmres(Z, v, preconditioner=Pc, precond_side=side, tol=1e-10, maxiter=600, restart=restart)
        assert s0.solved
        for rep in range(3):
            x1, s1 = dm.solve_gmres(Z, v, preconditioner=Pc, precond_side=side, tol=1e-10, maxiter=600,
                                    restart=restart, graph=True)
            print(f"graph {side} restart={restart} rep {rep}: {s1.niter} it (launched {s0.niter})", flush=True)
            assert s1.solved and s1.niter == s0.niter
            assert np.linalg.norm(x1 - x0) <= 1e-13 * np.linalg.norm(x0)
            Z.matmul(x1)  # other device work between the replays
    xm0, sm0 = dm.solve_mixed(Z, v, preconditioner=P, tol=1e-9)
    for rep in range(2):
        xm1, sm1 = dm.solve_mixed(Z, v, preconditioner=P, tol=1e-9, graph=True)
        assert sm1.solved and (sm1.niter, sm1.outer) == (sm0.niter, sm0.outer)
        assert np.linalg.norm(xm1 - xm0) <= 1e-13 * np.linalg.norm(xm0)
    # end to end on a sphere with the compressed complex64 form
    ms = dm.make_icosphere(0.1, 3)
    rs = dm.build_rwg(ms)
    Ps = dm.build_nearfield_preconditioner(None, ms, rs, 2.0 * float(np.mean(rs.len)))
    pw = [dm.make_plane_wave([0, 0, -K0], 1.0, [1, 0, 0])]
    X0, st0, _ = dm.solve_planewaves(ms, rs, K0, pw, "mixed", Ps, tol=1e-6, maxiter=500, batch=-1,
                                     compress_tol=5e-8)
    X1, st1, _ = dm.solve_planewaves(ms, rs, K0, pw, "mixed", Ps, tol=1e-6, maxiter=500, batch=-1,
                                     compress_tol=5e-8, graph=True)
    assert st1.solved and st1.niter == st0.niter
    assert np.linalg.norm(X1 - X0) <= 1e-13 * np.linalg.norm(X0)


def test_relaxed_products_reach_the_true_residual(dm):
    """Inexact Krylov on the compressed form (dm_solve_opts.relax): the inner
    FGMRES reads looser prefixes of the cross-approximation factors as its
    residual falls; the FP64 refinement against the dense complex128 matrix
    still brings the true residual to the tolerance, with about the same
    number of iterations as the unrelaxed solve."""
    m = dm.make_icosphere(0.1, 4)
    r = dm.build_rwg(m)
    P = dm.build_nearfield_preconditioner(None, m, r, 2.0 * float(np.mean(r.len)))
    pw = [dm.make_plane_wave([0, 0, -K0], 1.0, [1, 0, 0])]
    X0, st0, _ = dm.solve_planewaves(m, r, K0, pw, "mixed", P, tol=1e-6, maxiter=500, batch=-1, compress_tol=5e-8)
    for c in (0.003, 0.01):
        X1, st1, _, Z = dm.solve_planewaves(m, r, K0, pw, "mixed", P, tol=1e-6, maxiter=500, batch=-1,
                                            compress_tol=5e-8, relax=c, keep_matrix=True)
        v = dm.assemble_v_plane_wave(m, r, [0, 0, -K0], 1.0, [1, 0, 0])
        res = np.linalg.norm(Z.matmul(X1[:, 0], form=dm.DM_Z64P) - v) / np.linalg.norm(v)
        print(f"relax {c}: {st1.niter} it, {st1.outer} refinements, true residual {res:.2e} "
              f"(unrelaxed {st0.niter} it, {st0.outer})", flush=True)
        assert st1.solved and res <= 1e-6
        assert st1.niter <= 1.05 * st0.niter + 2 and st1.outer <= st0.outer + 1
        assert np.linalg.norm(X1 - X0) <= 1e-4 * np.linalg.norm(X0)
        Z.free()
    with pytest.raises(ValueError):
        dm.solve_planewaves(m, r, K0, pw, "mixed", P, tol=1e-6, batch=-1, compress_tol=5e-8, relax=-1.0)
