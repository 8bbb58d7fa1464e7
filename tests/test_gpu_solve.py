"""GPU solves (dense LU, GMRES with/without preconditioners, mixed precision,
solve_scattering) against the oracle's FP64 solution and the reference's own
bars: LU residual < 1e-10 (test/runtests.jl:2308), GMRES vs direct < 1e-6
(test/runtests.jl:7426,8793,9986), sphere RCS vs Mie (test/runtests.jl:5237-5240)."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, oracle_rwg

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
    """configs[1]-shaped solve at reduced size (plate 0.5 m, 30x30): GMRES to
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
