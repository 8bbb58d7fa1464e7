"""GPU parity of assembly, products, excitation and radiation vectors against the
CPU oracle (FP64 bar: relative Frobenius <= 1e-10; complex64 path: <= 1e-5)."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, oracle_rwg, rel_fro

pytestmark = pytest.mark.gpu
GOLD = os.path.join(ROOT, "tests", "golden")
K0 = 2 * math.pi / 0.1


def _oracle_Z(O, mesh, rwg, k, quad_order=3):
    return O.assemble_Z_efie(oracle_rwg(O, mesh, rwg), k, quad_order=quad_order)


def test_c0_plate_z64_matches_fixture_and_oracle(dm, O):
    npz = np.load(os.path.join(GOLD, "c0_plate.npz"))
    m = dm.make_rect_plate(0.1, 0.1, 6, 6)
    r = dm.build_rwg(m)
    assert r.nedges == 96
    assert np.array_equal(r.tplus, npz["tplus"]) and np.array_equal(r.vminus_opp, npz["vminus_opp"])
    Z = dm.assemble_Z_efie(m, r, K0)
    assert rel_fro(Z, npz["Z"]) <= 1e-10
    assert rel_fro(Z, _oracle_Z(O, m, r, K0)) <= 1e-10
    assert np.array_equal(Z, Z.T)


def test_c0_plate_z32_within_bar(dm, O):
    m = dm.make_rect_plate(0.1, 0.1, 6, 6)
    r = dm.build_rwg(m)
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    Zo = _oracle_Z(O, m, r, K0)
    e32 = rel_fro(Zd.to_numpy(dm.DM_Z32), Zo)
    e64 = rel_fro(Zd.to_numpy(dm.DM_Z64), Zo)
    print(f"rel Frobenius: fp64 {e64:.2e}  fp32 {e32:.2e}")
    assert e64 <= 1e-10 and e32 <= 1e-5


@pytest.mark.parametrize("seed", [1, 2])
def test_caller_basis_order(dm, O, seed):
    """A caller's RWGData in arbitrary order (Julia Dict order) is honoured,
    including the order-dependent adjacent-cell outer/inner choice."""
    m = dm.make_rect_plate(0.12, 0.1, 7, 5)
    base = dm.build_rwg(m)
    perm = np.random.default_rng(seed).permutation(base.nedges)
    r = dm.rwg_from_arrays(m, base.tplus[perm], base.tminus[perm], base.vplus_opp[perm], base.vminus_opp[perm],
                           base.len[perm], base.area_plus[perm], base.area_minus[perm])
    Z = dm.assemble_Z_efie(m, r, K0)
    assert rel_fro(Z, _oracle_Z(O, m, r, K0)) <= 1e-10


@pytest.mark.parametrize("case", [0, 1], ids=["real_k", "lossy_k"])
def test_entries_vs_mp30_fixture(dm, case):
    """The HIP FP64 and FP32 matrices of the bent, jittered 2 x 2-cell plate in
    the fixture's permuted caller order against its 30-digit entries
    (tests/golden/make_cells_mp.py, written from src/assembly/EFIE.jl:429-684
    without the C++ oracle), Z[m, n] and Z[n, m]."""
    import json
    fx = json.load(open(os.path.join(GOLD, "cells_mp.json")))
    m = dm.TriMesh(np.array(fx["xyz"], float), np.array(fx["tri"], np.int64))
    a = fx["rwg"]
    r = dm.rwg_from_arrays(m, np.array(a["tplus"]), np.array(a["tminus"]), np.array(a["vplus_opp"]),
                           np.array(a["vminus_opp"]), np.array(a["len"]), np.array(a["area_plus"]),
                           np.array(a["area_minus"]))
    c = fx["cases"][case]
    k = complex(*c["k"])
    Zd = dm.assemble_Z_efie(m, r, k if k.imag else k.real, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    Z64, Z32 = Zd.to_numpy(dm.DM_Z64), Zd.to_numpy(dm.DM_Z32)
    E = c["entries"]
    mm = np.array([e["m"] for e in E]); nn = np.array([e["n"] for e in E])
    want = np.array([complex(float(e["value"][0]), float(e["value"][1])) for e in E])
    for got64, got32 in ((Z64[mm, nn], Z32[mm, nn]), (Z64[nn, mm], Z32[nn, mm])):
        e64 = np.abs(got64 - want) / np.abs(want)
        e32 = np.linalg.norm(got32 - want) / np.linalg.norm(want)
        print(f"max entry rel fp64 {e64.max():.2e}, fp32 rel {e32:.2e}")
        assert e64.max() <= 1e-12 and e32 <= 1e-5


@pytest.mark.parametrize("quad_order", [1, 4, 7])
def test_quad_orders(dm, O, quad_order):
    m = dm.make_rect_plate(0.1, 0.1, 5, 5)
    r = dm.build_rwg(m)
    Z = dm.assemble_Z_efie(m, r, K0, quad_order=quad_order)
    assert rel_fro(Z, _oracle_Z(O, m, r, K0, quad_order)) <= 1e-10


def test_complex_wavenumber(dm, O):
    m = dm.make_rect_plate(0.1, 0.1, 5, 4)
    r = dm.build_rwg(m)
    k = K0 * (1 - 0.05j)
    Z = dm.assemble_Z_efie(m, r, k)
    assert rel_fro(Z, _oracle_Z(O, m, r, k)) <= 1e-10


@pytest.mark.parametrize("subdiv", [2, 3])
def test_sphere_full_matrix(dm, O, subdiv):
    m = dm.make_icosphere(0.1, subdiv)
    r = dm.build_rwg(m)
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    Zo = _oracle_Z(O, m, r, K0)
    e64 = rel_fro(Zd.to_numpy(dm.DM_Z64), Zo)
    e32 = rel_fro(Zd.to_numpy(dm.DM_Z32), Zo)
    print(f"sphere s={subdiv} N={r.nedges}: fp64 {e64:.2e} fp32 {e32:.2e}")
    assert e64 <= 1e-10 and e32 <= 1e-5


def test_plate_30x30_full_matrix(dm, O):
    m = dm.make_rect_plate(0.25, 0.25, 30, 30)
    r = dm.build_rwg(m)
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    Zo = _oracle_Z(O, m, r, K0)
    assert rel_fro(Zd.to_numpy(dm.DM_Z64), Zo) <= 1e-10
    assert rel_fro(Zd.to_numpy(dm.DM_Z32), Zo) <= 1e-5


def test_c1_plate_sampled_entries(dm, O):
    """The 0.5 m plate (60x60, N=10680): 2000 sampled entries including
    the diagonal and near entries, FP64 and complex64."""
    m = dm.make_rect_plate(0.5, 0.5, 60, 60)
    r = dm.build_rwg(m)
    assert r.nedges == 10680
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    rng = np.random.default_rng(11)
    M = rng.integers(0, r.nedges, 2000)
    Nn = np.concatenate([rng.integers(0, r.nedges, 1000), M[1000:1500], np.clip(M[1500:] + rng.integers(-3, 4, 500), 0, r.nedges - 1)])
    zo = O.EFIE(oracle_rwg(O, m, r), K0).entries(M, Nn)
    z64 = Zd.entries(M, Nn, dm.DM_Z64)
    z32 = Zd.entries(M, Nn, dm.DM_Z32)
    assert np.linalg.norm(z64 - zo) / np.linalg.norm(zo) <= 1e-10
    assert np.max(np.abs(z64 - zo) / np.abs(zo)) <= 1e-9
    assert np.linalg.norm(z32 - zo) / np.linalg.norm(zo) <= 1e-5


def test_products_and_matrix_free(dm, O):
    m = dm.make_rect_plate(0.12, 0.12, 10, 9)
    r = dm.build_rwg(m)
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    Z = Zd.to_numpy()
    rng = np.random.default_rng(5)
    X = rng.standard_normal((r.nedges, 3)) + 1j * rng.standard_normal((r.nedges, 3))
    Y = Zd.matmul(X)
    assert np.linalg.norm(Y - Z @ X) / np.linalg.norm(Z @ X) < 1e-12
    Ya = Zd.matmul(X, adjoint=True)
    assert np.linalg.norm(Ya - Z.conj().T @ X) / np.linalg.norm(Ya) < 1e-12
    Y32 = Zd.matmul(X, form=dm.DM_Z32)
    assert np.linalg.norm(Y32 - Z @ X) / np.linalg.norm(Z @ X) < 1e-5
    A = dm.matrixfree_efie_operator(m, r, K0)
    x = X[:, 0]
    y = A @ x
    assert np.linalg.norm(y - Z @ x) / np.linalg.norm(y) < 1e-10  # test/runtests.jl:8748
    ya = A.H @ x
    assert np.linalg.norm(ya - Z.conj().T @ x) / np.linalg.norm(ya) < 1e-10


@pytest.mark.parametrize("subdiv,nrhs", [(2, 8), (2, 70), (4, 72)])
def test_multi_rhs_products_on_matrix_cores(dm, subdiv, nrhs):
    """dm_apply with >= 8 right-hand sides runs Z X on the matrix cores
    (complex64: three-part bf16 split of Z and X, six part products per tile
    accumulated in FP32, FP64 sum of 4096-column chunks; complex128: FP64
    MFMA).  subdiv 4 (N = 7680) spans two K chunks; 70 and 72 leave partial
    64-column blocks.  The complex64 product is checked against the exact
    product of the complex64 entries to FP32 accuracy (a dropped part product
    would leave ~1e-5)."""
    m = dm.make_icosphere(0.1, subdiv)
    r = dm.build_rwg(m)
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32, device=True)
    Z = Zd.to_numpy()
    rng = np.random.default_rng(nrhs)
    X = rng.standard_normal((r.nedges, nrhs)) + 1j * rng.standard_normal((r.nedges, nrhs))
    ref = Z @ X
    Y64 = Zd.matmul(X, form=dm.DM_Z64)
    assert np.linalg.norm(Y64 - ref) / np.linalg.norm(ref) < 1e-12
    Y32 = Zd.matmul(X, form=dm.DM_Z32)
    assert np.linalg.norm(Y32 - ref) / np.linalg.norm(ref) < 1e-5
    Z32 = Zd.to_numpy(dm.DM_Z32).astype(np.complex128)
    ref32 = Z32 @ X.astype(np.complex64).astype(np.complex128)
    assert np.linalg.norm(Y32 - ref32) / np.linalg.norm(ref32) < 2e-6
    # every column, not only the aggregate, and against the one-vector GEMV path
    col_err = np.linalg.norm(Y32 - ref, axis=0) / np.linalg.norm(ref, axis=0)
    assert col_err.max() < 1e-5
    y_gemv = Zd.matmul(X[:, -1], form=dm.DM_Z32)
    assert np.linalg.norm(Y32[:, -1] - y_gemv) / np.linalg.norm(y_gemv) < 1e-5
    Ya = Zd.matmul(X, form=dm.DM_Z64, adjoint=True)
    refa = Z.conj().T @ X
    assert np.linalg.norm(Ya - refa) / np.linalg.norm(refa) < 1e-12


def test_excitation_parity(dm, O):
    for mk in (lambda: dm.make_rect_plate(0.1, 0.1, 6, 6), lambda: dm.make_icosphere(0.1, 2)):
        m = mk()
        r = dm.build_rwg(m)
        orr = oracle_rwg(O, m, r)
        th = math.radians(30)
        kv = K0 * np.array([math.sin(th), 0, math.cos(th)])
        pol = np.array([math.cos(th), 0, -math.sin(th)])
        v = dm.assemble_v_plane_wave(m, r, kv, 1.5, pol)
        vo = O.assemble_v_plane_wave(orr, kv, 1.5, pol)
        assert np.linalg.norm(v - vo) / np.linalg.norm(vo) < 1e-13
    with pytest.raises(ValueError):
        dm.assemble_v_plane_wave(m, r, kv, 1.0, [1.0, 0, 0.1])  # not unit / not transverse


def test_radiation_vectors_parity(dm, O):
    m = dm.make_icosphere(0.05, 2)
    r = dm.build_rwg(m)
    grid = dm.make_sph_grid(9, 8)
    G = dm.radiation_vectors(m, r, grid, K0)
    Go = O.radiation_vectors(oracle_rwg(O, m, r), grid.rhat, K0)
    assert np.linalg.norm(G - Go) / np.linalg.norm(Go) < 1e-13


def test_error_paths(dm):
    m = dm.make_rect_plate(0.1, 0.1, 3, 3)
    r = dm.build_rwg(m)
    with pytest.raises(ValueError):
        dm.assemble_Z_efie(m, r, float("nan"))
    with pytest.raises(ValueError):
        dm.assemble_Z_efie(m, r, K0, eta0=0.0)
    with pytest.raises(ValueError):
        dm.assemble_Z_efie(m, r, K0, max_output_bytes=r.nedges ** 2 * 16 - 1)
    with pytest.raises(ValueError):
        dm.assemble_Z_efie(m, r, K0, quad_order=5)
    other = dm.make_rect_plate(0.1, 0.1, 3, 3)
    with pytest.raises(ValueError):
        dm.assemble_Z_efie(other, r, K0)  # mesh/RWG pair mismatch (src/Types.jl:154-159)
    bad = dm.TriMesh(np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0.0]]), np.array([[1, 2, 3]]))
    with pytest.raises(ValueError):
        dm.build_rwg(bad)  # degenerate triangle fails the precheck


@pytest.mark.parametrize("mk", ["plate", "sphere"])
def test_packed_symmetric_storage_and_product(dm, mk):
    """Packed symmetric forms hold the entries of the full forms (upper triangle
    of 128 x 128 tiles; separate assemblies agree to the rounding of the tile
    accumulation, whose LDS atomics have no fixed order); the symmetric product
    (one read per stored tile, used for Z x and Z^T x) matches the dense product
    of the same stored entries to FP64 rounding."""
    m = dm.make_rect_plate(0.3, 0.25, 30, 22) if mk == "plate" else dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    allf = dm.DM_Z64 | dm.DM_Z32 | dm.DM_Z64P | dm.DM_Z32P
    Zd = dm.assemble_Z_efie(m, r, K0, forms=allf, device=True)
    Z64, Z32 = Zd.to_numpy(dm.DM_Z64), Zd.to_numpy(dm.DM_Z32)
    P64, P32 = Zd.to_numpy(dm.DM_Z64P), Zd.to_numpy(dm.DM_Z32P)
    assert np.array_equal(P64, P64.T) and np.array_equal(P32, P32.T)
    assert np.max(np.abs(P64 - Z64)) <= 1e-13 * np.max(np.abs(Z64))
    assert rel_fro(P64, Z64) <= 1e-14
    assert rel_fro(P32, Z32) <= 1e-6
    rng = np.random.default_rng(11)
    idx = rng.integers(0, r.nedges, (2, 500))
    assert np.array_equal(Zd.entries(idx[0], idx[1], dm.DM_Z64P), P64[idx[0], idx[1]])
    X = rng.standard_normal((r.nedges, 3)) + 1j * rng.standard_normal((r.nedges, 3))
    ref = P64 @ X
    Y = Zd.matmul(X, form=dm.DM_Z64P)
    assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-13
    assert np.max(np.abs(Y - ref)) <= 1e-12 * np.max(np.abs(ref))
    Y32 = Zd.matmul(X, form=dm.DM_Z32P)
    assert np.linalg.norm(Y32 - P32 @ X) / np.linalg.norm(ref) < 1e-13  # complex64 entries, FP64 sums
    assert np.linalg.norm(Y32 - ref) / np.linalg.norm(ref) < 1e-5
    Ya = Zd.matmul(X, form=dm.DM_Z64P, adjoint=True)
    assert np.linalg.norm(Ya - P64.conj().T @ X) / np.linalg.norm(ref) < 1e-13
    # diagonal, block and near-field preconditioners read packed storage too
    for fact, cut in (("diag", 0.0), ("block", 0.0), ("spai", 2.0 * float(np.mean(r.len)))):
        P = dm.build_nearfield_preconditioner(Zd, m, r, cut, factorization=fact, block_size=64)
        for f_full, f_packed, tol in ((dm.DM_Z64, dm.DM_Z64P, 1e-11), (dm.DM_Z32, dm.DM_Z32P, 1e-4)):
            x = X[:, 0]
            yf = dm.NearFieldOperator(P, Zd, f_full) @ x
            yp = dm.NearFieldOperator(P, Zd, f_packed) @ x
            assert np.linalg.norm(yp - yf) / np.linalg.norm(yf) < tol, (fact, f_packed)


def test_multi_mesh_persistent_assembly(dm, O):
    """One persistent launch over 12 seeded random plates (0.05-0.3 m sides,
    4-30 cells) and a sphere: each matrix equals its own single-mesh assembly
    (to the rounding of the tile accumulation order) and the oracle."""
    rng = np.random.default_rng(20261202)
    meshes = [dm.make_rect_plate(*rng.uniform(0.05, 0.3, 2), *map(int, rng.integers(4, 31, 2))) for _ in range(12)]
    meshes.append(dm.make_icosphere(0.1, 2))
    rwgs = [dm.build_rwg(m) for m in meshes]
    Zs = dm.assemble_Z_efie_multi(meshes, rwgs, K0, forms=dm.DM_Z64 | dm.DM_Z32P)
    for i, (m, r, Z) in enumerate(zip(meshes, rwgs, Zs)):
        single = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64 | dm.DM_Z32P, device=True)
        A, B = Z.to_numpy(dm.DM_Z64), single.to_numpy(dm.DM_Z64)
        assert rel_fro(A, B) <= 1e-14, i
        assert rel_fro(Z.to_numpy(dm.DM_Z32P), single.to_numpy(dm.DM_Z32P)) <= 1e-6, i
        if i in (0, 12):
            assert rel_fro(A, _oracle_Z(O, m, r, K0)) <= 1e-10


def test_64_seeded_plates_in_one_launch(dm, O):
    """The many-plate workload: 64 seeded plates (0.05-0.3 m sides, 4-30
    cells, src/geometry/Mesh.jl:309-396) from the many-angle generator stream,
    assembled in one persistent launch.  Every matrix equals its single-mesh
    assembly (to the rounding of the accumulation order) and 40 sampled
    entries of each (diagonal, near and far) match the oracle to 1e-10."""
    _, specs = dm.many_angle_workload(K0, 256, 64)
    meshes = [dm.make_rect_plate(*sp) for sp in specs]
    rwgs = [dm.build_rwg(m) for m in meshes]
    Ns = [r.nedges for r in rwgs]
    print(f"64 plates, N {min(Ns)}..{max(Ns)}, total {sum(Ns)}", flush=True)
    Zs = dm.assemble_Z_efie_multi(meshes, rwgs, K0, forms=dm.DM_Z64)
    rng = np.random.default_rng(64)
    worst = 0.0
    for i, (m, r, Z) in enumerate(zip(meshes, rwgs, Zs)):
        single = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64, device=True)
        n = r.nedges
        M = rng.integers(0, n, 40)
        Nn = np.concatenate([M[:10], np.clip(M[10:25] + rng.integers(-2, 3, 15), 0, n - 1), rng.integers(0, n, 15)])
        zm = Z.entries(M, Nn, dm.DM_Z64)
        assert np.array_equal(zm, single.entries(M, Nn, dm.DM_Z64)) or \
            np.abs(zm - single.entries(M, Nn, dm.DM_Z64)).max() <= 1e-14 * np.abs(zm).max(), i
        zo = O.EFIE(oracle_rwg(O, m, r), K0).entries(M, Nn)
        err = np.linalg.norm(zm - zo) / np.linalg.norm(zo)
        worst = max(worst, err)
        assert err <= 1e-10, (i, specs[i], err)
        single.free()
        Z.free()
    print(f"worst sampled-entry error vs oracle {worst:.2e}", flush=True)


@pytest.mark.parametrize("mk", ["plate", "sphere"])
def test_tuned_symmetric_product_shapes(dm, mk):
    """The tuner times the symmetric product's launch shapes on first use and
    records the choice.  Every allowed shape gives the same product (FP64 sums
    in another order): the LDS-staged form with 2..64 tiles per workgroup and
    the wave form (1000 + tiles per wave, runs of tiles crossing tile columns);
    shapes loaded with tune_set are used as is, out-of-range ones are refused."""
    m = dm.make_rect_plate(0.3, 0.25, 30, 22) if mk == "plate" else dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P, device=True)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((r.nedges, 2)) + 1j * rng.standard_normal((r.nedges, 2))
    ctx = dm.default_context()
    for form, tag in ((dm.DM_Z64P, "c128"), (dm.DM_Z32P, "c64")):
        P = Zd.to_numpy(form).astype(np.complex128)
        ref = P @ X
        Zd.matmul(X, form=form)
        key = f"symv/{tag}/{r.nedges}"
        got = ctx.tune_get()
        assert 2 <= got[key] <= 64 or 1001 <= got[key] <= 2024
        for shape in (2, 3, 16, 64, 1001, 1002, 1003, 1007, 1100):
            ctx.tune_set({key: shape})
            assert ctx.tune_get()[key] == shape
            Y = Zd.matmul(X, form=form)
            assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 1e-13, (tag, shape)
        for bad in (1, 65, 1000, 2025):
            ctx.tune_set({key: bad})
            with pytest.raises(Exception):
                Zd.matmul(X, form=form)
        ctx.tune_set(got)


def test_tuned_assembly_threads_and_zx_chunks(dm):
    """Tuner keys of the FP64 tile kernel's threads per workgroup
    (asm/<form>/q<points>/<N>: 1024 or 512) and of the bf16 Z X's FP32
    accumulation chunk (zx/c64/<N>/<R>: 4096 or 8192 columns): recorded on
    first use and exported by tune_get; every allowed value gives the same
    matrix (FP64 sums in another order) or the same product to FP32 accuracy;
    values outside the candidates are refused."""
    ctx = dm.default_context()
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    forms = dm.DM_Z64P | dm.DM_Z32P | dm.DM_Z32_ROUNDED
    Z = dm.assemble_Z_efie(m, r, K0, forms=forms, device=True)
    key = f"asm/z64p+z32/q3/{r.nedges}"
    got = ctx.tune_get()
    assert got[key] in (512, 1024)
    P64 = Z.to_numpy(dm.DM_Z64P)
    for t in (512, 1024):
        ctx.tune_set({key: t})
        Zt = dm.assemble_Z_efie(m, r, K0, forms=forms, device=True)
        Pt = Zt.to_numpy(dm.DM_Z64P)
        assert np.linalg.norm(Pt - P64) <= 1e-14 * np.linalg.norm(P64), t
        assert np.array_equal(Zt.to_numpy(dm.DM_Z32P), Pt.astype(np.complex64).astype(np.complex128))
        Zt.free()
    ctx.tune_set({key: 256})
    with pytest.raises(Exception):
        dm.assemble_Z_efie(m, r, K0, forms=forms, device=True)
    ctx.tune_set(got)
    # Z X on a matrix wider than one 4096-column chunk
    ms = dm.make_icosphere(0.1, 4)
    rs = dm.build_rwg(ms)
    Zs = dm.assemble_Z_efie(ms, rs, K0, forms=dm.DM_Z64 | dm.DM_Z32 | dm.DM_Z32_ROUNDED, device=True)
    rng = np.random.default_rng(13)
    X = rng.standard_normal((rs.nedges, 16)) + 1j * rng.standard_normal((rs.nedges, 16))
    ref = Zs.matmul(X, form=dm.DM_Z64)
    Y = Zs.matmul(X, form=dm.DM_Z32)
    zkey = f"zx/c64/{rs.nedges}/16"
    got = ctx.tune_get()
    assert got[zkey] in (4096, 8192)
    for kc in (4096, 8192):
        ctx.tune_set({zkey: kc})
        Yk = Zs.matmul(X, form=dm.DM_Z32)
        err = np.linalg.norm(Yk - ref) / np.linalg.norm(ref)
        print(f"Z X chunk {kc}: error vs complex128 product {err:.2e}")
        assert err <= 1e-6, kc
    assert np.linalg.norm(Y - ref) <= 1e-6 * np.linalg.norm(ref)
    ctx.tune_set({zkey: 2048})
    with pytest.raises(Exception):
        Zs.matmul(X, form=dm.DM_Z32)
    ctx.tune_set(got)
    Zs.free()


@pytest.mark.parametrize("mk", ["plate", "sphere"])
def test_compressed_form_error_bound(dm, mk):
    """dm_zmat_compress: off-diagonal stored tiles become cross approximations
    U V^T with ||Z32P - Z32H||_F <= eps ||Z32P||_F up to FP32 rounding of the
    updates (checked on the whole matrix, recovered column by column through
    the compressed product), the diagonal tiles stay dense, and the product
    matches the dense product of the same complex64 entries to that bound."""
    m = dm.make_rect_plate(0.3, 0.25, 30, 22) if mk == "plate" else dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    N = r.nedges
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P | dm.DM_Z32_ROUNDED, device=True)
    assert Zd.forms == dm.DM_Z64P | dm.DM_Z32P
    P = Zd.to_numpy(dm.DM_Z32P).astype(np.complex128)
    assert np.array_equal(P, Zd.to_numpy(dm.DM_Z64P).astype(np.complex64))
    for eps in (1e-3, 1e-6):
        st = Zd.compress(eps=eps, rmax=48)
        nb = -(-N // 128)
        assert st["tiles"] == nb * (nb + 1) // 2
        assert 0 < st["lowrank_tiles"] <= st["tiles"] - nb
        assert 0 < st["bytes_ratio"] < 1
        H = Zd.matmul(np.eye(N, dtype=np.complex128), form=dm.DM_Z32H)
        err = np.linalg.norm(H - P) / np.linalg.norm(P)
        print(f"{mk} N={N} eps {eps}: {st}, error {err:.2e}")
        # the bound holds for the tracked FP32 residual; rounding of the
        # FP32 updates adds a few 1e-8 ||Z||_F at most
        assert err <= eps + 1e-7
        assert err > 0 or st["lowrank_tiles"] == 0
        # Z_IJ ~ U V^T is used as Z_JI ~ V U^T: the product stays symmetric
        assert np.linalg.norm(H - H.T) / np.linalg.norm(P) <= 1e-12
    x = np.random.default_rng(2).standard_normal(N) + 0j
    assert np.linalg.norm(Zd.matmul(x, form=dm.DM_Z32H) - P @ x) / np.linalg.norm(P @ x) < 1e-5
    # the FP32-quadrature matrix carries quadrature noise above FP32 rounding,
    # which the cross approximation cannot drop: it compresses no better
    Zq = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z32P, device=True)
    Zr = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P | dm.DM_Z32_ROUNDED, device=True)
    sq, sr = Zq.compress(eps=1e-6), Zr.compress(eps=1e-6)
    print(f"eps 1e-6: FP32 quadrature {sq}, rounded FP64 {sr}")
    assert sr["bytes_ratio"] <= sq["bytes_ratio"]
    with pytest.raises(ValueError):
        dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z32P | dm.DM_Z32_ROUNDED, device=True)
    with pytest.raises(Exception):
        Zd.compress(eps=0.0)
    with pytest.raises(Exception):
        Zd.compress(eps=1e-6, rmax=65)


def test_compressed_form_levels(dm):
    """The looser levels of one cross approximation (dm_zmat_hlevel): level l
    reads the prefix of each tile's factors at 10^(l/2) times the tolerance,
    so its whole-matrix error is <= eps 10^(l/2) (up to FP32 rounding of the
    updates); ranks and bytes do not grow with the level; level 0 is the
    compression itself."""
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    N = r.nedges
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P | dm.DM_Z32_ROUNDED, device=True)
    P = Zd.to_numpy(dm.DM_Z32P).astype(np.complex128)
    st0 = Zd.compress(eps=1e-6, rmax=48)
    prev = None
    for lv in range(11):
        st = Zd.compress_level(lv)
        H = Zd.matmul(np.eye(N, dtype=np.complex128), form=dm.DM_Z32H)
        err = np.linalg.norm(H - P) / np.linalg.norm(P)
        print(f"level {lv}: eps {st['eps']:.0e}, low-rank tiles {st['lowrank_tiles']}, mean rank "
              f"{st['mean_rank']:.2f}, bytes {st['bytes_ratio']:.3f}, error {err:.2e}")
        assert abs(st["eps"] - 1e-6 * 10.0**(lv / 2)) <= 1e-12 * st["eps"]
        assert err <= st["eps"] + 1e-7, lv
        if lv == 0:
            assert st["bytes_ratio"] == st0["bytes_ratio"] and st["lowrank_tiles"] == st0["lowrank_tiles"]
        else:
            assert st["bytes_ratio"] <= prev["bytes_ratio"] and st["lowrank_tiles"] >= prev["lowrank_tiles"]
        prev = st
    with pytest.raises(Exception):
        Zd.compress_level(11)
    Zd.compress_level(0)


def test_compressed_product_shapes(dm):
    """Every wave count of the compressed product (a multiple of the resident
    wave slots; the waves take tile ranges of about equal bytes) gives the
    same result (FP64 sums in another order); out-of-range multiples are
    refused."""
    m = dm.make_icosphere(0.1, 3)
    r = dm.build_rwg(m)
    Zd = dm.assemble_Z_efie(m, r, K0, forms=dm.DM_Z64P | dm.DM_Z32P | dm.DM_Z32_ROUNDED, device=True)
    Zd.compress(eps=1e-4)
    rng = np.random.default_rng(9)
    x = rng.standard_normal(r.nedges) + 1j * rng.standard_normal(r.nedges)
    ctx = dm.default_context()
    ref = Zd.matmul(x, form=dm.DM_Z32H)
    key = f"symvh/c64/{r.nedges}"
    got = ctx.tune_get()
    assert got[key] in (1, 2, 4, 8)
    for mult in (1, 2, 3, 5, 64):
        ctx.tune_set({key: mult})
        y = Zd.matmul(x, form=dm.DM_Z32H)
        assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13, mult
    for bad in (0, 65):
        ctx.tune_set({key: bad})
        with pytest.raises(Exception):
            Zd.matmul(x, form=dm.DM_Z32H)
    ctx.tune_set(got)


def test_farfield_of_many_currents(dm, O):
    """dm_farfield: E = G I for a batch of currents on the FP64 matrix cores
    (compute_farfield, src/postprocessing/FarField.jl:584-605), against the
    oracle's radiation vectors; a 12000-direction grid on N = 1920 spans two
    direction chunks, checked on a spread subset of directions."""
    m = dm.make_icosphere(0.05, 2)
    r = dm.build_rwg(m)
    rng = np.random.default_rng(4)
    grid = dm.make_sph_grid(9, 8)
    Go = O.radiation_vectors(oracle_rwg(O, m, r), grid.rhat, K0)
    I = rng.standard_normal((r.nedges, 70)) + 1j * rng.standard_normal((r.nedges, 70))
    E = dm.farfield(m, r, grid, K0, I)
    assert E.shape == (3, 72, 70)
    ref = (Go @ I).reshape((3, 72, 70), order="F")
    assert np.linalg.norm(E - ref) / np.linalg.norm(ref) < 1e-13
    e1 = dm.farfield(m, r, grid, K0, I[:, 5])
    assert np.linalg.norm(e1 - dm.compute_farfield(Go, I[:, 5], 72)) / np.linalg.norm(e1) < 1e-13
    m2 = dm.make_icosphere(0.1, 3)
    r2 = dm.build_rwg(m2)
    big = dm.make_sph_grid(120, 100)
    I2 = rng.standard_normal((r2.nedges, 9)) + 1j * rng.standard_normal((r2.nedges, 9))
    E2 = dm.farfield(m2, r2, big, K0, I2)
    S = np.unique(np.concatenate([np.arange(0, 12000, 97), [11999, 11650, 11651]]))
    G2 = dm.radiation_vectors(m2, r2, big.rhat[S], K0)
    ref2 = (G2 @ I2).reshape((3, len(S), 9), order="F")
    assert np.linalg.norm(E2[:, S, :] - ref2) / np.linalg.norm(ref2) < 1e-13
    with pytest.raises(ValueError):
        dm.farfield(m, r, np.zeros((2, 3)), K0, I)
