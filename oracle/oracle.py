"""CPU oracle for the surface-EFIE path (TEST INFRASTRUCTURE ONLY).

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it, and only to compare against or
time beside the HIP library.  The product library never calls it.

Two layers:
* ``liboracle.so`` (``oracle.cpp``): FP64 restatement of the reference's Green
  functions, quadrature, singular integrals, entry-wise EFIE assembly, plane-wave
  RHS and radiation vectors (citations in that file).
* this file: ctypes glue plus numpy/scipy restatements of the host-side solver
  pieces:
    - dense direct solve            src/solver/Solve.jl:1124-1198 (LAPACK getrf/getrs)
    - GMRES (Krylov.jl 0.9 semantics: full, MGS + Givens, left/right preconditioner,
      stop on ||M r|| <= atol + rtol ||M b||)   src/solver/IterativeSolve.jl:341-392
      Krylov.jl itself is not vendored in the reference; this follows its published
      algorithm.  Iteration counts are "parity unpinned".
    - near-field preconditioner     src/solver/NearFieldPreconditioner.jl:247-464,670-717
      (UMFPACK sparse LU replaced by SuperLU from scipy: same Z_nf^{-1}).
    - rwg_centers                   src/solver/NearFieldPreconditioner.jl:247-257
    - make_sph_grid                 src/postprocessing/FarField.jl:216-277
    - compute_farfield / bistatic_rcs src/postprocessing/FarField.jl:584-605,
                                    src/postprocessing/Diagnostics.jl:144-180,460-473
    - Mie PEC sphere RCS (own series; the reference's src/postprocessing/Mie.jl:1473
      is used there only as a physics cross-check).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    """Load liboracle.so, building it with make when it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    L = C.CDLL(path)
    vp = C.c_void_p
    L.or_make_rect_plate.argtypes = [C.c_double, C.c_double, C.c_int64, C.c_int64, f64p, i64p]
    L.or_make_icosphere.argtypes = [C.c_double, C.c_int, f64p, i64p]
    L.or_triangle_area.argtypes = [f64p, i64p, C.c_int64]
    L.or_triangle_area.restype = C.c_double
    L.or_build_rwg.argtypes = [f64p, C.c_int64, i64p] + [vp] * 8
    L.or_build_rwg.restype = C.c_int64
    L.or_adjacent_pair_count.argtypes = [C.c_int64, f64p, C.c_int64, i64p]
    L.or_adjacent_pair_count.restype = C.c_int64
    L.or_quad_rule.argtypes = [C.c_int, f64p, f64p]
    L.or_quad_rule.restype = C.c_int
    L.or_greens.argtypes = [f64p, f64p, C.c_double, C.c_double, f64p]
    L.or_greens_smooth.argtypes = [f64p, f64p, C.c_double, C.c_double, f64p]
    L.or_int_1overR.argtypes = [f64p, f64p, f64p, f64p]
    L.or_int_1overR.restype = C.c_double
    rwg_args = [C.c_int64, f64p, C.c_int64, i64p, C.c_int64, i64p, i64p, i64p, i64p,
                f64p, f64p, f64p, f64p, f64p]
    L.or_efie_create.argtypes = rwg_args + [C.c_double, C.c_double, C.c_double, C.c_int]
    L.or_efie_create.restype = vp
    L.or_efie_free.argtypes = [vp]
    L.or_efie_entries.argtypes = [vp, i64p, i64p, C.c_int64, f64p]
    i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    L.or_pair_terms.argtypes = [vp, i64p, i32p, i64p, i32p, C.c_int64, f64p]
    L.or_assemble_z_rows.argtypes = [vp, C.c_int64, C.c_int64, f64p, C.c_int64]
    L.or_mf_apply.argtypes = [vp, f64p, f64p]
    L.or_assemble_rows_sum.argtypes = [vp, C.c_int64, C.c_int64, f64p]
    L.or_assemble_rows_sum.restype = C.c_int64
    L.or_plane_wave.argtypes = rwg_args + [f64p, C.c_double, f64p, C.c_int, f64p]
    L.or_radiation_vectors.argtypes = rwg_args + [f64p, C.c_int64, C.c_double, C.c_double,
                                                  C.c_int, f64p]
    _LIB = L
    return L


# ----------------------------------------------------------------------------- mesh
class Mesh:
    """xyz: (Nv, 3) float64 (row i = vertex i+1); tri: (Nt, 3) int64, 1-based."""

    def __init__(self, xyz, tri):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int64)

    @property
    def nv(self):
        return self.xyz.shape[0]

    @property
    def nt(self):
        return self.tri.shape[0]


def make_rect_plate(Lx, Ly, Nx, Ny):
    nv, nt = (Nx + 1) * (Ny + 1), 2 * Nx * Ny
    xyz = np.zeros((nv, 3)); tri = np.zeros((nt, 3), dtype=np.int64)
    lib().or_make_rect_plate(float(Lx), float(Ly), int(Nx), int(Ny), xyz, tri)
    return Mesh(xyz, tri)


def make_icosphere(radius, subdivisions):
    nv, nt = 10 * 4 ** subdivisions + 2, 20 * 4 ** subdivisions
    xyz = np.zeros((nv, 3)); tri = np.zeros((nt, 3), dtype=np.int64)
    lib().or_make_icosphere(float(radius), int(subdivisions), xyz, tri)
    return Mesh(xyz, tri)


def triangle_area(mesh, t1):
    return lib().or_triangle_area(mesh.xyz, mesh.tri, int(t1))


class RWG:
    """Arrays mirror RWGData (src/Types.jl:138-152); indices 1-based."""

    def __init__(self, mesh, tplus, tminus, evert, vplus_opp, vminus_opp, length, area_plus,
                 area_minus, coeff_plus=None, coeff_minus=None):
        self.mesh = mesh
        self.tplus = np.ascontiguousarray(tplus, dtype=np.int64)
        self.tminus = np.ascontiguousarray(tminus, dtype=np.int64)
        self.evert = np.ascontiguousarray(evert, dtype=np.int64)
        self.vplus_opp = np.ascontiguousarray(vplus_opp, dtype=np.int64)
        self.vminus_opp = np.ascontiguousarray(vminus_opp, dtype=np.int64)
        self.len = np.ascontiguousarray(length, dtype=np.float64)
        self.area_plus = np.ascontiguousarray(area_plus, dtype=np.float64)
        self.area_minus = np.ascontiguousarray(area_minus, dtype=np.float64)
        n = len(self.tplus)
        self.coeff_plus = np.ones(n) if coeff_plus is None else np.ascontiguousarray(coeff_plus, dtype=np.float64)
        self.coeff_minus = np.ones(n) if coeff_minus is None else np.ascontiguousarray(coeff_minus, dtype=np.float64)

    @property
    def nedges(self):
        return len(self.tplus)

    def permuted(self, perm):
        """Basis functions reordered: new[i] = old[perm[i]] (tests RWG-order independence)."""
        p = np.asarray(perm)
        return RWG(self.mesh, self.tplus[p], self.tminus[p], self.evert[p], self.vplus_opp[p],
                   self.vminus_opp[p], self.len[p], self.area_plus[p], self.area_minus[p],
                   self.coeff_plus[p], self.coeff_minus[p])

    def args(self):
        m = self.mesh
        return (m.nv, m.xyz, m.nt, m.tri, self.nedges, self.tplus, self.tminus, self.vplus_opp,
                self.vminus_opp, self.len, self.area_plus, self.area_minus, self.coeff_plus,
                self.coeff_minus)


def build_rwg(mesh):
    L = lib()
    n = L.or_build_rwg(mesh.xyz, mesh.nt, mesh.tri, *([None] * 8))
    tp = np.zeros(n, np.int64); tm = np.zeros(n, np.int64); ev = np.zeros((n, 2), np.int64)
    vp = np.zeros(n, np.int64); vm = np.zeros(n, np.int64)
    ln = np.zeros(n); ap = np.zeros(n); am = np.zeros(n)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    L.or_build_rwg(mesh.xyz, mesh.nt, mesh.tri, ptr(tp), ptr(tm), ptr(ev), ptr(vp), ptr(vm),
                   ptr(ln), ptr(ap), ptr(am))
    return RWG(mesh, tp, tm, ev, vp, vm, ln, ap, am)


def adjacent_pair_count(mesh):
    return lib().or_adjacent_pair_count(mesh.nv, mesh.xyz, mesh.nt, mesh.tri)


# ------------------------------------------------------------------ scalar kernels
def quad_rule(order):
    xi = np.zeros(14); w = np.zeros(7)
    nq = lib().or_quad_rule(int(order), xi, w)
    if nq == 0:
        raise ValueError(f"Unsupported quadrature order {order}")
    return xi[: 2 * nq].reshape(nq, 2), w[:nq]


def greens(r, rp, k):
    out = np.zeros(2); k = complex(k)
    lib().or_greens(np.asarray(r, float), np.asarray(rp, float), k.real, k.imag, out)
    return complex(out[0], out[1])


def greens_smooth(r, rp, k):
    out = np.zeros(2); k = complex(k)
    lib().or_greens_smooth(np.asarray(r, float), np.asarray(rp, float), k.real, k.imag, out)
    return complex(out[0], out[1])


def int_1overR(P, V1, V2, V3):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (P, V1, V2, V3)]
    return lib().or_int_1overR(*a)


# --------------------------------------------------------------------------- EFIE
class EFIE:
    """Entry-wise FP64 EFIE operator (src/assembly/EFIE.jl:332-753,867-909)."""

    def __init__(self, rwg, k, eta0=376.730313668, quad_order=3):
        k = complex(k)
        self.rwg = rwg
        self.N = rwg.nedges
        self.h = lib().or_efie_create(*rwg.args(), k.real, k.imag, float(eta0), int(quad_order))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_efie_free(self.h)
            self.h = None

    def entries(self, m, n):
        m = np.ascontiguousarray(m, dtype=np.int64); n = np.ascontiguousarray(n, dtype=np.int64)
        out = np.zeros(2 * len(m))
        lib().or_efie_entries(self.h, m, n, len(m), out)
        return out.view(np.complex128)

    def pair_terms(self, m, sm, n, sn):
        """Triangle-pair terms of _efie_entry before -i k eta0
        (src/assembly/EFIE.jl:453-508): side sm of basis m (0 = T+, 1 = T-) is
        the outer triangle, side sn of basis n the inner one (0-based bases)."""
        m = np.ascontiguousarray(m, dtype=np.int64); n = np.ascontiguousarray(n, dtype=np.int64)
        sm = np.ascontiguousarray(sm, dtype=np.int32); sn = np.ascontiguousarray(sn, dtype=np.int32)
        out = np.zeros(2 * len(m))
        lib().or_pair_terms(self.h, m, sm, n, sn, len(m), out)
        return out.view(np.complex128)

    def assemble(self, row0=0, row1=None):
        """Dense Z (complex128, N x N). With a row range, only rows [row0,row1)
        and their mirrored columns inside the block are filled."""
        N = self.N
        row1 = N if row1 is None else row1
        Z = np.zeros((N, N), dtype=np.complex128, order="F")
        buf = Z.reshape(-1, order="F").view(np.float64)
        lib().or_assemble_z_rows(self.h, int(row0), int(row1), buf, N)
        return Z

    def rows_sum(self, row0, row1):
        """Evaluate entries (m, n >= m) of rows [row0, row1) without storing them;
        returns (entry count, checksum).  Used to time the CPU baseline."""
        cs = np.zeros(1)
        cnt = lib().or_assemble_rows_sum(self.h, int(row0), int(row1), cs)
        return cnt, float(cs[0])

    def mf_apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.N, dtype=np.complex128)
        lib().or_mf_apply(self.h, x.view(np.float64), y.view(np.float64))
        return y


def assemble_Z_efie(rwg, k, quad_order=3, eta0=376.730313668):
    return EFIE(rwg, k, eta0, quad_order).assemble()


def assemble_v_plane_wave(rwg, k_vec, E0, pol, quad_order=3):
    v = np.zeros(2 * rwg.nedges)
    lib().or_plane_wave(*rwg.args(), np.asarray(k_vec, float), float(E0), np.asarray(pol, float),
                        int(quad_order), v)
    return v.view(np.complex128)


# ---------------------------------------------------------------------- far field
def make_sph_grid(Ntheta, Nphi):
    dth = math.pi / Ntheta; dph = 2 * math.pi / Nphi
    it = np.arange(1, Ntheta + 1); ip = np.arange(1, Nphi + 1)
    th = np.repeat((it - 0.5) * dth, Nphi); ph = np.tile((ip - 0.5) * dph, Ntheta)
    rhat = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    return dict(rhat=np.ascontiguousarray(rhat), theta=th, phi=ph, w=np.sin(th) * dth * dph)


def radiation_vectors(rwg, rhat, k, eta0=376.730313668, quad_order=3):
    rhat = np.ascontiguousarray(rhat, dtype=np.float64)
    no = rhat.shape[0]
    G = np.zeros(2 * 3 * no * rwg.nedges)
    lib().or_radiation_vectors(*rwg.args(), rhat, no, float(k), float(eta0), int(quad_order), G)
    return G.view(np.complex128).reshape((3 * no, rwg.nedges), order="F")


def compute_farfield(G, I):
    return (G @ I).reshape((3, -1), order="F")


def bistatic_rcs(E, E0=1.0):
    return 4 * math.pi * np.sum(np.abs(E) ** 2, axis=0) / abs(E0) ** 2


# --------------------------------------------------------------------- solvers
def solve_direct(Z, v):
    """Dense LU solve (LAPACK zgesv through numpy): src/solver/Solve.jl:1124-1198."""
    return np.linalg.solve(Z, v)


def gmres(A, b, M=None, Nr=None, rtol=1e-8, atol=0.0, itmax=200):
    """Full GMRES with modified Gram-Schmidt and Givens rotations, as Krylov.jl's
    gmres (restart=false) used by src/solver/IterativeSolve.jl:364-384.
    A, M (left) and Nr (right) are callables.  Returns (x, niter, history, solved)."""
    n = len(b)
    Mf = M if M is not None else (lambda v: v)
    Nf = Nr if Nr is not None else (lambda v: v)
    r0 = Mf(np.asarray(b, dtype=np.complex128))
    beta = np.linalg.norm(r0)
    hist = [beta]
    if beta == 0:
        return np.zeros(n, complex), 0, hist, True
    eps = atol + rtol * beta
    V = [r0 / beta]
    H = np.zeros((itmax + 1, itmax), complex)
    cs = np.zeros(itmax); sn = np.zeros(itmax, complex)
    z = np.zeros(itmax + 1, complex); z[0] = beta
    k = 0; solved = False
    while k < itmax:
        w = Mf(A(Nf(V[k])))
        for i in range(k + 1):
            H[i, k] = np.vdot(V[i], w)
            w = w - H[i, k] * V[i]
        hnext = np.linalg.norm(w)
        H[k + 1, k] = hnext
        for i in range(k):
            t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
            H[i + 1, k] = -np.conj(sn[i]) * H[i, k] + cs[i] * H[i + 1, k]
            H[i, k] = t
        a, bb = H[k, k], H[k + 1, k]
        rho = math.hypot(abs(a), abs(bb))
        if abs(a) == 0:
            cs[k] = 0.0
            sn[k] = np.conj(bb) / abs(bb) if abs(bb) else 1.0
        else:
            cs[k] = abs(a) / rho
            sn[k] = (a / abs(a)) * np.conj(bb) / rho
        H[k, k] = cs[k] * a + sn[k] * bb
        H[k + 1, k] = 0
        z[k + 1] = -np.conj(sn[k]) * z[k]
        z[k] = cs[k] * z[k]
        k += 1
        hist.append(abs(z[k]))
        if abs(z[k]) <= eps:
            solved = True
            break
        if hnext == 0:
            break
        V.append(w / hnext)
    y = np.linalg.solve(np.triu(H[:k, :k]), z[:k])
    x = Nf(np.array(V[:k]).T @ y)
    return x, k, hist, solved


def true_residual(Z, x, b):
    """||b - Z x|| / ||b||: src/solver/IterativeSolve.jl:278-302."""
    return np.linalg.norm(Z @ x - b) / np.linalg.norm(b)


def rwg_centers(rwg):
    m = rwg.mesh
    cen = m.xyz[m.tri - 1].mean(axis=1)  # (Nt, 3) centroids
    return 0.5 * (cen[rwg.tplus - 1] + cen[rwg.tminus - 1])


def nearfield_pairs(centers, cutoff):
    """Spatial-hash enumeration of (m, n) with |c_m - c_n| <= cutoff (or m == n):
    src/solver/NearFieldPreconditioner.jl:259-338,408-427."""
    N = len(centers)
    if cutoff <= 0:
        i = np.arange(N)
        return i, i
    origin = centers.min(axis=0)
    keys = np.floor((centers - origin) / cutoff).astype(np.int64)
    buckets = {}
    for i, k in enumerate(map(tuple, keys)):
        buckets.setdefault(k, []).append(i)
    I, J = [], []
    for m in range(N):
        k = keys[m]
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    lst = buckets.get((k[0] + dx, k[1] + dy, k[2] + dz))
                    if not lst:
                        continue
                    lst = np.asarray(lst)
                    d = np.sqrt(((centers[lst] - centers[m]) ** 2).sum(axis=1))
                    sel = lst[(d <= cutoff) | (lst == m)]
                    I.extend([m] * len(sel)); J.extend(sel.tolist())
    return np.asarray(I), np.asarray(J)


def nearfield_pairs_tree(centers, cutoff):
    """The same (m, n) set as nearfield_pairs, from a k-d tree (scipy, compiled):
    for timing the near-field build at full size."""
    from scipy.spatial import cKDTree
    N = len(centers)
    if cutoff <= 0:
        i = np.arange(N)
        return i, i
    P = cKDTree(centers).query_pairs(cutoff, output_type="ndarray")
    i = np.arange(N)
    I = np.concatenate([i, P[:, 0], P[:, 1]])
    J = np.concatenate([i, P[:, 1], P[:, 0]])
    return I, J


class NearFieldLU:
    """Sparse near-field LU preconditioner: src/solver/NearFieldPreconditioner.jl:670-717.
    Z: the dense matrix, or an EFIE whose entries are computed on demand (the
    reference's matrix-free build, :847-901)."""

    def __init__(self, Z, rwg, cutoff):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spl
        c = rwg_centers(rwg)
        I, J = nearfield_pairs_tree(c, cutoff) if isinstance(Z, EFIE) else nearfield_pairs(c, cutoff)
        N = rwg.nedges
        vals = Z.entries(I, J) if isinstance(Z, EFIE) else Z[I, J]
        A = sp.csc_matrix((vals, (I, J)), shape=(N, N))
        self.nnz_ratio = A.nnz / N / N
        self.lu = spl.splu(A)

    def __call__(self, v):
        return self.lu.solve(np.asarray(v, dtype=np.complex128))


# ------------------------------------------------------------------------- Mie
def mie_bistatic_rcs_pec(k, a, khat, pol, rhat, nmax=None):
    """PEC-sphere bistatic RCS from the Mie series (Bohren & Huffman S1/S2 with
    PEC coefficients a_n = psi_n'/xi_n', b_n = psi_n/xi_n)."""
    from scipy.special import spherical_jn, spherical_yn
    x = k * a
    if nmax is None:
        nmax = int(x + 4 * x ** (1 / 3) + 2) + 5
    n = np.arange(1, nmax + 1)
    jn = spherical_jn(n, x); jnp_ = spherical_jn(n, x, derivative=True)
    yn = spherical_yn(n, x); ynp_ = spherical_yn(n, x, derivative=True)
    hn = jn + 1j * yn; hnp = jnp_ + 1j * ynp_
    psi = x * jn; dpsi = jn + x * jnp_
    xi = x * hn; dxi = hn + x * hnp
    an = dpsi / dxi; bn = psi / xi
    khat = np.asarray(khat, float) / np.linalg.norm(khat)
    pol = np.asarray(pol, float)
    rhat = np.asarray(rhat, float) / np.linalg.norm(rhat)
    mu = float(np.clip(np.dot(khat, rhat), -1, 1))
    # angular functions pi_n, tau_n by recurrence
    pi = np.zeros(nmax + 1); tau = np.zeros(nmax + 1)
    pi[1] = 1.0
    if nmax >= 2:
        pi[2] = 3 * mu
    for m in range(3, nmax + 1):
        pi[m] = ((2 * m - 1) / (m - 1)) * mu * pi[m - 1] - (m / (m - 1)) * pi[m - 2]
    for m in range(1, nmax + 1):
        tau[m] = m * mu * pi[m] - (m + 1) * pi[m - 1]
    f = (2 * n + 1) / (n * (n + 1))
    S1 = np.sum(f * (an * pi[1:] + bn * tau[1:]))
    S2 = np.sum(f * (an * tau[1:] + bn * pi[1:]))
    # azimuth between the scattering plane and the incident polarisation
    perp = rhat - mu * khat
    pn = np.linalg.norm(perp)
    cphi2 = 1.0 if pn < 1e-12 else (np.dot(perp, pol) / pn) ** 2
    sphi2 = 1.0 - cphi2
    return 4 * math.pi / k ** 2 * (abs(S2) ** 2 * cphi2 + abs(S1) ** 2 * sphi2)
