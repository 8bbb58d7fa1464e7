"""Python host mirror of DiffMoM.jl's surface-EFIE API over libdiffmom (HIP, gfx950).

Names, argument meaning and errors follow the Julia functions they mirror
(citations are path:line in DiffMoM.jl):

    make_rect_plate            src/geometry/Mesh.jl:309-396
    make_icosphere             test/runtests.jl:373-427 (icosphere test helper)
    build_rwg                  src/basis/RWG.jl:102-143
    assemble_Z_efie            src/assembly/EFIE.jl:699-753
    matrixfree_efie_operator   src/assembly/EFIE.jl:975-1000
    assemble_v_plane_wave / assemble_excitation / assemble_multiple_excitations
                               src/assembly/Excitation.jl:2684-2896
    solve_forward              src/solver/Solve.jl:1342-1382
    solve_gmres                src/solver/IterativeSolve.jl:341-392
    build_nearfield_preconditioner  src/solver/NearFieldPreconditioner.jl:732-759
    solve_scattering           src/Workflow.jl:67-317
    make_sph_grid, radiation_vectors, compute_farfield, bistatic_rcs
                               src/postprocessing/FarField.jl:216-605,
                               src/postprocessing/Diagnostics.jl:460-473

Every numerical kernel runs on the GPU through libdiffmom.so; importing this
package never falls back to host arithmetic.  When the library is missing the
first call raises RuntimeError.

Array conventions: a mesh's ``xyz`` is (Nv, 3) float64 and ``tri`` (Nt, 3)
int64 with 1-based vertex ids — the same memory as Julia's 3xNv / 3xNt
column-major matrices, so buffers pass to the C ABI unchanged.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiffmom.so")

# error codes (include/diffmom.h)
DM_OK, DM_ERR_ARGUMENT, DM_ERR_OVERFLOW, DM_ERR_DIMENSION = 0, 1, 2, 3
DM_ERR_CONVERGENCE, DM_ERR_DEVICE, DM_ERR_MEMORY, DM_ERR_DOMAIN = 4, 5, 6, 7
DM_Z64, DM_Z32 = 1, 2
# packed symmetric storage of the same matrices (upper triangle of 128 x 128 tiles)
DM_Z64P, DM_Z32P = 4, 8
DM_Z32H = 16  # compressed complex64 (products and solves only)
DM_Z32_ROUNDED = 32  # assemble_Z_efie flag: complex64 forms = rounded FP64 assembly
DM_METHOD_LU, DM_METHOD_GMRES, DM_METHOD_MIXED = 0, 1, 2
DM_PRECOND_NONE, DM_PRECOND_DIAG, DM_PRECOND_NEARFIELD, DM_PRECOND_LEARNED, DM_PRECOND_BLOCK = 0, 1, 2, 3, 4

ETA0 = 376.730313668
C0_DEFAULT = 299792458.0
_DEFAULT_MAX_DENSE_PAYLOAD_BYTES = 2_000_000_000
_DEFAULT_MAX_EFIE_CACHE_BYTES = 2_000_000_000  # src/assembly/EFIE.jl:14
_DEFAULT_MAX_EFIE_ADJACENCY_PAIRS = 20_000_000  # src/assembly/EFIE.jl:15
_DEFAULT_MAX_EXCITATION_WORK_BYTES = 512 * 1024 * 1024  # src/assembly/Excitation.jl:19
_DEFAULT_MAX_EXCITATION_TERMS = 200_000_000  # src/assembly/Excitation.jl:20
_DEFAULT_MAX_GMRES_WORKSPACE_BYTES = 512 * 1024 * 1024  # src/solver/IterativeSolve.jl:10


def _resource_limit(name, value, allow_zero=False):
    """_validated_resource_limit / _validated_nonnegative_resource_limit
    (src/Types.jl:41-65)."""
    v = int(value)
    if v != value or not (v >= 0 if allow_zero else v > 0):
        raise ValueError(f"{name} must be {'nonnegative' if allow_zero else 'positive'}, got {value}")
    return v


def _efie_limits(mesh, rwg, max_cache_bytes, max_adjacency_pairs):
    """The EFIE cache gates of _build_triangle_adjacency
    (src/assembly/EFIE.jl:140-231): the edge-adjacent triangle pairs against
    max_adjacency_pairs, and the device tables this library builds (per
    triangle: quadrature points of both rules, area, neighbours; per basis:
    triangles, signs, lengths; the near-field tables) against max_cache_bytes."""
    cache = _resource_limit("max_cache_bytes", max_cache_bytes)
    pairs_limit = _resource_limit("max_adjacency_pairs", max_adjacency_pairs, allow_zero=True)
    pairs = rwg.adjacent_pair_count()
    if pairs > pairs_limit:
        raise ValueError(f"triangle adjacency exceeds max_adjacency_pairs={pairs_limit}")
    nt, n = mesh.nt, rwg.nedges
    est = nt * (8 * 3 * (1 + 3 + 3 + 7) + 8 + 8 * 3 + 16 * 9 * 4) + n * (8 * 2 + 8 * 2 + 8 * 3)
    if est > cache:
        raise ValueError(f"EFIE cache requires at most {est} estimated bytes, exceeding max_cache_bytes={cache}")


class ConvergenceError(RuntimeError):
    """GMRES did not converge (the reference raises ErrorException)."""


class DeviceError(RuntimeError):
    pass


class DomainError(ValueError):
    pass


_EXC = {DM_ERR_ARGUMENT: ValueError, DM_ERR_OVERFLOW: OverflowError, DM_ERR_DIMENSION: ValueError,
        DM_ERR_CONVERGENCE: ConvergenceError, DM_ERR_DEVICE: DeviceError, DM_ERR_MEMORY: MemoryError,
        DM_ERR_DOMAIN: DomainError}


class DimensionMismatch(ValueError):
    pass


_EXC[DM_ERR_DIMENSION] = DimensionMismatch


class dm_solve_opts(C.Structure):
    _fields_ = [("method", C.c_int), ("precond", C.c_int), ("precond_side", C.c_int), ("maxiter", C.c_int),
                ("tol", C.c_double), ("nf_cutoff", C.c_double), ("restart", C.c_int),
                ("check_convergence", C.c_int), ("inner_tol", C.c_double), ("max_outer", C.c_int),
                ("block_size", C.c_int), ("batch", C.c_int), ("learned", C.c_void_p), ("adjoint", C.c_int),
                ("compress_tol", C.c_double), ("storage", C.c_int), ("graph", C.c_int), ("relax", C.c_double)]


class dm_solve_stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int), ("solved", C.c_int),
                ("residual", C.c_double), ("precond_time_s", C.c_double), ("solve_time_s", C.c_double)]


_L = None
SIGNATURES = None  # name -> (argtypes, restype) of every entry point the package binds


def lib():
    """Load libdiffmom.so (built by diffmom.jl_amd/build.py); raises when absent."""
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdiffmom.so is not built ({LIB_PATH}); run python diffmom.jl_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, i64, d, i32 = C.c_void_p, C.c_int64, C.c_double, C.c_int
    pp = C.POINTER(C.c_void_p)
    sig = {
        "dm_create": ([i32, pp], i32), "dm_destroy": ([vp], i32), "dm_last_error": ([], C.c_char_p),
        "dm_version": ([], i32), "dm_sync": ([vp], i32),
        "dm_mesh_create": ([vp, vp, i64, vp, i64, pp], i32), "dm_mesh_free": ([vp], i32),
        "dm_mesh_check": ([vp, i32, i32, d], i32), "dm_mesh_areas": ([vp, vp], i32),
        "dm_build_rwg": ([vp, i32, i32, i32, d, pp], i32),
        "dm_rwg_create": ([vp, i64] + [vp] * 10 + [pp], i32),
        "dm_rwg_size": ([vp, C.POINTER(i64)], i32), "dm_rwg_export": ([vp] + [vp] * 8, i32),
        "dm_rwg_free": ([vp], i32), "dm_rwg_adjacent_pairs": ([vp, C.POINTER(i64)], i32),
        "dm_assemble_z": ([vp, vp, d, d, d, i32, i32, i64, pp], i32),
        "dm_assemble_z_multi": ([vp, i64, vp, d, d, d, i32, i32, vp], i32),
        "dm_zmat_size": ([vp, C.POINTER(i64), C.POINTER(i32)], i32),
        "dm_zmat_get": ([vp, i32, vp], i32), "dm_zmat_entries": ([vp, i32, vp, vp, i64, vp], i32),
        "dm_zmat_free": ([vp], i32),
        "dm_apply": ([vp, vp, i32, i32, vp, i64, i64, vp, i64], i32),
        "dm_apply_noalloc": ([vp, vp, i32, i32, vp, vp], i32),
        "dm_zmat_compress": ([vp, C.c_double, i32, vp], i32),
        "dm_farfield": ([vp, vp, vp, i64, C.c_double, C.c_double, i32, vp, i64, i64, vp, i64], i32),
        "dm_mf_create": ([vp, vp, d, d, d, i32, pp], i32), "dm_mf_apply": ([vp, vp, i32, vp, vp], i32),
        "dm_mf_free": ([vp], i32),
        "dm_excitation_planewave": ([vp, vp, vp, vp, vp, i64, i32, vp], i32),
        "dm_solve_opts_default": ([C.POINTER(dm_solve_opts)], None),
        "dm_solve": ([vp, vp, vp, i64, i64, C.POINTER(dm_solve_opts), vp, i64, C.POINTER(dm_solve_stats)], i32),
        "dm_radiation_vectors": ([vp, vp, vp, i64, d, d, i32, vp], i32),
        "dm_solve_scattering_pw": ([vp, vp, d, d, d, i32, vp, vp, vp, i64, C.POINTER(dm_solve_opts), vp,
                                    C.POINTER(dm_solve_stats), C.POINTER(d), pp], i32),
        "dm_precond_create": ([vp, vp, i32, d, i32, i32, pp], i32), "dm_precond_apply": ([vp, vp, vp, vp], i32),
        "dm_precond_free": ([vp], i32),
        "dm_nearfield_pattern_info": ([vp, d, C.POINTER(i64), C.POINTER(d), C.POINTER(i32), C.POINTER(i64)], i32),
        "dm_profile_enable": ([vp, i32], i32), "dm_profile_reset": ([vp], i32),
        "dm_profile_read": ([vp, C.c_char_p, C.POINTER(i64), C.POINTER(d)], i32),
        "dm_tune_get": ([vp, C.c_char_p, i64, C.POINTER(i64)], i32), "dm_tune_set": ([vp, C.c_char_p], i32),
        "dm_zmat_hlevel": ([vp, i32, vp], i32),
        "dm_mpnn_param_count": ([], i64), "dm_mpnn_create": ([vp, vp, pp], i32),
        "dm_mpnn_get_params": ([vp, vp], i32), "dm_mpnn_free": ([vp], i32),
        "dm_rwg_graph": ([vp, d, vp, vp, vp], i32),
        "dm_mpnn_data_create": ([vp, i64, vp, vp, vp, vp, vp, i64, vp, pp], i32),
        "dm_mpnn_data_free": ([vp], i32),
        "dm_mpnn_data_set_matrix": ([vp, i64, i64, vp, i64], i32),
        "dm_mpnn_residual_operator": ([vp, vp, vp, vp], i32),
        "dm_mpnn_loss_grad": ([vp, vp, vp, i64, i64, C.POINTER(d), vp], i32),
        "dm_mpnn_train": ([vp, vp, vp, i32, i64, d, C.c_uint64, vp], i32),
        "dm_mpnn_apply": ([vp, vp, vp, d, vp, vp, vp], i32),
    }
    global SIGNATURES
    SIGNATURES = sig
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _L = L
    return L


def _check(rc):
    if rc != DM_OK:
        msg = lib().dm_last_error().decode(errors="replace")
        raise _EXC.get(rc, RuntimeError)(msg)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Context:
    """One GPU device and its HIP streams (dm_create)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().dm_create(int(device), C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        _check(lib().dm_sync(self.h))

    def profile(self, on=True):
        """Bracket tracked kernel launches with HIP events on the library's
        stream: True (every tracked launch), "products" (only the operator
        products, "symv*" / "gemv*"), False (off)."""
        mode = 2 if on == "products" else (1 if on else 0)
        _check(lib().dm_profile_enable(self.h, mode))

    def profile_reset(self):
        _check(lib().dm_profile_reset(self.h))

    def tune_get(self):
        """Launch shapes chosen by the library's tuner so far, as a dict."""
        n = C.c_int64()
        _check(lib().dm_tune_get(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        _check(lib().dm_tune_get(self.h, buf, n.value + 1, C.byref(n)))
        import json
        return json.loads(buf.value.decode())

    def tune_set(self, choices):
        """Load tuner choices (e.g. saved from tune_get in an earlier process)."""
        import json
        _check(lib().dm_tune_set(self.h, json.dumps({str(k): int(v) for k, v in choices.items()}).encode()))

    def profile_read(self, name):
        """(launch count, total milliseconds) of one tracked kernel name."""
        c, t = C.c_int64(), C.c_double()
        _check(lib().dm_profile_read(self.h, name.encode(), C.byref(c), C.byref(t)))
        return c.value, t.value

    def close(self):
        if getattr(self, "h", None):
            lib().dm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CTX = None


def default_context():
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


# ---------------------------------------------------------------------- mesh
class TriMesh:
    """Vertices (Nv, 3) float64 and 1-based triangles (Nt, 3) int64 (src/Types.jl:123-129)."""

    def __init__(self, xyz, tri, ctx=None):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int64)
        if self.xyz.ndim != 2 or self.xyz.shape[1] != 3 or self.tri.ndim != 2 or self.tri.shape[1] != 3:
            raise ValueError("xyz must be (Nv, 3) and tri (Nt, 3)")
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        _check(lib().dm_mesh_create(self.ctx.h, _ptr(self.xyz), self.nv, _ptr(self.tri), self.nt, C.byref(h)))
        self.h = h

    @property
    def nv(self):
        return self.xyz.shape[0]

    @property
    def nt(self):
        return self.tri.shape[0]

    def areas(self):
        out = np.zeros(self.nt)
        _check(lib().dm_mesh_areas(self.h, _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().dm_mesh_free(self.h)
            self.h = None


def nvertices(mesh):
    return mesh.nv


def ntriangles(mesh):
    return mesh.nt


def make_rect_plate(Lx, Ly, Nx, Ny, ctx=None):
    """Plate in the xy-plane centred at the origin, (Nx+1)(Ny+1) vertices and
    2 Nx Ny triangles (src/geometry/Mesh.jl:309-396)."""
    Lx, Ly = float(Lx), float(Ly)
    if not (math.isfinite(Lx) and Lx > 0 and math.isfinite(Ly) and Ly > 0):
        raise ValueError("Lx and Ly must be positive and finite")
    if int(Nx) < 1 or int(Ny) < 1:
        raise ValueError("Nx and Ny must be positive")
    Nx, Ny = int(Nx), int(Ny)
    hx, hy = Lx / 2, Ly / 2

    def axis(half, cells):
        i = np.arange(cells + 1, dtype=np.float64)
        c = half * (2 * (i / cells) - 1)
        c[0], c[-1] = -half, half
        return c

    X, Y = np.meshgrid(axis(hx, Nx), axis(hy, Ny))  # rows = iy
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    jy, jx = np.meshgrid(np.arange(Ny), np.arange(Nx), indexing="ij")
    v1 = (jy * (Nx + 1) + jx + 1).ravel()
    v2, v3, v4 = v1 + 1, v1 + Nx + 2, v1 + Nx + 1
    tri = np.empty((2 * Nx * Ny, 3), dtype=np.int64)
    tri[0::2] = np.stack([v1, v2, v3], axis=1)
    tri[1::2] = np.stack([v1, v3, v4], axis=1)
    return TriMesh(xyz, tri, ctx)


def make_icosphere(radius=0.05, subdivisions=2, ctx=None):
    """Icosahedron refined by edge midpoints projected to the sphere, scaled by
    radius (the reference's test helper, test/runtests.jl:373-427)."""
    phi = (1 + math.sqrt(5.0)) / 2
    verts = [(-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0), (0, -1, phi), (0, 1, phi),
             (0, -1, -phi), (0, 1, -phi), (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(1, 12, 6), (1, 6, 2), (1, 2, 8), (1, 8, 11), (1, 11, 12), (2, 6, 10), (6, 12, 5), (12, 11, 3),
             (11, 8, 7), (8, 2, 9), (4, 10, 5), (4, 5, 3), (4, 3, 7), (4, 7, 9), (4, 9, 10), (5, 10, 6),
             (3, 5, 12), (7, 3, 11), (9, 7, 8), (10, 9, 2)]
    for _ in range(int(subdivisions)):
        mid = {}
        nf = []

        def midpoint(i, j):
            key = (min(i, j), max(i, j))
            if key in mid:
                return mid[key]
            m = (verts[i - 1] + verts[j - 1]) / 2
            m = m / math.sqrt(float(m @ m))
            verts.append(m)
            mid[key] = len(verts)
            return len(verts)

        for (i, j, k) in faces:
            a, b, c = midpoint(i, j), midpoint(j, k), midpoint(k, i)
            nf += [(i, a, c), (j, b, a), (k, c, b), (a, b, c)]
        faces = nf
    xyz = float(radius) * np.array(verts)
    return TriMesh(xyz, np.array(faces, dtype=np.int64), ctx)


# ----------------------------------------------------------------------- RWG
class RWGData:
    """RWG basis arrays (src/Types.jl:138-152), indices 1-based, any order."""

    def __init__(self, mesh, handle, arrays):
        self.mesh = mesh
        self.h = handle
        for k, v in arrays.items():
            setattr(self, k, v)

    @property
    def nedges(self):
        return len(self.tplus)

    def adjacent_pair_count(self):
        c = C.c_int64()
        _check(lib().dm_rwg_adjacent_pairs(self.h, C.byref(c)))
        return c.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().dm_rwg_free(self.h)
            self.h = None


def _export_rwg(mesh, h):
    n = C.c_int64()
    _check(lib().dm_rwg_size(h, C.byref(n)))
    n = n.value
    a = dict(tplus=np.zeros(n, np.int64), tminus=np.zeros(n, np.int64), evert=np.zeros((n, 2), np.int64),
             vplus_opp=np.zeros(n, np.int64), vminus_opp=np.zeros(n, np.int64), len=np.zeros(n),
             area_plus=np.zeros(n), area_minus=np.zeros(n))
    _check(lib().dm_rwg_export(h, *[_ptr(a[k]) for k in ("tplus", "tminus", "evert", "vplus_opp", "vminus_opp",
                                                       "len", "area_plus", "area_minus")]))
    a["coeff_plus"] = np.ones(n)
    a["coeff_minus"] = np.ones(n)
    return a


def build_rwg(mesh, precheck=True, allow_boundary=True, require_closed=False, area_tol_rel=1e-12):
    """One basis per interior edge, T+ = lower-index triangle (src/basis/RWG.jl:102-143).
    Basis order: sorted (v1 < v2) edge keys."""
    h = C.c_void_p()
    _check(lib().dm_build_rwg(mesh.h, int(precheck), int(allow_boundary), int(require_closed),
                              float(area_tol_rel), C.byref(h)))
    return RWGData(mesh, h, _export_rwg(mesh, h))


def rwg_from_arrays(mesh, tplus, tminus, vplus_opp, vminus_opp, length, area_plus, area_minus,
                    evert=None, coeff_plus=None, coeff_minus=None):
    """RWGData as a caller holds it (e.g. a Julia RWGData in Dict order)."""
    arr = dict(tplus=np.ascontiguousarray(tplus, np.int64), tminus=np.ascontiguousarray(tminus, np.int64),
               vplus_opp=np.ascontiguousarray(vplus_opp, np.int64),
               vminus_opp=np.ascontiguousarray(vminus_opp, np.int64), len=np.ascontiguousarray(length, np.float64),
               area_plus=np.ascontiguousarray(area_plus, np.float64),
               area_minus=np.ascontiguousarray(area_minus, np.float64))
    n = len(arr["tplus"])
    arr["evert"] = np.ascontiguousarray(evert, np.int64) if evert is not None else np.zeros((n, 2), np.int64)
    arr["coeff_plus"] = np.ascontiguousarray(coeff_plus, np.float64) if coeff_plus is not None else np.ones(n)
    arr["coeff_minus"] = np.ascontiguousarray(coeff_minus, np.float64) if coeff_minus is not None else np.ones(n)
    h = C.c_void_p()
    _check(lib().dm_rwg_create(mesh.h, n, _ptr(arr["tplus"]), _ptr(arr["tminus"]), _ptr(arr["evert"]),
                               _ptr(arr["vplus_opp"]), _ptr(arr["vminus_opp"]), _ptr(arr["len"]),
                               _ptr(arr["area_plus"]), _ptr(arr["area_minus"]), _ptr(arr["coeff_plus"]),
                               _ptr(arr["coeff_minus"]), C.byref(h)))
    return RWGData(mesh, h, arr)


def _check_pair(mesh, rwg):
    if rwg.mesh is not mesh:
        raise ValueError("RWG data must be built from the supplied mesh object.")


def assert_mesh_quality(mesh, allow_boundary=True, require_closed=False, area_tol_rel=1e-12):
    _check(lib().dm_mesh_check(mesh.h, int(allow_boundary), int(require_closed), float(area_tol_rel)))


# ------------------------------------------------------------------ matrices
class ZMatrix:
    """Device-resident EFIE matrix (complex128 and/or complex64 forms)."""

    def __init__(self, rwg, handle, forms, k, eta0, quad_order):
        self.rwg = rwg
        self.h = handle
        self.forms = forms
        self.k = k
        self.eta0 = eta0
        self.quad_order = quad_order
        self.N = rwg.nedges
        self.ctx = rwg.mesh.ctx

    @property
    def shape(self):
        return (self.N, self.N)

    def _form(self, form):
        if form is not None:
            return form
        for f in (DM_Z64, DM_Z64P, DM_Z32, DM_Z32P):
            if self.forms & f:
                return f
        raise ValueError("matrix holds no stored form")

    def compress(self, eps=5e-8, rmax=48):
        """Cross-approximate the off-diagonal tiles of the packed complex64 form
        (dm_zmat_compress): ||Z32P - Z32H||_F <= eps ||Z32P||_F.  Adds DM_Z32H,
        which mixed-precision solves then use for the inner products.  Returns
        the tile statistics."""
        st = np.zeros(4)
        _check(lib().dm_zmat_compress(self.h, float(eps), int(rmax), _ptr(st)))
        self.forms |= DM_Z32H
        return dict(tiles=int(st[0]), lowrank_tiles=int(st[1]), mean_rank=float(st[2]), bytes_ratio=float(st[3]))

    def compress_level(self, level):
        """Select the compressed form's level (0..10) for later DM_Z32H
        products (dm_zmat_hlevel): level l uses the prefix of every tile's
        cross-approximation factors at 10^(l/2) times the tolerance (error
        bound eps 10^(l/2)).  Returns that level's statistics and its bound."""
        st = np.zeros(5)
        _check(lib().dm_zmat_hlevel(self.h, int(level), _ptr(st)))
        return dict(tiles=int(st[0]), lowrank_tiles=int(st[1]), mean_rank=float(st[2]), bytes_ratio=float(st[3]),
                    eps=float(st[4]))

    def to_numpy(self, form=None):
        out = np.zeros((self.N, self.N), dtype=np.complex128, order="F")
        _check(lib().dm_zmat_get(self.h, self._form(form), out.ctypes.data_as(C.c_void_p)))
        return out

    def entries(self, m, n, form=None):
        m = np.ascontiguousarray(m, np.int64)
        n = np.ascontiguousarray(n, np.int64)
        out = np.zeros(len(m), np.complex128)
        _check(lib().dm_zmat_entries(self.h, self._form(form), _ptr(m), _ptr(n), len(m), _ptr(out)))
        return out

    def matmul(self, X, form=None, adjoint=False):
        X = np.asarray(X, dtype=np.complex128)
        vec = X.ndim == 1
        Xf = np.asfortranarray(X.reshape(self.N, -1))
        Y = np.zeros_like(Xf, order="F")
        _check(lib().dm_apply(self.ctx.h, self.h, self._form(form), int(adjoint), Xf.ctypes.data_as(C.c_void_p),
                              self.N, Xf.shape[1], Y.ctypes.data_as(C.c_void_p), self.N))
        return Y[:, 0] if vec else Y

    def __matmul__(self, X):
        return self.matmul(X)

    def free(self):
        if getattr(self, "h", None):
            lib().dm_zmat_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _k_parts(k):
    k = complex(k)
    return k.real, k.imag


def assemble_Z_efie(mesh, rwg, k, quad_order=3, eta0=ETA0, mesh_precheck=True, allow_boundary=True,
                    require_closed=False, area_tol_rel=1e-12, max_output_bytes=_DEFAULT_MAX_DENSE_PAYLOAD_BYTES,
                    max_cache_bytes=_DEFAULT_MAX_EFIE_CACHE_BYTES,
                    max_adjacency_pairs=_DEFAULT_MAX_EFIE_ADJACENCY_PAIRS, forms=DM_Z64, device=False):
    """Dense EFIE matrix (src/assembly/EFIE.jl:699-753).  Returns a complex128
    numpy matrix like the reference, or the device-resident ZMatrix when
    device=True (forms selects complex128 and/or complex64 storage; with
    DM_Z32_ROUNDED the complex64 forms are the rounding of the FP64 assembly,
    which compresses far better than the FP32-quadrature ones)."""
    _check_pair(mesh, rwg)
    N = rwg.nedges
    if max_output_bytes is not None and N * N * 16 > max_output_bytes:
        raise ValueError(f"dense EFIE matrix requires {N * N * 16} bytes, exceeding max_output_bytes={max_output_bytes}")
    if mesh_precheck:
        assert_mesh_quality(mesh, allow_boundary, require_closed, area_tol_rel)
    _efie_limits(mesh, rwg, max_cache_bytes, max_adjacency_pairs)
    kr, ki = _k_parts(k)
    h = C.c_void_p()
    _check(lib().dm_assemble_z(mesh.ctx.h, rwg.h, kr, ki, float(eta0), int(quad_order), int(forms),
                               int(max_output_bytes or 0), C.byref(h)))
    Z = ZMatrix(rwg, h, int(forms) & ~DM_Z32_ROUNDED, complex(k), float(eta0), quad_order)
    if device:
        return Z
    out = Z.to_numpy()
    Z.free()
    return out


class MatrixFreeEFIEOperator:
    """Matrix-free EFIE operator (src/assembly/EFIE.jl:759-1007): every product
    recomputes all entries on the GPU in FP64; nothing of size N^2 is stored."""

    def __init__(self, rwg, h, k, eta0, quad_order, adjoint=False):
        self.rwg, self.h, self.k, self.eta0, self.quad_order = rwg, h, k, eta0, quad_order
        self.N = rwg.nedges
        self._adjoint = adjoint
        self.ctx = rwg.mesh.ctx

    @property
    def shape(self):
        return (self.N, self.N)

    def __matmul__(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.shape != (self.N,):
            raise DimensionMismatch(f"x length {x.shape} != {self.N}")
        y = np.zeros(self.N, np.complex128)
        _check(lib().dm_mf_apply(self.ctx.h, self.h, int(self._adjoint), _ptr(x), _ptr(y)))
        return y

    def adjoint(self):
        return MatrixFreeEFIEOperator(self.rwg, self.h, self.k, self.eta0, self.quad_order, not self._adjoint)

    @property
    def H(self):
        return self.adjoint()


def assemble_Z_efie_multi(meshes, rwgs, k, quad_order=3, eta0=ETA0, forms=DM_Z64):
    """Device matrices of many meshes from one persistent assembly launch per
    stored form (all workgroups share one tile queue over every mesh; no
    reference counterpart).  Returns a list of ZMatrix."""
    if len(meshes) != len(rwgs) or not meshes:
        raise ValueError("need one RWG set per mesh")
    for m, r in zip(meshes, rwgs):
        _check_pair(m, r)
    ctx = meshes[0].ctx
    hs = (C.c_void_p * len(rwgs))(*[r.h.value for r in rwgs])
    outs = (C.c_void_p * len(rwgs))()
    kr, ki = _k_parts(k)
    _check(lib().dm_assemble_z_multi(ctx.h, len(rwgs), C.cast(hs, C.c_void_p), kr, ki, float(eta0), int(quad_order),
                                     int(forms), C.cast(outs, C.c_void_p)))
    return [ZMatrix(r, C.c_void_p(h), forms, complex(k), float(eta0), quad_order) for r, h in zip(rwgs, outs)]


def matrixfree_efie_operator(mesh, rwg, k, quad_order=3, eta0=ETA0, mesh_precheck=True, allow_boundary=True,
                             require_closed=False, area_tol_rel=1e-12, max_cache_bytes=_DEFAULT_MAX_EFIE_CACHE_BYTES,
                             max_adjacency_pairs=_DEFAULT_MAX_EFIE_ADJACENCY_PAIRS):
    """src/assembly/EFIE.jl:975-1000."""
    _check_pair(mesh, rwg)
    if mesh_precheck:
        assert_mesh_quality(mesh, allow_boundary, require_closed, area_tol_rel)
    _efie_limits(mesh, rwg, max_cache_bytes, max_adjacency_pairs)
    kr, ki = _k_parts(k)
    h = C.c_void_p()
    _check(lib().dm_mf_create(mesh.ctx.h, rwg.h, kr, ki, float(eta0), int(quad_order), C.byref(h)))
    return MatrixFreeEFIEOperator(rwg, h, complex(k), float(eta0), quad_order)


# ---------------------------------------------------------------- excitation
@dataclass
class PlaneWaveExcitation:
    """src/assembly/Excitation.jl:86-101"""
    k_vec: np.ndarray
    E0: float
    pol: np.ndarray


def make_plane_wave(k_vec, E0, pol):
    return PlaneWaveExcitation(np.asarray(k_vec, float), float(E0), np.asarray(pol, float))


MANY_ANGLE_SEED = 20261202


def seeded_planewaves(k, count, rng):
    """count plane waves from a numpy Generator: directions are normalised
    N(0, I3) draws, polarisations random unit vectors projected perpendicular
    to the direction, E0 = 1 (the many-angle workload)."""
    out = []
    for _ in range(count):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        p = rng.standard_normal(3)
        p -= d * (p @ d)
        p /= np.linalg.norm(p)
        out.append(make_plane_wave(k * d, 1.0, p))
    return out


def seeded_plate_specs(count, rng):
    """(Lx, Ly, Nx, Ny) of count plates with sides ~ U[0.05, 0.3] m and
    Nx, Ny ~ U{4..30} (make_rect_plate, src/geometry/Mesh.jl:309-396), drawn
    from a numpy Generator."""
    out = []
    for _ in range(count):
        lx, ly = rng.uniform(0.05, 0.3, 2)
        nx, ny = rng.integers(4, 31, 2)
        out.append((float(lx), float(ly), int(nx), int(ny)))
    return out


def many_angle_workload(k, nwaves=256, nplates=64, seed=MANY_ANGLE_SEED):
    """The many-angle / many-plate workload: nwaves seeded plane waves, then
    nplates seeded plate specs from the same generator stream."""
    rng = np.random.default_rng(seed)
    waves = seeded_planewaves(k, nwaves, rng)
    return waves, seeded_plate_specs(nplates, rng)


def assemble_multiple_excitations(mesh, rwg, excitations, quad_order=3,
                                  max_output_bytes=_DEFAULT_MAX_DENSE_PAYLOAD_BYTES,
                                  max_work_bytes=_DEFAULT_MAX_EXCITATION_WORK_BYTES,
                                  max_terms=_DEFAULT_MAX_EXCITATION_TERMS):
    """src/assembly/Excitation.jl:2738-2797 for plane waves: the output byte
    gate as the reference's; the work and term limits are range-checked (the
    device path keeps no per-term workspace)."""
    _check_pair(mesh, rwg)
    M = len(excitations)
    out = _resource_limit("max_output_bytes", max_output_bytes)
    if 16 * rwg.nedges * M > out:
        raise ValueError(f"multiple-excitation RHS matrix requires {16 * rwg.nedges * M} raw bytes, exceeding "
                         f"max_output_bytes={out}")
    _resource_limit("max_work_bytes", max_work_bytes)
    _resource_limit("max_terms", max_terms)
    kv = np.ascontiguousarray(np.array([e.k_vec for e in excitations], float).reshape(M, 3))
    pol = np.ascontiguousarray(np.array([e.pol for e in excitations], float).reshape(M, 3))
    E0 = np.ascontiguousarray([float(e.E0) for e in excitations])
    V = np.zeros((rwg.nedges, M), dtype=np.complex128, order="F")
    _check(lib().dm_excitation_planewave(mesh.ctx.h, rwg.h, _ptr(kv), _ptr(pol), _ptr(E0), M, int(quad_order),
                                         V.ctypes.data_as(C.c_void_p)))
    return V


def assemble_v_plane_wave(mesh, rwg, k_vec, E0, pol, quad_order=3):
    return assemble_multiple_excitations(mesh, rwg, [make_plane_wave(k_vec, E0, pol)], quad_order)[:, 0]


def assemble_excitation(mesh, rwg, excitation, quad_order=3):
    if not isinstance(excitation, PlaneWaveExcitation):
        raise NotImplementedError("only plane-wave excitations are on the GPU path")
    return assemble_v_plane_wave(mesh, rwg, excitation.k_vec, excitation.E0, excitation.pol, quad_order)


# ------------------------------------------------------------------- solvers
@dataclass
class NearFieldPreconditionerData:
    """Near-field preconditioner spec (src/solver/NearFieldPreconditioner.jl:84-110).

    factorization="spai": a sparse approximate inverse M of the near-field
    matrix (entries with RWG-centre distance <= cutoff), minimising
    ||Z M - I||_F column by column over the cutoff pattern (the nearest 80
    bases within the cutoff per column), applied as one SpMV per iteration.
    The reference factors the same near-field matrix with a sparse LU
    (factorization=:lu, :700-712) or ILU (:732-759); neither factorisation is
    offered on the GPU path, and asking for one raises ValueError.
    factorization="diag": Jacobi (src/solver/NearFieldPreconditioner.jl:623-668).
    factorization="block": block Jacobi over block_size spatially consecutive bases."""
    cutoff: float
    factorization: str = "spai"
    block_size: int = 128

    def pattern_info(self, rwg):
        """(nnz, mean radius covered per column in metres, column cap, columns
        the cap cut short of the cutoff) of the SPAI pattern this cutoff gives
        on rwg."""
        nnz, rad, cap, capped = C.c_int64(), C.c_double(), C.c_int(), C.c_int64()
        _check(lib().dm_nearfield_pattern_info(rwg.h, float(self.cutoff), C.byref(nnz), C.byref(rad), C.byref(cap),
                                               C.byref(capped)))
        return nnz.value, rad.value, cap.value, capped.value


_NO_FACTOR = ("factorization={!r}: the GPU path does not factor the near-field matrix; it builds a sparse "
              "approximate inverse of it (factorization='spai'), or Jacobi ('diag') / block Jacobi ('block')")


def build_nearfield_preconditioner(Z, mesh, rwg, cutoff, factorization="spai", block_size=128):
    if factorization in ("lu", "ilu"):
        raise ValueError(_NO_FACTOR.format(factorization))
    if factorization not in ("spai", "diag", "block"):
        raise ValueError(f"unsupported factorization {factorization!r} on the GPU path (use 'spai', 'diag' or "
                         "'block')")
    if not (cutoff >= 0 and math.isfinite(cutoff)):
        raise ValueError("cutoff must be finite and nonnegative")
    return NearFieldPreconditionerData(float(cutoff), factorization, int(block_size))


class NearFieldOperator:
    """The preconditioner as an operator, y = M x (NearFieldOperator's mul!,
    src/solver/NearFieldPreconditioner.jl:1034-1060), built on the GPU from the
    device matrix Z (form DM_Z64 or DM_Z32)."""

    def __init__(self, P, Z, form=None):
        kind = {"diag": DM_PRECOND_DIAG, "block": DM_PRECOND_BLOCK}.get(P.factorization, DM_PRECOND_NEARFIELD)
        h = C.c_void_p()
        _check(lib().dm_precond_create(Z.ctx.h, Z.h, kind, float(P.cutoff), int(P.block_size), Z._form(form),
                                       C.byref(h)))
        self.h, self.Z, self.N = h, Z, Z.N

    def __matmul__(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.shape != (self.N,):
            raise DimensionMismatch(f"x length {x.shape} != {self.N}")
        y = np.zeros(self.N, np.complex128)
        _check(lib().dm_precond_apply(self.Z.ctx.h, self.h, _ptr(x), _ptr(y)))
        return y

    def __del__(self):
        if getattr(self, "h", None):
            lib().dm_precond_free(self.h)
            self.h = None


@dataclass
class SolveStats:
    niter: int
    solved: bool
    residual: float
    outer: int = 0
    precond_time_s: float = 0.0
    solve_time_s: float = 0.0
    residuals: list = field(default_factory=list)


def _solve(Z, B, method, precond=None, side="right", tol=1e-6, maxiter=300, restart=0, check=True,
           inner_tol=0.0, max_outer=10, batch=0, adjoint=False, graph=False):
    if not isinstance(Z, ZMatrix):
        raise TypeError("GPU solves take the device matrix (assemble_Z_efie(..., device=True))")
    B = np.asarray(B, dtype=np.complex128)
    vec = B.ndim == 1
    Bf = np.asfortranarray(B.reshape(Z.N, -1))
    if Bf.shape[0] != Z.N:
        raise DimensionMismatch(f"rhs length {Bf.shape[0]} != {Z.N}")
    X = np.zeros_like(Bf, order="F")
    o = _opts(method, precond, side, tol, maxiter, restart, check, inner_tol, max_outer, batch, graph=graph)
    o.adjoint = int(adjoint)
    st = dm_solve_stats()
    _check(lib().dm_solve(Z.ctx.h, Z.h, Bf.ctypes.data_as(C.c_void_p), Z.N, Bf.shape[1], C.byref(o),
                          X.ctypes.data_as(C.c_void_p), Z.N, C.byref(st)))
    stats = SolveStats(st.iterations, bool(st.solved), st.residual, st.outer_iterations, st.precond_time_s,
                       st.solve_time_s)
    return (X[:, 0] if vec else X), stats


def _opts(method, precond=None, side="right", tol=1e-6, maxiter=300, restart=0, check=True, inner_tol=0.0,
          max_outer=10, batch=0, compress_tol=0.0, storage="packed", graph=False, relax=0.0):
    """dm_solve_opts.  batch: 0 = lockstep batches of right-hand sides when there
    are >= 8 of them, 1 = always, -1 = one right-hand side at a time.
    compress_tol > 0: the mixed solve compresses its complex64 matrix.
    storage: "packed" (symmetric tiles, half the bytes) or "full" for the
    one-column FP64 GMRES path of solve_planewaves.  graph: one right-hand
    side at a time, each Krylov step after the first of a cycle replays a
    captured HIP graph.  relax > 0 (mixed): inner products relaxed as the
    residual falls -- prefixes of the cross-approximation factors (compressed,
    one right-hand side at a time) or three instead of six bf16 part products
    (lockstep batches)."""
    if storage not in ("packed", "full"):
        raise ValueError(f"storage must be 'packed' or 'full', got {storage!r}")
    o = dm_solve_opts()
    lib().dm_solve_opts_default(C.byref(o))
    o.method, o.tol, o.maxiter, o.restart = method, float(tol), int(maxiter), int(restart)
    o.batch = int(batch)
    o.storage = 0 if storage == "packed" else 1
    o.graph = 1 if graph else 0
    o.relax = float(relax)
    o.compress_tol = float(compress_tol)
    o.check_convergence, o.inner_tol, o.max_outer = int(check), float(inner_tol), int(max_outer)
    o.precond_side = 1 if side == "right" else 0
    if precond is None or precond == "none":
        o.precond = DM_PRECOND_NONE
    elif isinstance(precond, NearFieldPreconditionerData):
        o.precond = {"diag": DM_PRECOND_DIAG, "block": DM_PRECOND_BLOCK}.get(precond.factorization,
                                                                             DM_PRECOND_NEARFIELD)
        o.block_size, o.nf_cutoff = precond.block_size, precond.cutoff
    elif precond == "diag":
        o.precond = DM_PRECOND_DIAG
    elif isinstance(precond, LearnedPreconditioner):
        if side != "right":
            raise ValueError("the learned preconditioner is nonlinear: use precond_side='right'")
        o.precond = DM_PRECOND_LEARNED
        o.nf_cutoff = precond.cutoff
        o.learned = precond.net.h
    else:
        raise ValueError(f"unknown preconditioner {precond!r}")
    return o


def solve_planewaves(mesh, rwg, k, excitations, method="mixed", preconditioner=None, tol=1e-6, maxiter=300,
                     quad_order=3, eta0=ETA0, check_convergence=True, keep_matrix=False, batch=0, restart=0,
                     compress_tol=0.0, storage="packed", graph=False, relax=0.0, inner_tol=0.0):
    """End-to-end device solve for plane waves (the dense branches of
    solve_scattering, src/Workflow.jl:67-317): assembly, excitation,
    preconditioner and Krylov solve in one call.  method: "direct" (dense LU),
    "fp64" (FP64 GMRES on the complex128 matrix) or "mixed" (FGMRES on the
    complex64 matrix with FP64 Arnoldi and FP64 refinement against the
    complex128 matrix; one FP64 assembly writes both).  compress_tol > 0 (mixed,
    one wave at a time): the complex64 matrix is compressed first
    (ZMatrix.compress) and the inner products use the compressed form.
    storage ("fp64", one wave at a time): "packed" keeps the upper triangle of
    128 x 128 tiles (Z = Z^T exactly; each product reads half the bytes),
    "full" the N x N matrix as the reference stores it.  graph (one wave at
    a time): Krylov steps replay a captured HIP graph.  relax > 0 (mixed):
    the inner FGMRES relaxes its products as its residual falls (inexact
    Krylov) -- one wave at a time with compress_tol > 0 it reads a prefix of
    the cross-approximation factors, a lockstep batch drops its complex64
    Z X from six to three bf16 part products once the worst active column
    allows a 5e-5 product error; inner_tol: the first inner solve's relative
    tolerance (0: tol / 2).
    Returns (X, stats, assembly_time_s[, ZMatrix])."""
    _check_pair(mesh, rwg)
    M = len(excitations)
    kv = np.ascontiguousarray(np.array([e.k_vec for e in excitations], float).reshape(M, 3))
    pol = np.ascontiguousarray(np.array([e.pol for e in excitations], float).reshape(M, 3))
    E0 = np.ascontiguousarray([float(e.E0) for e in excitations])
    meth = {"direct": DM_METHOD_LU, "fp64": DM_METHOD_GMRES, "mixed": DM_METHOD_MIXED}[method]
    o = _opts(meth, preconditioner, "right", tol, maxiter, restart=restart, check=check_convergence, batch=batch,
              compress_tol=compress_tol, storage=storage, graph=graph, relax=relax, inner_tol=inner_tol)
    X = np.zeros((rwg.nedges, M), dtype=np.complex128, order="F")
    st = dm_solve_stats()
    t_asm = C.c_double()
    kr, ki = _k_parts(k)
    zh = C.c_void_p()
    _check(lib().dm_solve_scattering_pw(mesh.ctx.h, rwg.h, kr, ki, float(eta0), int(quad_order), _ptr(kv), _ptr(pol),
                                        _ptr(E0), M, C.byref(o), X.ctypes.data_as(C.c_void_p), C.byref(st),
                                        C.byref(t_asm), C.byref(zh) if keep_matrix else None))
    stats = SolveStats(st.iterations, bool(st.solved), st.residual, st.outer_iterations, st.precond_time_s,
                       st.solve_time_s)
    if keep_matrix:
        n = C.c_int64()
        forms = C.c_int()
        _check(lib().dm_zmat_size(zh, C.byref(n), C.byref(forms)))
        return X, stats, t_asm.value, ZMatrix(rwg, zh, forms.value, complex(k), float(eta0), quad_order)
    return X, stats, t_asm.value


def solve_forward(Z, v, solver="direct", **kw):
    """src/solver/Solve.jl:1342-1382: dense LU (:direct) or GMRES with a true-residual guard."""
    if solver == "direct":
        x, _ = _solve(Z, v, DM_METHOD_LU)
        return x
    if solver == "gmres":
        x, st = solve_gmres(Z, v, **kw)
        if not st.residual <= 100 * kw.get("tol", 1e-8):
            raise ConvergenceError(f"forward GMRES true residual too large: relative_residual={st.residual}")
        return x
    raise ValueError(f"unknown solver {solver!r}")


def solve_gmres(Z, rhs, preconditioner=None, precond_side="left", tol=1e-8, maxiter=200, memory=20,
                max_workspace_bytes=_DEFAULT_MAX_GMRES_WORKSPACE_BYTES, verbose=False,
                check_gmres_convergence=True, restart=0, batch=0, graph=False):
    """src/solver/IterativeSolve.jl:341-392 (full GMRES unless restart > 0).
    Several right-hand sides with right or no preconditioning run as lockstep
    batches (one Z X product per step) unless batch=-1.  graph: one
    right-hand side at a time, Krylov steps replay a captured HIP graph."""
    if not (math.isfinite(tol) and tol > 0):
        raise ValueError(f"GMRES tol must be finite and positive, got {tol}")
    if maxiter < 1:
        raise ValueError(f"GMRES maxiter must be at least 1, got {maxiter}")
    if memory < 1:
        raise ValueError(f"GMRES memory must be at least 1, got {memory}")
    if precond_side not in ("left", "right"):
        raise ValueError(f"Invalid precond_side: {precond_side}")
    _gmres_workspace_gate(Z.N, memory, max_workspace_bytes)
    return _solve(Z, rhs, DM_METHOD_GMRES, preconditioner, precond_side, tol, maxiter, restart,
                  check_gmres_convergence, batch=batch, graph=graph)


def _gmres_workspace_gate(N, memory, max_workspace_bytes):
    """_preflight_gmres_workspace / _gmres_workspace_bytes
    (src/solver/IterativeSolve.jl:12-36): Krylov.jl's GMRES storage for
    min(N, memory) basis vectors against max_workspace_bytes."""
    lim = _resource_limit("max_workspace_bytes", max_workspace_bytes)
    m = min(N, memory)
    need = 16 * ((m + 5) * N + 2 * m + m * (m + 1) // 2) + 8 * m
    if need > lim:
        raise ValueError(f"GMRES Krylov workspace requires {need} raw bytes, exceeding max_workspace_bytes={lim}")


def solve_gmres_adjoint(Z, rhs, preconditioner=None, precond_side="left", tol=1e-8, maxiter=200, memory=20,
                        max_workspace_bytes=_DEFAULT_MAX_GMRES_WORKSPACE_BYTES, verbose=False,
                        check_gmres_convergence=True, restart=0, batch=0):
    """Z^H x = rhs (src/solver/IterativeSolve.jl:409-460).  Z is exactly
    symmetric, so the library solves Z conj(x) = conj(rhs) with the forward
    operator and preconditioner and conjugates the result."""
    if not (math.isfinite(tol) and tol > 0):
        raise ValueError(f"GMRES tol must be finite and positive, got {tol}")
    if maxiter < 1:
        raise ValueError(f"GMRES maxiter must be at least 1, got {maxiter}")
    if precond_side not in ("left", "right"):
        raise ValueError(f"Invalid precond_side: {precond_side}")
    _gmres_workspace_gate(Z.N, memory, max_workspace_bytes)
    return _solve(Z, rhs, DM_METHOD_GMRES, preconditioner, precond_side, tol, maxiter, restart,
                  check_gmres_convergence, batch=batch, adjoint=True)


def solve_mixed(Z, rhs, preconditioner=None, tol=1e-6, maxiter=300, inner_tol=0.0, max_outer=10,
                check_convergence=True, restart=0, batch=0, graph=False):
    """Mixed precision: FGMRES on the complex64 matrix with FP64 Arnoldi, FP64
    iterative refinement on the complex128 matrix (no reference counterpart).
    Several right-hand sides run as lockstep batches (Z X on MFMA) unless batch=-1."""
    return _solve(Z, rhs, DM_METHOD_MIXED, preconditioner, "right", tol, maxiter, restart, check_convergence,
                  inner_tol, max_outer, batch, graph=graph)


# --------------------------------------------------------------- far field
@dataclass
class SphGrid:
    rhat: np.ndarray  # (NΩ, 3)
    theta: np.ndarray
    phi: np.ndarray
    w: np.ndarray


def make_sph_grid(Ntheta, Nphi):
    """Uniform (θ, φ) grid, src/postprocessing/FarField.jl:216-277."""
    if Ntheta <= 0 or Nphi <= 0:
        raise ValueError("Ntheta and Nphi must be positive")
    dth, dph = math.pi / Ntheta, 2 * math.pi / Nphi
    th = np.repeat((np.arange(1, Ntheta + 1) - 0.5) * dth, Nphi)
    ph = np.tile((np.arange(1, Nphi + 1) - 0.5) * dph, Ntheta)
    rhat = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    return SphGrid(np.ascontiguousarray(rhat), th, ph, np.sin(th) * dth * dph)


def radiation_vectors(mesh, rwg, grid, k, quad_order=3, eta0=ETA0):
    """G (3NΩ x N): E∞ = G I (src/postprocessing/FarField.jl:425-574), on the GPU."""
    _check_pair(mesh, rwg)
    rhat = np.ascontiguousarray(grid.rhat if isinstance(grid, SphGrid) else grid, dtype=np.float64)
    no = rhat.shape[0]
    G = np.zeros((3 * no, rwg.nedges), dtype=np.complex128, order="F")
    _check(lib().dm_radiation_vectors(mesh.ctx.h, rwg.h, _ptr(rhat), no, float(k), float(eta0), int(quad_order),
                                      G.ctypes.data_as(C.c_void_p)))
    return G


def compute_farfield(G, I, NΩ):
    """E∞ (3, NΩ) = G I (src/postprocessing/FarField.jl:584-605)."""
    if G.shape[0] != 3 * NΩ:
        raise DimensionMismatch(f"G has {G.shape[0]} rows, expected {3 * NΩ}")
    return (G @ np.asarray(I, np.complex128)).reshape((3, NΩ), order="F")


def farfield(mesh, rwg, grid, k, I, quad_order=3, eta0=ETA0):
    """E∞ for one current vector (3, NΩ) or many (3, NΩ, M): compute_farfield
    of radiation_vectors (src/postprocessing/FarField.jl:425-605) without G
    leaving the GPU (chunks of directions, E = G I on the FP64 matrix cores)."""
    _check_pair(mesh, rwg)
    rhat = np.ascontiguousarray(grid.rhat if isinstance(grid, SphGrid) else grid, dtype=np.float64)
    no = rhat.shape[0]
    I = np.asarray(I, dtype=np.complex128)
    vec = I.ndim == 1
    If = np.asfortranarray(I.reshape(rwg.nedges, -1))
    M = If.shape[1]
    E = np.zeros((3 * no, M), dtype=np.complex128, order="F")
    _check(lib().dm_farfield(mesh.ctx.h, rwg.h, _ptr(rhat), no, float(k), float(eta0), int(quad_order),
                             If.ctypes.data_as(C.c_void_p), rwg.nedges, M, E.ctypes.data_as(C.c_void_p), 3 * no))
    E = E.reshape((3, no, M), order="F")
    return E[:, :, 0] if vec else E


def bistatic_rcs(E_ff, E0=1.0):
    """σ = 4π |E∞|² / |E0|² (src/postprocessing/Diagnostics.jl:460-473)."""
    return 4 * math.pi * np.sum(np.abs(E_ff) ** 2, axis=0) / abs(E0) ** 2


# -------------------------------------------------------------- workflow
@dataclass
class ScatteringResult:
    """src/Types.jl:1297-1308"""
    I_coeffs: np.ndarray
    method: str
    N: int
    assembly_time_s: float
    solve_time_s: float
    preconditioner_time_s: float
    gmres_iters: int
    gmres_residual: float
    mesh_report: dict
    warnings: list


def mesh_resolution_report(mesh, freq_hz, points_per_wavelength=10.0, c0=C0_DEFAULT):
    """src/geometry/Mesh.jl:3418-3471"""
    lam = c0 / freq_hz
    t = mesh.tri - 1
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    lens = np.sort(np.linalg.norm(mesh.xyz[e[:, 0]] - mesh.xyz[e[:, 1]], axis=1))
    target = lam / points_per_wavelength
    return dict(wavelength_m=lam, target_max_edge_m=target, edge_min_m=lens[0], edge_max_m=lens[-1],
                edge_median_m=float(np.median(lens)), edge_max_over_lambda=lens[-1] / lam,
                meets_target=bool(lens[-1] <= target))


def solve_scattering(mesh, freq_hz, excitation, method="auto", dense_direct_limit=2000, dense_gmres_limit=10000,
                     mlfma_threshold=50000, check_resolution=True, points_per_wavelength=10.0,
                     error_on_underresolved=False, gmres_tol=1e-6, gmres_maxiter=300, check_gmres_convergence=True,
                     nf_cutoff_lambda=1.0, preconditioner="auto", aca_tol=1e-6, aca_leaf_size=64, aca_eta=1.5,
                     aca_max_rank=50, verbose=False, quad_order=3, c0=C0_DEFAULT,
                     max_dense_matrix_bytes=_DEFAULT_MAX_DENSE_PAYLOAD_BYTES, precision="fp64", nf_block_size=128):
    """src/Workflow.jl:67-317 for the dense methods (:dense_direct, :dense_gmres).
    ``precision="mixed"`` selects the complex64 + FP64-refinement solve.  The
    ACA/MLFMA branches of the reference are not part of this library."""
    f = float(freq_hz)
    if not (math.isfinite(f) and f > 0):
        raise ValueError(f"solve_scattering: freq_hz must be finite and positive, got {freq_hz}")
    if method not in ("auto", "dense_direct", "dense_gmres", "aca_gmres", "mlfma"):
        raise ValueError("solve_scattering: method must be :auto, :dense_direct, :dense_gmres, :aca_gmres, or :mlfma")
    if preconditioner not in ("auto", "lu", "ilu", "diag", "none", "spai"):
        raise ValueError("solve_scattering: preconditioner must be :auto, :lu, :ilu, :diag, or :none")
    if preconditioner in ("lu", "ilu"):
        raise ValueError("solve_scattering: " + _NO_FACTOR.format(preconditioner) +
                         "; preconditioner=:auto builds the sparse approximate inverse")
    if not (math.isfinite(nf_cutoff_lambda) and nf_cutoff_lambda > 0):
        raise ValueError(f"solve_scattering: nf_cutoff_lambda must be finite and positive, got {nf_cutoff_lambda}")
    warnings = []
    lam = c0 / f
    k = 2 * math.pi * f / c0
    if isinstance(excitation, PlaneWaveExcitation):
        kn = float(np.linalg.norm(excitation.k_vec))
        if not abs(kn - k) <= 1e-8 * k:
            raise ValueError("solve_scattering: plane-wave |k_vec| does not match the frequency")
    report = mesh_resolution_report(mesh, f, points_per_wavelength, c0)
    if check_resolution and not report["meets_target"]:
        msg = (f"Mesh under-resolved: edge_max/lambda={report['edge_max_over_lambda']:.3f}, "
               f"target <= {1.0 / points_per_wavelength:.3f}. Results may be inaccurate.")
        warnings.append(msg)
        if error_on_underresolved:
            raise RuntimeError(f"solve_scattering: {msg}")
    rwg = build_rwg(mesh)
    N = rwg.nedges
    sel = method
    if method == "auto":
        sel = ("dense_direct" if N <= dense_direct_limit else "dense_gmres" if N <= dense_gmres_limit
               else "aca_gmres" if N <= mlfma_threshold else "mlfma")
    if sel in ("aca_gmres", "mlfma"):
        raise NotImplementedError(f"solve_scattering: method {sel} (N={N}) is outside the dense GPU path; "
                                  "force method='dense_gmres'")
    if N * N * 16 > max_dense_matrix_bytes:
        raise ValueError("solve_scattering dense EFIE matrix exceeds max_dense_matrix_bytes")
    if isinstance(excitation, PlaneWaveExcitation):
        v = assemble_excitation(mesh, rwg, excitation, quad_order)
    else:
        v = np.asarray(excitation, dtype=np.complex128)
        if v.shape != (N,):
            raise DimensionMismatch(f"solve_scattering: excitation length {v.shape} != N={N}")
    if not np.all(np.isfinite(v)):
        raise ValueError("solve_scattering: excitation vector must contain only finite values")
    P = None
    if sel == "dense_gmres":
        if preconditioner in ("auto", "spai"):
            P = build_nearfield_preconditioner(None, mesh, rwg, nf_cutoff_lambda * lam, "spai")
            nnz, rad, cap, capped = P.pattern_info(rwg)
            if capped:
                warnings.append(f"near-field preconditioner: sparse approximate inverse over nf_cutoff_lambda="
                                f"{nf_cutoff_lambda:g} ({P.cutoff:.4g} m) keeps the {cap} nearest bases per column "
                                f"({capped} of {N} columns cut short; mean radius {rad:.4g} m, {nnz} nonzeros)")
        elif preconditioner == "diag":
            P = build_nearfield_preconditioner(None, mesh, rwg, nf_cutoff_lambda * lam, "diag")
    method_gpu = "direct" if sel == "dense_direct" else ("mixed" if precision == "mixed" else "fp64")
    t0 = time.perf_counter()
    if isinstance(excitation, PlaneWaveExcitation):
        X, st, t_asm = solve_planewaves(mesh, rwg, k, [excitation], method_gpu, P, tol=gmres_tol,
                                        maxiter=gmres_maxiter, quad_order=quad_order,
                                        check_convergence=check_gmres_convergence)
        I = X[:, 0]
    else:
        forms = DM_Z64 | DM_Z32 if method_gpu == "mixed" else DM_Z64
        Z = assemble_Z_efie(mesh, rwg, k, quad_order=quad_order, mesh_precheck=False, max_output_bytes=None,
                            forms=forms, device=True)
        mesh.ctx.sync()
        t_asm = time.perf_counter() - t0
        if method_gpu == "direct":
            I = solve_forward(Z, v, solver="direct")
            st = SolveStats(-1, True, float("nan"))
        elif method_gpu == "mixed":
            I, st = solve_mixed(Z, v, P, tol=gmres_tol, maxiter=gmres_maxiter,
                                check_convergence=check_gmres_convergence)
        else:
            I, st = solve_gmres(Z, v, P, precond_side="right", tol=gmres_tol, maxiter=gmres_maxiter,
                                check_gmres_convergence=check_gmres_convergence)
        Z.free()
    t_total = time.perf_counter() - t0
    iters = st.niter if sel == "dense_gmres" else -1
    resid = st.residual if sel == "dense_gmres" else float("nan")
    t_pre = st.precond_time_s
    t_solve = t_total - t_asm - t_pre
    return ScatteringResult(I, sel, N, t_asm, t_solve, t_pre, iters, resid, report, warnings)


# ------------------------------------------------------ learned preconditioner
# No reference counterpart.  A 4-layer message-passing network over the RWG
# edge-adjacency graph corrects the near-field SPAI inside FGMRES:
# M r = S r + network(S r, r).  Training runs on the GPU (forward, backward and
# Adam in the library) on (residual, error) pairs from the library's own GMRES
# solves; convergence is always judged by the true FP64 residual.

def mpnn_param_count():
    return int(lib().dm_mpnn_param_count())


def mpnn_init_params(seed=0):
    """Initial parameters: W_e ~ N(0, 1/6), W_s, W_n ~ N(0, 1/(4*64)), b = 0 and
    W_d = 0 (so an untrained network leaves the SPAI output unchanged)."""
    rng = np.random.default_rng(seed)
    H, F, Ls = 64, 6, 4
    parts = [rng.standard_normal(H * F) / math.sqrt(F)]
    for _ in range(Ls):
        parts += [0.5 * rng.standard_normal(H * H) / math.sqrt(H), 0.5 * rng.standard_normal(H * H) / math.sqrt(H),
                  np.zeros(H)]
    parts.append(np.zeros(2 * H))
    out = np.concatenate(parts).astype(np.float32)
    assert out.size == mpnn_param_count()
    return out


def rwg_graph(rwg, k):
    """(nbr (N, 4) int32 0-based caller indices with -1 padding, sgn (N, 4)
    float32, geom (N, 2) float32) of the network for one mesh."""
    N = rwg.nedges
    nbr = np.zeros((N, 4), np.int32)
    sgn = np.zeros((N, 4), np.float32)
    geom = np.zeros((N, 2), np.float32)
    _check(lib().dm_rwg_graph(rwg.h, float(k), _ptr(nbr), _ptr(sgn), _ptr(geom)))
    return nbr, sgn, geom


class MPNN:
    """The network on the device (FP32 master parameters, bfloat16 copy used by
    the kernels, Adam state)."""

    def __init__(self, params=None, ctx=None, seed=0):
        self.ctx = ctx or default_context()
        p = mpnn_init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        if p.shape != (mpnn_param_count(),):
            raise ValueError(f"expected {mpnn_param_count()} parameters, got {p.shape}")
        h = C.c_void_p()
        _check(lib().dm_mpnn_create(self.ctx.h, _ptr(p), C.byref(h)))
        self.h = h

    def params(self):
        out = np.zeros(mpnn_param_count(), np.float32)
        _check(lib().dm_mpnn_get_params(self.h, _ptr(out)))
        return out

    def loss_grad(self, data, s0=0, s1=None):
        """Mean objective over samples [s0, s1) of `data` and its gradient."""
        s1 = data.nsample if s1 is None else s1
        loss = C.c_double()
        g = np.zeros(mpnn_param_count(), np.float32)
        _check(lib().dm_mpnn_loss_grad(self.ctx.h, self.h, data.h, int(s0), int(s1), C.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, data, epochs=10, batch_samples=32, lr=1e-3, seed=0):
        """Adam over shuffled mini-batches; returns the mean loss of each epoch."""
        hist = np.zeros(int(epochs))
        _check(lib().dm_mpnn_train(self.ctx.h, self.h, data.h, int(epochs), int(batch_samples), float(lr),
                                   int(seed), _ptr(hist)))
        return hist

    def apply(self, rwg, k, z, r):
        """The network's correction d(z, r) for one mesh (caller order, in the
        units of r); the preconditioner returns z + S d."""
        z = np.ascontiguousarray(z, np.complex128)
        r = np.ascontiguousarray(r, np.complex128)
        out = np.zeros_like(z)
        _check(lib().dm_mpnn_apply(self.ctx.h, self.h, rwg.h, float(k), _ptr(z), _ptr(r), _ptr(out)))
        return out

    def save(self, path):
        np.savez(path, params=self.params())

    @classmethod
    def load(cls, path, ctx=None):
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"], ctx=ctx)

    def __del__(self):
        if getattr(self, "h", None):
            lib().dm_mpnn_free(self.h)
            self.h = None


class MPNNData:
    """Training samples on the device.  Each sample is a dict with z = S r and
    r (complex, N), nbr, sgn, geom of its mesh (rwg_graph) and the mesh's
    operator A = Z S under "A" (N x N complex64, caller order; the samples of
    one mesh share one array and are consecutive), of which the device keeps
    a copy.  The objective of a sample is ||Z M r - r||^2 / ||r||^2, the
    relative residual one application of the preconditioner leaves."""

    def __init__(self, samples, ctx=None):
        self.ctx = ctx or default_context()
        for s in samples:
            if s.get("A") is None:
                raise ValueError("every training sample needs its mesh's operator A = Z S under 'A'")
        sizes = [len(s["z"]) for s in samples]
        ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        nbr = np.concatenate([np.where(s["nbr"] >= 0, s["nbr"] + off, -1) for s, off in zip(samples, ptr[:-1])])
        cat = lambda key, dt: np.ascontiguousarray(np.concatenate([s[key] for s in samples]), dt)  # noqa: E731
        nbr = np.ascontiguousarray(nbr, np.int32)
        sgn, geom = cat("sgn", np.float32), cat("geom", np.float32)
        z, r = cat("z", np.complex128), cat("r", np.complex128)
        h = C.c_void_p()
        _check(lib().dm_mpnn_data_create(self.ctx.h, int(ptr[-1]), _ptr(nbr), _ptr(sgn), _ptr(geom), _ptr(z), _ptr(r),
                                         len(samples), _ptr(ptr), C.byref(h)))
        self.h, self.nsample, self.nnode = h, len(samples), int(ptr[-1])
        a = 0
        while a < len(samples):
            A = samples[a]["A"]
            b = a + 1
            while b < len(samples) and samples[b]["A"] is A:
                b += 1
            A = np.asfortranarray(A, np.complex64)
            if A.shape != (sizes[a], sizes[a]):
                raise DimensionMismatch(f"operator shape {A.shape} != ({sizes[a]}, {sizes[a]})")
            _check(lib().dm_mpnn_data_set_matrix(self.h, a, b, A.ctypes.data_as(C.c_void_p), sizes[a]))
            a = b

    def __del__(self):
        if getattr(self, "h", None):
            lib().dm_mpnn_data_free(self.h)
            self.h = None


def residual_operator(S):
    """A = Z S of a near-field operator S (NearFieldOperator built on a matrix
    that holds DM_Z64), complex64, caller order: the operator whose residual
    the learned preconditioner is trained on."""
    out = np.zeros((S.N, S.N), np.complex64, order="F")
    _check(lib().dm_mpnn_residual_operator(S.Z.ctx.h, S.Z.h, S.h, out.ctypes.data_as(C.c_void_p)))
    return out


@dataclass
class LearnedPreconditioner:
    """FGMRES preconditioner M r = S (r + d(S r, r)), d the network's
    correction; S is the near-field SPAI over `cutoff` metres (<= 0: the
    library default)."""
    net: MPNN
    cutoff: float = 0.0


def make_training_meshes(count, seed=0, wavelength=0.1, jitter=0.1, max_sphere_subdiv=3, ctx=None):
    """Seeded random plates (sides U[0.05, 0.3] m, 4..30 cells per side, vertex
    jitter in the plane) and icospheres (2..max_sphere_subdiv subdivisions,
    radius U[0.3, 1.5] wavelengths, 3-D jitter); jitter is the standard
    deviation in mean edge lengths.  Meshes that fail the quality check are
    redrawn."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        if rng.random() < 0.5:
            Lx, Ly = rng.uniform(0.05, 0.3, 2)
            Nx, Ny = rng.integers(4, 31, 2)
            base = make_rect_plate(Lx, Ly, int(Nx), int(Ny), ctx=ctx)
            h = 0.5 * (Lx / Nx + Ly / Ny)
            xyz = base.xyz.copy()
            xyz[:, :2] += rng.normal(0.0, jitter * h, (xyz.shape[0], 2))
        else:
            sub = int(rng.integers(2, max_sphere_subdiv + 1))
            base = make_icosphere(float(rng.uniform(0.3, 1.5)) * wavelength, sub, ctx=ctx)
            e = base.xyz[base.tri[:, :2] - 1]
            h = float(np.mean(np.linalg.norm(e[:, 0] - e[:, 1], axis=1)))
            xyz = base.xyz + rng.normal(0.0, jitter * h, base.xyz.shape)
        m = TriMesh(xyz, base.tri, ctx=ctx)
        try:
            assert_mesh_quality(m)
        except Exception:
            continue
        out.append(m)
    return out


def generate_training_samples(meshes, k, iters=(1, 2, 4, 8, 16, 32), cutoff_edges=2.0, seed=0, tol=1e-10,
                              inputs="residual"):
    """Training samples from the library's GMRES on each mesh with a random
    plane wave b: inputs="residual": the converged x* (right
    SPAI-preconditioned GMRES to tol) and the iterates x_k after k = iters
    steps give the residuals r = b - Z x_k (with errors e = x* - x_k);
    inputs="arnoldi": the Arnoldi basis vectors v_k (k = iters) of the
    SPAI-preconditioned Krylov space of b, the vectors FGMRES hands its
    preconditioner.  Each sample holds r (the input), z = S r and the mesh's
    operator A = Z S (shared by its samples)."""
    if inputs not in ("residual", "arnoldi"):
        raise ValueError("inputs must be 'residual' or 'arnoldi'")
    rng = np.random.default_rng(seed)
    samples = []
    for m in meshes:
        rwg = build_rwg(m)
        Z = assemble_Z_efie(m, rwg, k, forms=DM_Z64, device=True, mesh_precheck=False, max_output_bytes=None)
        P = build_nearfield_preconditioner(None, m, rwg, cutoff_edges * float(np.mean(rwg.len)))
        S = NearFieldOperator(P, Z)
        A = residual_operator(S)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        pol = rng.standard_normal(3)
        pol -= d * (pol @ d)
        pol /= np.linalg.norm(pol)
        b = assemble_v_plane_wave(m, rwg, k * d, 1.0, pol)
        nbr, sgn, geom = rwg_graph(rwg, k)
        if inputs == "arnoldi":
            V = [b / np.linalg.norm(b)]
            for j in range(1, max(iters)):
                w = Z.matmul(S @ V[-1])
                for _ in range(2):
                    for v in V:
                        w = w - (np.vdot(v, w)) * v
                nw = np.linalg.norm(w)
                if not nw > 1e-10 * np.linalg.norm(b):
                    break
                V.append(w / nw)
            for it in iters:
                if it <= len(V):
                    v = V[it - 1]
                    samples.append(dict(z=S @ v, r=v, nbr=nbr, sgn=sgn, geom=geom, A=A))
            del S
            Z.free()
            continue
        xs, _ = solve_gmres(Z, b, preconditioner=P, precond_side="right", tol=tol, maxiter=1000)
        for it in iters:
            xk, _ = solve_gmres(Z, b, preconditioner=P, precond_side="right", tol=1e-15, maxiter=int(it),
                                check_gmres_convergence=False)
            r = b - Z.matmul(xk)
            if not np.linalg.norm(r) > 1e-12 * np.linalg.norm(b):
                continue
            samples.append(dict(z=S @ r, r=r, e=xs - xk, nbr=nbr, sgn=sgn, geom=geom, A=A))
        del S
        Z.free()
    return samples
