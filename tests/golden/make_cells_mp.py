"""Writes tests/golden/cells_mp.json: EFIE triangle-pair terms and whole entries
of a small jittered, non-planar mesh at 30 significant digits (mpmath).

    python tests/golden/make_cells_mp.py

Written from the reference's own definitions, independently of the C++ oracle:
  * quadrature tables         src/basis/Quadrature.jl:17-54 (orders 3 and 7)
  * affine quadrature points  src/basis/Quadrature.jl:76-129
  * RWG values / divergence   src/basis/RWG.jl:819-856
  * G, G_smooth               src/basis/Greens.jl:364-400, evaluated from their
                              closed forms exp(-ikR)/(4 pi R) and
                              (exp(-ikR) - 1)/(4 pi R) (R = 0: 0 and -ik/(4 pi))
  * S(P) = int_T 1/|P - r'|   NOT the Wilton/Graglia closed form the reference
                              uses: an exact polar reduction around the foot of
                              P, one mpmath.quad per edge of
                              int [sqrt(rho_max(a)^2 + h^2) - |h|] da
  * self cell                 src/assembly/EFIE.jl:534-606 (7x7 smooth part with
                              G_smooth, 7-point outer rule, analytic S, bounded
                              remainder f_n(r') - f_n(r))
  * adjacent cell             src/assembly/EFIE.jl:611-684 (user-rule smooth
                              part, 7-point outer rule on the OUTER triangle,
                              S over the INNER one, 7x7 1/R vector term)
  * far pairs                 src/assembly/EFIE.jl:487-506
  * entry                     src/assembly/EFIE.jl:439-512 with (m, n) taken
                              as (min, max) (:443-445) and -i k eta0 (:429-437)

The mesh is a 2 x 2-cell plate of 0.02 m bent out of plane and jittered (so
neighbouring triangles are not coplanar and the two orders of an adjacent
pair give different values), with a seeded permutation of the basis order
(so the entry rule "outer = triangles of basis min(m, n)" is not the same as
"outer = lower triangle").  k = 2 pi / 0.1 (real) and a lossy complex k.

Fixture contents: mesh, RWG arrays (1-based, caller order), for each k the
pair terms I(outer half, inner half) of every self and adjacent triangle pair
in both orders, the far terms of a few far pairs, and all entries Z[m, n],
m <= n, in the reference's order rule, plus the same entries under the swapped
rule (outer = triangles of max(m, n)) so a test can check it separates them.
"""
import json
import math
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 30

ETA0 = mp.mpf("376.730313668")


def quad_rule(order):
    """src/basis/Quadrature.jl:17-54 (the reference's decimal literals)."""
    if order == 3:
        xi = [(mp.mpf(1) / 6, mp.mpf(1) / 6), (mp.mpf(2) / 3, mp.mpf(1) / 6), (mp.mpf(1) / 6, mp.mpf(2) / 3)]
        return xi, [mp.mpf(1) / 6] * 3
    if order == 7:
        a1, b1 = mp.mpf("0.059715871789770"), mp.mpf("0.470142064105115")
        a2, b2 = mp.mpf("0.797426985353087"), mp.mpf("0.101286507323456")
        xi = [(mp.mpf(1) / 3, mp.mpf(1) / 3), (b1, b1), (a1, b1), (b1, a1), (b2, b2), (a2, b2), (b2, a2)]
        w0, w1, w2 = mp.mpf("0.1125"), mp.mpf("0.0661970763942530"), mp.mpf("0.0629695902724135")
        return xi, [w0, w1, w1, w1, w2, w2, w2]
    raise ValueError(order)


def vec(a):
    return [mp.mpf(float(x)) for x in a]


def sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def norm(a):
    return mp.sqrt(dot(a, a))


def scale(a, s):
    return [a[0] * s, a[1] * s, a[2] * s]


def area(V):
    return norm(cross(sub(V[1], V[0]), sub(V[2], V[0]))) / 2


def qpoints(V, xi):
    """r = (1 - x1 - x2) v1 + x1 v2 + x2 v3 (src/basis/Quadrature.jl:76-113)."""
    return [[(1 - a - b) * V[0][c] + a * V[1][c] + b * V[2][c] for c in range(3)] for a, b in xi]


def G(r, rp, k):
    R = norm(sub(r, rp))
    if R == 0:
        return mp.mpc(0)
    return mp.exp(-1j * k * R) / (4 * mp.pi * R)


def Gs(r, rp, k):
    R = norm(sub(r, rp))
    if R == 0:
        return -1j * k / (4 * mp.pi)
    return mp.expm1(-1j * k * R) / (4 * mp.pi * R)


def S_polar(P, V):
    """int_T 1/|P - r'| dS' by the polar reduction around the foot P0 of P in
    the triangle's plane: sum over edges (A, B) of the signed sub-triangle
    (P0, A, B) integral int_a [sqrt(d^2 / cos(a)^2 + h^2) - |h|] da, with d the
    distance from P0 to the edge's line and a the angle from the foot of the
    perpendicular."""
    n = cross(sub(V[1], V[0]), sub(V[2], V[0]))
    n = scale(n, 1 / norm(n))
    h = dot(sub(P, V[0]), n)
    P0 = sub(P, scale(n, h))
    ah = abs(h)
    total = mp.mpf(0)
    for A, B in ((V[0], V[1]), (V[1], V[2]), (V[2], V[0])):
        orient = dot(cross(sub(A, P0), sub(B, P0)), n)
        if orient == 0:
            continue
        t = sub(B, A)
        t = scale(t, 1 / norm(t))
        F = [A[c] + dot(sub(P0, A), t) * t[c] for c in range(3)]
        d = norm(sub(P0, F))
        aA = mp.atan2(dot(sub(A, F), t), d)
        aB = mp.atan2(dot(sub(B, F), t), d)
        lo, hi = min(aA, aB), max(aA, aB)
        f = lambda a: mp.sqrt(d * d / mp.cos(a) ** 2 + h * h) - ah
        val = mp.quad(f, [lo, hi])
        total += val if orient > 0 else -val
    return total


class Problem:
    def __init__(self, xyz, tri, rwg, k, quad_order=3):
        self.xyz, self.tri, self.rwg, self.k = xyz, tri, rwg, k
        self.V = [[vec(xyz[v - 1]) for v in tri[t]] for t in range(len(tri))]
        self.A = [area(V) for V in self.V]
        self.xi, self.w = quad_rule(quad_order)
        self.xih, self.wh = quad_rule(7)
        self.qp = [qpoints(V, self.xi) for V in self.V]
        self.qph = [qpoints(V, self.xih) for V in self.V]
        self.S_cache = {}
        self.adj = set()
        edges = {}
        for t, tv in enumerate(tri):
            for j in range(3):
                a, b = tv[j], tv[(j + 1) % 3]
                edges.setdefault((min(a, b), max(a, b)), []).append(t)
        for ts in edges.values():
            for i in range(len(ts)):
                for j in range(len(ts)):
                    if ts[i] != ts[j]:
                        self.adj.add((ts[i], ts[j]))

    def half(self, n, side):
        """(triangle 0-based, f at user-rule points, f at 7 points, div) of basis
        n's side (0 = T+, 1 = T-), src/basis/RWG.jl:819-856."""
        r = self.rwg
        L = mp.mpf(float(r["len"][n]))
        if side == 0:
            t = r["tplus"][n] - 1
            p = vec(self.xyz[r["vplus_opp"][n] - 1])
            c = L / (2 * mp.mpf(float(r["area_plus"][n])))
            f = lambda q: scale(sub(q, p), c)
            dv = L / mp.mpf(float(r["area_plus"][n]))
        else:
            t = r["tminus"][n] - 1
            p = vec(self.xyz[r["vminus_opp"][n] - 1])
            c = L / (2 * mp.mpf(float(r["area_minus"][n])))
            f = lambda q: scale(sub(p, q), c)
            dv = -L / mp.mpf(float(r["area_minus"][n]))
        return t, [f(q) for q in self.qp[t]], [f(q) for q in self.qph[t]], dv

    def S(self, tP, qi, tT):
        key = (tP, qi, tT)
        if key not in self.S_cache:
            self.S_cache[key] = S_polar(self.qph[tP][qi], self.V[tT])
        return self.S_cache[key]

    def pair_term(self, m, sm, n, sn):
        """The triangle-pair term of _efie_entry for basis m's side sm as the
        OUTER triangle and basis n's side sn as the inner one (before -i k eta0)."""
        k, w, wh = self.k, self.w, self.wh
        tm, fm, fmh, dvm = self.half(m, sm)
        tn, fn, fnh, dvn = self.half(n, sn)
        Am, An = self.A[tm], self.A[tn]
        dv = dvm * dvn / (k * k)
        inv4pi = 1 / (4 * mp.pi)
        val = mp.mpc(0)
        if tm == tn:  # src/assembly/EFIE.jl:534-606
            q = self.qph[tm]
            for a in range(7):
                for b in range(7):
                    g = Gs(q[a], q[b], k)
                    val += (dot(fmh[a], fnh[b]) * g - dv * g) * wh[a] * wh[b] * (2 * Am) * (2 * Am)
            for a in range(7):
                inner = inv4pi * self.S(tm, a, tm)
                lead = dot(fmh[a], fnh[a]) * inner
                rem = mp.mpf(0)
                for b in range(7):
                    R = norm(sub(q[a], q[b]))
                    if R == 0:
                        continue
                    rem += dot(fmh[a], sub(fnh[b], fnh[a])) * (inv4pi / R) * wh[b] * (2 * Am)
                val += ((lead + rem) - dv * inner) * wh[a] * (2 * Am)
        elif (tm, tn) in self.adj:  # src/assembly/EFIE.jl:611-684
            qm, qn = self.qp[tm], self.qp[tn]
            for a in range(len(w)):
                for b in range(len(w)):
                    g = Gs(qm[a], qn[b], k)
                    val += (dot(fm[a], fn[b]) * g - dv * g) * w[a] * w[b] * (2 * Am) * (2 * An)
            qmh, qnh = self.qph[tm], self.qph[tn]
            for a in range(7):
                inner = inv4pi * self.S(tm, a, tn)
                vs = mp.mpf(0)
                for b in range(7):
                    R = norm(sub(qmh[a], qnh[b]))
                    vs += dot(fmh[a], fnh[b]) * (inv4pi / R) * wh[b] * (2 * An)
                val += (vs - dv * inner) * wh[a] * (2 * Am)
        else:  # src/assembly/EFIE.jl:487-506
            qm, qn = self.qp[tm], self.qp[tn]
            for a in range(len(w)):
                for b in range(len(w)):
                    g = G(qm[a], qn[b], k)
                    val += (dot(fm[a], fn[b]) * g - dv * g) * w[a] * w[b] * (2 * Am) * (2 * An)
        return val

    def entry(self, m, n, swapped=False):
        """Z[m, n] with the reference's rule (outer = basis min(m, n)), or the
        swapped rule (outer = basis max(m, n)) when swapped."""
        a, b = (min(m, n), max(m, n))
        if swapped:
            a, b = b, a
        val = mp.mpc(0)
        for sa in range(2):
            for sb in range(2):
                val += self.pair_term(a, sa, b, sb)
        return -1j * self.k * ETA0 * val


def make_mesh(seed=20261203):
    """2 x 2-cell plate, 0.02 m, bent (z = 0.25 (x^2 - 0.6 y^2) / L) and jittered
    (0.1 edge in every coordinate); vertex and face order of make_rect_plate
    (src/geometry/Mesh.jl:355-392)."""
    rng = np.random.default_rng(seed)
    Lx = Ly = 0.02
    nx = ny = 2
    xyz = []
    for iy in range(ny + 1):
        for ix in range(nx + 1):
            x = Lx * (ix / nx - 0.5)
            y = Ly * (iy / ny - 0.5)
            xyz.append([x, y, 0.25 * (x * x - 0.6 * y * y) / Lx])
    xyz = np.array(xyz) + rng.normal(0.0, 0.1 * Lx / nx, (len(xyz), 3))
    tri = []
    for iy in range(ny):
        for ix in range(nx):
            v1 = iy * (nx + 1) + ix + 1
            v2, v3, v4 = v1 + 1, v1 + nx + 2, v1 + nx + 1
            tri += [[v1, v2, v3], [v1, v3, v4]]
    return xyz, np.array(tri, dtype=np.int64)


def build_rwg(xyz, tri, perm_seed=7):
    """One basis per interior edge, T+ = lower-index triangle, opposite
    vertices, FP64 lengths and areas (src/basis/RWG.jl:102-143), then a seeded
    permutation of the basis order."""
    edges = {}
    for t, tv in enumerate(tri):
        for j in range(3):
            a, b = int(tv[j]), int(tv[(j + 1) % 3])
            edges.setdefault((min(a, b), max(a, b)), []).append(t)

    def opp(t, a, b):
        return int([v for v in tri[t] if v != a and v != b][0])

    def farea(t):
        p = xyz[tri[t] - 1]
        return 0.5 * float(np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0])))

    out = {k: [] for k in ("tplus", "tminus", "vplus_opp", "vminus_opp", "len", "area_plus", "area_minus", "evert")}
    for (a, b), ts in sorted(edges.items()):
        if len(ts) != 2:
            continue
        tp, tm = min(ts), max(ts)
        out["tplus"].append(tp + 1)
        out["tminus"].append(tm + 1)
        out["vplus_opp"].append(opp(tp, a, b))
        out["vminus_opp"].append(opp(tm, a, b))
        out["len"].append(float(np.linalg.norm(xyz[a - 1] - xyz[b - 1])))
        out["area_plus"].append(farea(tp))
        out["area_minus"].append(farea(tm))
        out["evert"].append([a, b])
    n = len(out["tplus"])
    perm = np.random.default_rng(perm_seed).permutation(n)
    return {k: [v[i] for i in perm] for k, v in out.items()}


def cstr(z):
    return [mp.nstr(mp.re(z), 30), mp.nstr(mp.im(z), 30)]


def main():
    xyz, tri = make_mesh()
    rwg = build_rwg(xyz, tri)
    N = len(rwg["tplus"])
    fixture = {"about": "EFIE pair terms and entries at 30 digits, see make_cells_mp.py",
               "xyz": xyz.tolist(), "tri": tri.tolist(), "rwg": rwg, "quad_order": 3,
               "eta0": "376.730313668", "cases": []}
    k0 = 2 * math.pi / 0.1
    for kr, ki in ((k0, 0.0), (k0, -0.35 * k0)):
        k = mp.mpc(kr, ki)
        P = Problem(xyz, tri.tolist(), rwg, k)
        halves = [(n, s, P.half(n, s)[0]) for n in range(N) for s in range(2)]
        terms = []
        nfar = 0
        for (m, sm, tm) in halves:
            for (n, sn, tn) in halves:
                kind = "self" if tm == tn else "adjacent" if (tm, tn) in P.adj else "far"
                if kind == "far":
                    if nfar >= 12 or (m + n + sm + sn) % 3:
                        continue
                    nfar += 1
                terms.append({"m": m, "sm": sm, "n": n, "sn": sn, "outer": tm + 1, "inner": tn + 1, "kind": kind,
                              "value": cstr(P.pair_term(m, sm, n, sn))})
        entries = []
        for m in range(N):
            for n in range(m, N):
                entries.append({"m": m, "n": n, "value": cstr(P.entry(m, n)),
                                "swapped": cstr(P.entry(m, n, swapped=True))})
        fixture["cases"].append({"k": [kr, ki], "terms": terms, "entries": entries})
        print(f"k = {kr:.4f}{ki:+.4f}i: {len(terms)} pair terms, {len(entries)} entries", flush=True)
    with open(os.path.join(HERE, "cells_mp.json"), "w") as f:
        json.dump(fixture, f, indent=0)


if __name__ == "__main__":
    main()
