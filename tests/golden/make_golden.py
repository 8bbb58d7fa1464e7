"""Writes the committed golden fixtures under tests/golden/.

    python tests/golden/make_golden.py

* test44.json    — the reference's own golden values for the closed-form 1/R
                   integral (test/runtests.jl:14060,14134), copied as data.
* green_mp.json  — G and G_smooth at seeded points from mpmath at 50 digits,
                   from the closed forms exp(-ikR)/(4 pi R) and (exp(-ikR)-1)/(4 pi R).
* int1overR_mp.json — S(P) = int_T 1/|P - r'| dS' for seeded triangles / points:
                   in-plane points by the exact polar reduction (apex-P sub-triangles,
                   int rho_max(theta) dtheta, mpmath.quad), off-plane by mpmath 2-D quadrature.
* c0_plate.npz   — the 0.1 m plate problem (make_rect_plate(0.1,0.1,6,6), k=2pi/0.1,
                   quad 3, normal incidence x-pol): RWG arrays, Z, v, I from the oracle.
"""
import json
import math
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

mp.mp.dps = 50


def green_points(seed=20261202, count=20):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        r = rng.uniform(-0.2, 0.2, 3)
        rp = rng.uniform(-0.2, 0.2, 3)
        if i % 4 == 0:
            rp = r + rng.uniform(-1e-3, 1e-3, 3)  # |kR| < 0.5: exprel branch
        k = 2 * math.pi / 0.1 * rng.uniform(0.5, 2.0)
        ki = 0.0 if i % 3 else -rng.uniform(0.0, 5.0)
        kk = mp.mpc(k, ki)
        R = mp.sqrt(sum((mp.mpf(a) - mp.mpf(b)) ** 2 for a, b in zip(r, rp)))
        G = mp.exp(-1j * kk * R) / (4 * mp.pi * R)
        Gs = (mp.exp(-1j * kk * R) - 1) / (4 * mp.pi * R)
        out.append(dict(r=r.tolist(), rp=rp.tolist(), k=[k, ki], G=[float(G.real), float(G.imag)],
                        Gs=[float(Gs.real), float(Gs.imag)]))
    return out


def s_inplane(P, A, B, C):
    """int_T 1/|P - r'| for P in the triangle plane: sum over the three
    sub-triangles (P, X, Y) of int_0^alpha rho(theta) dtheta, rho the distance
    from P to the line XY along theta (signed by orientation)."""
    P, A, B, C = (mp.matrix([mp.mpf(float(v)) for v in x]) for x in (P, A, B, C))
    n = _cross(B - A, C - A)
    n = n / mp.norm(n)
    total = mp.mpf(0)
    for X, Y in ((A, B), (B, C), (C, A)):
        u, v = X - P, Y - P
        su = _dot(_cross(u, v), n)  # signed double area of (P, X, Y)
        if su == 0:
            continue
        lu, lv = mp.norm(u), mp.norm(v)
        alpha = mp.acos(max(-1, min(1, _dot(u, v) / (lu * lv))))
        e = Y - X
        d = abs(su) / mp.norm(e)  # distance from P to the line XY
        # angle at P between u and the foot of the perpendicular
        ue = _dot(u, e) / mp.norm(e)
        phi0 = mp.atan2(ue, d)  # signed angle of u from the perpendicular, toward e
        val = mp.quad(lambda t: d / mp.cos(phi0 + t), [0, alpha])
        total += mp.sign(su) * val
    return float(abs(total))


def s_offplane(P, A, B, C):
    P, A, B, C = (np.asarray(x, float) for x in (P, A, B, C))
    Pm = [mp.mpf(v) for v in P]
    Am, Bm, Cm = ([mp.mpf(v) for v in x] for x in (A, B, C))
    area2 = mp.norm(_cross(mp.matrix(Bm) - mp.matrix(Am), mp.matrix(Cm) - mp.matrix(Am)))

    def f(s, t):  # Duffy-free: (s, t) in unit square -> triangle via r = A + s(B-A) + s t (C-B)
        r = [Am[i] + s * (Bm[i] - Am[i]) + s * t * (Cm[i] - Bm[i]) for i in range(3)]
        R = mp.sqrt(sum((Pm[i] - r[i]) ** 2 for i in range(3)))
        return s * area2 / R

    return float(mp.quad(f, [0, 1], [0, 1]))


def _cross(a, b):
    return mp.matrix([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def int1overR_cases(seed=7):
    rng = np.random.default_rng(seed)
    tris = [np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.]]),
            np.array([[0.01, 0.02, 0.0], [0.014, 0.021, 0.001], [0.011, 0.026, -0.002]]),
            np.array([[0.3, -0.1, 0.2], [0.1, 0.2, 0.25], [-0.2, 0.0, 0.1]])]
    out = []
    for T in tris:
        n = np.cross(T[1] - T[0], T[2] - T[0])
        n /= np.linalg.norm(n)
        scale = np.linalg.norm(T[1] - T[0])
        for j in range(6):
            b = rng.dirichlet([2, 2, 2])
            P = b @ T
            if j % 2 == 0:
                S = s_inplane(P, *T)
                kind = "inplane"
            else:
                P = P + n * scale * rng.uniform(0.05, 0.8) * (1 if j % 4 == 1 else -1)
                S = s_offplane(P, *T)
                kind = "offplane"
            out.append(dict(P=P.tolist(), T=T.tolist(), S=S, kind=kind))
    return out


def c0_fixture():
    from oracle import oracle as O
    m = O.make_rect_plate(0.1, 0.1, 6, 6)
    r = O.build_rwg(m)
    k = 2 * math.pi / 0.1
    Z = O.assemble_Z_efie(r, k)
    v = O.assemble_v_plane_wave(r, [0, 0, -k], 1.0, [1, 0, 0])
    I = O.solve_direct(Z, v)
    np.savez_compressed(os.path.join(HERE, "c0_plate.npz"), xyz=m.xyz, tri=m.tri, tplus=r.tplus, tminus=r.tminus,
                        vplus_opp=r.vplus_opp, vminus_opp=r.vminus_opp, len=r.len, area_plus=r.area_plus,
                        area_minus=r.area_minus, Z=Z, v=v, I=I, k=k)


if __name__ == "__main__":
    with open(os.path.join(HERE, "test44.json"), "w") as f:
        json.dump({"source": "test/runtests.jl:14060-14061,14134-14135",
                   "S_centroid_unit_right_triangle": 2.4072299231640093,
                   "S_centroid_h0p5": 0.8515737774153572, "tol": 1e-10}, f, indent=1)
    with open(os.path.join(HERE, "green_mp.json"), "w") as f:
        json.dump(green_points(), f, indent=1)
    with open(os.path.join(HERE, "int1overR_mp.json"), "w") as f:
        json.dump(int1overR_cases(), f, indent=1)
    c0_fixture()
    print("fixtures written")
