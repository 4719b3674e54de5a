"""The reference's torus coil as quadratures: Torus (torus.cpp:9-117, R = 2, section 0.2,
20 x 4 x 4 hexahedra, 5 tetrahedra each) -> tetrahedraToQuadratures (integration.cpp:163-187)
with the 24-point Keast rule (basis_quadratures/keast 7 *.txt, tests/golden/keast7.npz)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_PI = 3.14159265359
TETS = ((0, 1, 3, 5), (0, 3, 5, 6), (0, 2, 3, 6), (0, 4, 5, 6), (3, 5, 6, 7))   # torus.cpp:110-116


def torus_quadratures(R=2.0, width=0.2, n_len=20, n_wid=4, n_hei=4):
    k = np.load(os.path.join(GOLDEN, "keast7.npz"))
    kx, kw = k["x"], k["w"]
    inner = R - width / 2.0
    step = 360.0 / n_len * REF_PI / 180.0
    hh = width / n_hei

    def inner_point(n):
        return np.array([inner * np.cos(step * n), inner * np.sin(step * n), 0.0])

    def width_point(p, n):
        return p + p / np.linalg.norm(p) * (width / n_wid) * n

    P, W = [], []
    for lc in range(n_len):
        a, b = inner_point(lc), inner_point(lc + 1)
        for wc in range(n_wid):
            r, nr = width_point(a, wc), width_point(a, wc + 1)
            le, nle = width_point(b, wc), width_point(b, wc + 1)
            for hc in range(n_hei):
                bot, top = -width / 2.0 + hh * hc, -width / 2.0 + hh * (hc + 1)
                pts = [np.array([q[0], q[1], z]) for z in (bot, top) for q in (r, nr, le, nle)]
                for t in TETS:
                    A, B, C, D = (pts[i] for i in t)
                    vol = abs(np.dot(A - D, np.cross(B - D, C - D))) / 6.0
                    for xq, wq in zip(kx, kw):
                        P.append(A + (B - A) * xq[0] + (C - A) * xq[1] + (D - A) * xq[2])
                        W.append(vol * wq)
    return np.array(P), np.array(W)


def telma():
    d = np.load(os.path.join(GOLDEN, "telma_results.npz"))
    return d["points"], d["H"] * 1.2566370614e-6     # H (A/m) -> B (T), main.cpp:110
