"""Writes the golden fixtures under tests/golden/ from data files the reference holds.

Run once in a container where the reference is mounted (read-only) at /root/reference:
    python tests/golden/make_fixtures.py
It only parses data files (no reference code is imported or executed):
  FMM/FMMGPU/matrices/regular_{0..7}.txt   complex M2M matrices, p = 10, float build
  FMM/FMMGPU/results/telma_results.txt      external field solution for the torus coil
  FMM/FMMGPU/basis_quadratures/keast 7 {x,w}.txt   24-point tetrahedron rule
"""
import os
import re

import numpy as np

REF = "/root/reference/FMM/FMMGPU"
OUT = os.path.dirname(os.path.abspath(__file__))


def parse_complex_matrix(fn):
    rows = []
    for line in open(fn):
        if line.strip():
            rows.append([complex(float(a), float(b)) for a, b in re.findall(r"\(([^,]+),([^)]+)\)", line)])
    return np.array(rows)


def main():
    data = {}
    for o in range(8):
        M = parse_complex_matrix(os.path.join(REF, "matrices", f"regular_{o}.txt"))
        r, c = np.nonzero(M)
        data[f"o{o}_row"] = r.astype(np.int16)
        data[f"o{o}_col"] = c.astype(np.int16)
        data[f"o{o}_re"] = M[r, c].real.astype(np.float64)
        data[f"o{o}_im"] = M[r, c].imag.astype(np.float64)
        data[f"o{o}_shape"] = np.array(M.shape, np.int16)
    np.savez_compressed(os.path.join(OUT, "m2m_regular_p10.npz"), **data)

    lines = open(os.path.join(REF, "results", "telma_results.txt")).read().split()
    n = int(lines[3])
    vals = np.array([float(v) for v in lines[4:4 + 7 * n]]).reshape(n, 7)
    np.savez_compressed(os.path.join(OUT, "telma_results.npz"), points=vals[:, 1:4], H=vals[:, 4:7])

    x = np.loadtxt(os.path.join(REF, "basis_quadratures", "keast 7 x.txt")).reshape(-1, 3)
    w = np.loadtxt(os.path.join(REF, "basis_quadratures", "keast 7 w.txt")).reshape(-1)
    np.savez_compressed(os.path.join(OUT, "keast7.npz"), x=x, w=w)


if __name__ == "__main__":
    main()
