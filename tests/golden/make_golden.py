"""Write the small golden fixtures under tests/golden/ (run from the repo root).

The reference ships no test data (no test suite, no fixtures), so the
fixtures are computed here from the mathematics, independently of the C
oracle:

* chi0_tiny.npz  -- a seeded HILBERT NONE chi0 case, expected output by a
  brute-force extended-precision (longdouble) triple sum with the weight
  written in the BASELINE spelling D = 2 w_k (e_n - e_n')/((e_n - e_n')^2
  + w^2) rather than the reference's -2 d/(w^2 + d^2);
* gaunt_table.npz -- Gaunt coefficients for l <= 4 from sympy's exact
  gaunt() under the convention map
  ref_gaunt(l1,l2,l3,m1,m2,m3) = (-1)^(m1+m3) sympy_gaunt(l1,l2,l3,-m1,m2,-m3);
* inverse_small.npz -- an HPD packed matrix and its inverse from LAPACK
  (scipy zpotrf/zpotri).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def chi0_tiny():
    g = np.random.Generator(np.random.PCG64(20261202))
    nk, nocc, nunocc, nb, nw = 2, 3, 5, 7, 3
    eo = np.sort(g.uniform(-0.5, 0.0, (nk, nocc)), axis=1)
    eu = np.sort(g.uniform(0.05, 3.0, (nk, nunocc)), axis=1)
    n = nk * nocc * nunocc
    M = (g.standard_normal((n, nb)) + 1j * g.standard_normal((n, nb))) / np.sqrt(nb)
    u, _ = np.polynomial.legendre.leggauss(nw)
    x = 0.5 * (u + 1)
    om = x / (1 - x)
    L = np.longdouble
    wk = L(1) / L(nk)
    chi = np.zeros((nw, nb, nb), np.clongdouble)
    t = 0
    D = np.zeros((n, nw), np.float64)
    for k in range(nk):
        for b1 in range(nocc):
            for b2 in range(nunocc):
                de = L(eo[k, b1]) - L(eu[k, b2])  # e_n - e_n'
                for f in range(nw):
                    d = 2 * wk * de / (de * de + L(om[f]) ** 2)
                    D[t, f] = float(d)
                    ci = M[t].astype(np.clongdouble)
                    chi[f] += d * np.outer(ci, np.conj(ci))
                t += 1
    iu = np.concatenate([np.arange(j + 1) for j in range(nb)])
    ju = np.concatenate([np.full(j + 1, j) for j in range(nb)])
    packed = np.stack([chi[f][iu, ju] for f in range(nw)]).astype(np.complex128)
    np.savez(OUT / "chi0_tiny.npz", nk=nk, nocc=nocc, nunocc=nunocc, nb=nb, nw=nw,
             e_occ=eo, e_unocc=eu, M=M, omega=om, D=D, chi_packed=packed)


def gaunt_table():
    from sympy.physics.wigner import gaunt

    rows = []
    L = 4
    for l1 in range(L + 1):
        for l2 in range(L + 1):
            for l3 in range(L + 1):
                for m1 in range(-l1, l1 + 1):
                    for m2 in range(-l2, l2 + 1):
                        m3 = m2 - m1
                        if abs(m3) > l3:
                            continue
                        v = float(((-1) ** (m1 + m3)) * gaunt(l1, l2, l3, -m1, m2, -m3))
                        rows.append((l1, l2, l3, m1, m2, m3, v))
    a = np.array(rows)
    np.savez(OUT / "gaunt_table.npz", idx=a[:, :6].astype(np.int32), val=a[:, 6])


def inverse_small():
    from scipy.linalg import lapack

    g = np.random.Generator(np.random.PCG64(7))
    nb = 9
    a = g.standard_normal((nb, nb)) + 1j * g.standard_normal((nb, nb))
    h = a @ a.conj().T + nb * np.eye(nb)
    c, info = lapack.zpotrf(h, lower=0)
    assert info == 0
    inv, info = lapack.zpotri(c, lower=0)
    assert info == 0
    iu = np.concatenate([np.arange(j + 1) for j in range(nb)])
    ju = np.concatenate([np.full(j + 1, j) for j in range(nb)])
    np.savez(OUT / "inverse_small.npz", packed=h[iu, ju], inv_packed=inv[iu, ju])


if __name__ == "__main__":
    chi0_tiny()
    gaunt_table()
    inverse_small()
    print("fixtures written to", OUT)
