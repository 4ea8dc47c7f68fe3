"""Transformation of the products to the Coulomb eigenbasis -- TEST
INFRASTRUCTURE, part of the oracle; only tests/ may import it.

Restates susceptibility.f:806-832 (noinv build):

  cprod = conjg(cprod);  cprod(:, :dim) = cprod . T;  cprod = conjg(cprod)

with T = ctrafo(:, :dim) when the products are <M phi|phi> (fullpw), else
T = ctrafo1: the muffin-tin rows of ctrafo as they are and every
interstitial column passed through olapvec (overlap.f:298-315) over the
mixed-basis plane waves gptm of the q point:

  ctrafo1(nbasp + i, c) = sum_j conj(cstep(G_j - G_i)) ctrafo(nbasp + j, c).

The reference does the product in row blocks of dim/4 transitions; the
result does not depend on the blocking.  Pinned in
tests/test_oracle_trafo.py by chi0 in the new basis being T^H chi0 T of the
old one, by the identity for theta = 1, and by olapvec against a real-space
integral.
"""
import numpy as np

from oracle.momentum import olapvec


def ctrafo_olap(ctrafo, nbasp, gptm, cstep, cext):
    """ctrafo1 of susceptibility.f:815-819."""
    T = np.array(ctrafo, complex)
    G = np.asarray(gptm, np.int64)
    out = T.copy()
    for c in range(T.shape[1]):
        out[nbasp:, c] = olapvec(T[nbasp:, c], G, cstep, cext)
    return out


def coulomb_trafo(cprod, T, dim):
    """The products in the Coulomb eigenbasis, (n, dim): conj(conj(cprod) T)."""
    c = np.conj(np.asarray(cprod, complex))
    out = c @ np.asarray(T, complex)[:, :dim]
    return np.conj(out)
