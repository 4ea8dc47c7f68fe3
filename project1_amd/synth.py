"""Seeded synthetic inputs for the polarisation-function path.

There is no DFT data in this environment, so every workload is generated
from a seed with numpy's PCG64 generator (bit-identical on every host):

* energies: occupied ~ U[-0.5, 0] Ha, unoccupied ~ U[0.05, 3] Ha, sorted
  per k (the unoccupied set stands for the states at k+q);
* imaginary frequencies: Gauss-Legendre nodes x on [0, 1] mapped to
  omega = x/(1-x), quadrature weights w/(2(1-x)^2);
* products M: complex128 with Re, Im ~ N(0,1)/sqrt(nb), stored (n, nb)
  with t = b2 + nunocc*(b1 + nocc*k);
* Coulomb matrix: V = Q diag(lam) Q^T with Q Haar-orthogonal and lam
  log-uniform in [1e-3, 1e3], so V^(1/2) = Q diag(sqrt(lam)) Q^T exactly;
* wave functions for the product kernels (plane-wave coefficients damped by
  exp(-|G|^2/8), muffin-tin coefficients, radial-integral tensors).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def gl_frequencies(nw: int):
    """Gauss-Legendre nodes on [0,1] mapped to omega = x/(1-x)."""
    u, wu = np.polynomial.legendre.leggauss(nw)
    x = 0.5 * (u + 1.0)
    wx = 0.5 * wu
    omega = x / (1.0 - x)
    wq = wx / (1.0 - x) ** 2
    return omega, wq


def mesh_frequencies(nfreq: int, freqmax: float):
    """SPEX MESH grid, correlation.f:2486-2496 (no additional points)."""
    f = np.empty(nfreq)
    for i in range(1, nfreq + 1):
        r = 0.9 * (i - 1) / (nfreq - 1) if freqmax > 0 else 1.0 * (i - 1) / nfreq
        f[i - 1] = r / (1 - r)
    if nfreq > 1:
        f *= abs(freqmax) / f[-1]
    return f


def energies(g: np.random.Generator, nk: int, nocc: int, nunocc: int):
    eo = np.sort(g.uniform(-0.5, 0.0, size=(nk, nocc)), axis=1)
    eu = np.sort(g.uniform(0.05, 3.0, size=(nk, nunocc)), axis=1)
    return eo, eu


def products(g: np.random.Generator, n: int, nb: int) -> np.ndarray:
    """(n, nb) complex128, C-contiguous (transition-major)."""
    s = 1.0 / np.sqrt(nb)
    m = np.empty((n, nb), np.complex128)
    m.real = g.standard_normal((n, nb)) * s
    m.imag = g.standard_normal((n, nb)) * s
    return m


def coulomb(g: np.random.Generator, nb: int, lo: float = 1e-3, hi: float = 1e3):
    """Random SPD V and its exact square root (both real symmetric)."""
    a = g.standard_normal((nb, nb))
    q, r = np.linalg.qr(a)
    q *= np.sign(np.diag(r))
    lam = np.exp(g.uniform(np.log(lo), np.log(hi), size=nb))
    vhalf = (q * np.sqrt(lam)) @ q.T
    vhalf = 0.5 * (vhalf + vhalf.T)
    v = (q * lam) @ q.T
    return v, vhalf, lam


@dataclass
class Chi0Case:
    nk: int
    nocc: int
    nunocc: int
    nb: int
    nw: int
    M: np.ndarray      # (n, nb) complex128, C order
    e_occ: np.ndarray  # (nk, nocc)
    e_unocc: np.ndarray  # (nk, nunocc) at k+q
    omega: np.ndarray  # (nw,)
    wq: np.ndarray     # quadrature weights

    @property
    def n(self) -> int:
        return self.nk * self.nocc * self.nunocc


def chi0_case(nk=8, nocc=8, nunocc=56, nb=256, nw=6, seed=1) -> Chi0Case:
    """The CPU reference case: 2x2x2 k mesh, 8/56 bands, nb=256, 6 frequencies."""
    g = rng(seed)
    eo, eu = energies(g, nk, nocc, nunocc)
    M = products(g, nk * nocc * nunocc, nb)
    om, wq = gl_frequencies(nw)
    return Chi0Case(nk, nocc, nunocc, nb, nw, M, eo, eu, om, wq)


def weights_none(e_occ, e_unocc, omega, nkpt=None):
    """Closed-form HILBERT NONE weights (n, nw), used by host-side callers.

    Same expression and operation order as the device kernel:
    -2*d/(w*w + d*d)/nkpt/nsym1*nkpts with nsym1 = nkpts = 1.
    """
    nk, nocc = e_occ.shape
    nunocc = e_unocc.shape[1]
    nkpt = nk if nkpt is None else nkpt
    d = (e_unocc[:, None, :] - e_occ[:, :, None]).reshape(-1)
    w = np.asarray(omega)[None, :]
    return (-2 * d[:, None] / (w * w + d[:, None] * d[:, None]) / nkpt / 1) * 1.0
