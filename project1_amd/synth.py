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


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def device_normal(count: int, seed: int) -> np.ndarray:
    """Host mirror of the device generator (spexhip_synth_normal_dev):
    complex draw i = (value 2i, value 2i+1) is one Box-Muller pair from the
    splitmix64 hashes of counters 2i, 2i+1.  Agrees with the device to a few
    ulp (log, sqrt and sincospi round differently)."""
    with np.errstate(over="ignore"):
        base = _splitmix64(np.array([np.uint64(seed) ^ np.uint64(0xD1B54A32D192ED03)]))[0]
        i = np.arange(count, dtype=np.uint64)
        h1 = _splitmix64(base + np.uint64(2) * i)
        h2 = _splitmix64(base + np.uint64(2) * i + np.uint64(1))
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 1.0) / 9007199254740992.0
    u2 = ((h2 >> np.uint64(11)).astype(np.float64) + 1.0) / 9007199254740992.0
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)


def gl_frequencies(nw: int):
    """Gauss-Legendre nodes on [0,1] mapped to omega = x/(1-x)."""
    u, wu = np.polynomial.legendre.leggauss(nw)
    x = 0.5 * (u + 1.0)
    wx = 0.5 * wu
    omega = x / (1.0 - x)
    wq = wx / (1.0 - x) ** 2
    return omega, wq


def mesh_frequencies(nfreq: int, freqmax: float, freqadd: int = 0):
    """SPEX MESH grid on the imaginary axis (keyword ``MESH n[+a] wmax``,
    correlation.f:2478-2507): x_i = 0.9(i-1)/(n-1) (wmax > 0) or (i-1)/n
    (wmax < 0), w = x/(1-x) rescaled so that the last point is |wmax|; with
    ``+a`` the m-th of a rounds puts m equally spaced points into the
    interval (w_{a-m+1}, w_{a-m+2}) (1-based), and the mesh is sorted."""
    if nfreq <= 0:
        raise ValueError("MESH: the number of frequencies must be positive")
    if nfreq == 1 and freqmax > 0:
        # the reference's 0.9*(i-1)/(nfreq-1) is 0/0 here
        raise ValueError("MESH: one frequency needs wmax < 0")
    f = np.empty(nfreq + freqadd * (freqadd + 1) // 2)
    for i in range(1, nfreq + 1):
        r = 0.9 * (i - 1) / (nfreq - 1) if freqmax > 0 else 1.0 * (i - 1) / nfreq
        f[i - 1] = r / (1 - r)
    if nfreq > 1:
        f[:nfreq] *= abs(freqmax) / f[nfreq - 1]
    j = nfreq
    for m in range(1, freqadd + 1):
        lo, hi = f[freqadd - m], f[freqadd - m + 1]  # freq(a-m+1), freq(a-m+2), 1-based
        for i in range(1, m + 1):
            f[j] = lo + (hi - lo) / (m + 1) * i
            j += 1
    return np.sort(f)


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


@dataclass
class ProductSystem:
    """Synthetic inputs of the wave-function products for one q-point."""

    nkpt: int
    kpt: np.ndarray      # (nkpt, 3) fractional
    q: np.ndarray        # (3,) = kpt[iq]
    k1: np.ndarray       # (nk,) packet k indices (0-based)
    k2: np.ndarray       # (nk,) index of k1+q folded into the mesh
    g2: np.ndarray       # (nk, 3) umklapp nint(kpt(k2) - q - kpt(k1))
    nband: int
    # plane waves
    gpt: np.ndarray      # (nkpt, maxgpt, 3)
    ngpt: np.ndarray     # (nkpt,)
    cpw: np.ndarray      # (maxgpt, nband, nkpt) complex, Fortran order
    gptm: np.ndarray     # (ngptm, 3)
    nfft: tuple
    vol: float
    # muffin tin
    types: list          # dicts: lcut, nindx, lcutm, nindxm, neq, integral
    cent: np.ndarray     # (ncent, 3)
    cmt: np.ndarray      # (maxlm, ncent, nband, nkpt) complex, Fortran order

    @property
    def nbasp(self):
        return sum(T["neq"] * sum((2 * l + 1) * T["nindxm"][l] for l in range(T["lcutm"] + 1))
                   for T in self.types)

    @property
    def ngptm(self):
        return self.gptm.shape[0]

    @property
    def maxgpt(self):
        return self.gpt.shape[1]


def kmesh(n: int):
    g = np.arange(n) / n
    kx, ky, kz = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)


def kptsum_index(kpt, n, a, b):
    """Index of kpt[a] + kpt[b] folded into [0,1)^3 (vector.f:107-124)."""
    v = np.mod(kpt[a] + kpt[b], 1.0)
    g = np.rint(v * n).astype(int) % n
    return int(g[0] * n * n + g[1] * n + g[2])


def integral_dims(lcut, nindx, lcutm, nindxm):
    """Extents of integral(ln, ln1, lln), wavefproducts.f:200-222."""
    lm = sum(nindx[: lcut + 1])
    lln = sum(nindxm[: lcutm + 1])
    lm1 = 0
    for ll in range(lcutm + 1):
        for l in range(lcut + 1):
            s = sum(nindx[l1] for l1 in range(abs(ll - l), min(ll + l, lcut) + 1, 2))
            lm1 = max(lm1, s)
    return lm, lm1, lln


def product_system(nmesh=2, nband=20, gcut=3.1, ngptm=None, gcutm=2.1, nfft=(16, 16, 16),
                   ntype=2, lcut=2, nrad=2, lcutm=2, nradm=2, iq=1, seed=5,
                   damping=8.0) -> ProductSystem:
    """Seeded product inputs on an nmesh^3 k mesh (simple cubic lattice).

    Plane waves: G with |k+G| <= gcut, coefficients CN(0,1) exp(-|G|^2/damping).
    Mixed-basis IPWs: the ngptm shortest G (or |G| <= gcutm).  MT: ntype
    types with one atom each, lcut/nrad radial functions, random radial
    integrals ~N(0,1)/10 and coefficients CN(0,1)/sqrt(maxlm).
    """
    g = rng(seed)
    kpt = kmesh(nmesh)
    nkpt = kpt.shape[0]
    q = kpt[iq].copy()
    k1 = np.arange(nkpt, dtype=np.int32)
    k2 = np.array([kptsum_index(kpt, nmesh, iq, k) for k in k1], np.int32)
    g2 = np.rint(kpt[k2] - q - kpt[k1]).astype(np.int32)
    r = int(np.ceil(gcut)) + 1
    ax = np.arange(-r, r + 1)
    G = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    lists = []
    for k in range(nkpt):
        kk = kpt[k] + G
        sel = G[np.einsum("ij,ij->i", kk, kk) <= gcut**2]
        lists.append(sel)
    maxgpt = max(len(s) for s in lists)
    gpt = np.zeros((nkpt, maxgpt, 3), np.int32)
    ngpt = np.zeros(nkpt, np.int32)
    cpw = np.zeros((maxgpt, nband, nkpt), np.complex128, order="F")
    for k, s in enumerate(lists):
        ngpt[k] = len(s)
        gpt[k, : len(s)] = s
        damp = np.exp(-np.einsum("ij,ij->i", s, s) / damping)
        c = (g.standard_normal((len(s), nband)) + 1j * g.standard_normal((len(s), nband)))
        cpw[: len(s), :, k] = c * damp[:, None]
    n2 = np.einsum("ij,ij->i", G, G)
    order = np.lexsort((G[:, 2], G[:, 1], G[:, 0], n2))
    if ngptm is None:
        gptm = G[order][n2[order] <= gcutm**2]
    else:
        gptm = G[order][:ngptm]
    types = []
    maxlm = 0
    for it in range(ntype):
        nindx = np.full(lcut + 1, nrad, np.int32)
        nindxm = np.full(lcutm + 1, nradm, np.int32)
        d = integral_dims(lcut, nindx, lcutm, nindxm)
        integ = np.asfortranarray(g.standard_normal(d) / 10)
        types.append(dict(lcut=lcut, nindx=nindx, lcutm=lcutm, nindxm=nindxm, neq=1, integral=integ))
        maxlm = max(maxlm, int(sum((2 * l + 1) * nindx[l] for l in range(lcut + 1))))
    ncent = sum(T["neq"] for T in types)
    cent = g.uniform(0, 1, (ncent, 3))
    cmt = np.asfortranarray((g.standard_normal((maxlm, ncent, nband, nkpt)) +
                             1j * g.standard_normal((maxlm, ncent, nband, nkpt))) / np.sqrt(maxlm))
    return ProductSystem(nkpt, kpt, q, k1, k2, g2, nband, gpt, ngpt, cpw,
                         np.ascontiguousarray(gptm, np.int32), tuple(nfft), 97.3, types, cent, cmt)


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
