"""CPU oracle for the SPEX polarisation-function path -- TEST INFRASTRUCTURE.

ctypes wrapper around ``oracle/liboracle.so`` (built from
``oracle/spex_oracle.c`` by ``oracle/Makefile``).  The C functions restate
the reference arithmetic line by line (file:line cited in the C source).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
library never links or calls it.

What pins each function (golden fixtures, closed forms, independent
algorithms) is stated by the tests that check it, tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    # the reference's own cuttetraN into oracle/_ref (skipped without the
    # reference sources)
    subprocess.run([str(HERE / "build_ref.sh")], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _declare(_lib)
    return _lib


_vp, _I, _L, _D = C.c_void_p, C.c_int, C.c_long, C.c_double


class MtType(C.Structure):
    _fields_ = [
        ("lcut", _I), ("nindx", _vp), ("lcutm", _I), ("nindxm", _vp),
        ("neq", _I), ("integral", _vp), ("d1", _I), ("d2", _I), ("d3", _I),
    ]


def _declare(L):
    sig = {
        "or_weights_none": (None, [_I, _I, _I, _vp, _vp, _I, _I, _vp, _I, _vp, _vp]),
        "or_chi0_pairsum": (None, [_L, _I, _vp, _L, _I, _vp, _vp, _L, _L, _I]),
        "or_chi0_herk": (None, [_L, _I, _vp, _L, _I, _vp, _vp, _I]),
        "or_scale": (None, [_L, _vp, _D]),
        "or_dielectric_diag": (None, [_I, _vp, _vp]),
        "or_dielectric_dense": (None, [_I, _vp, _vp, _vp]),
        "or_inverse_hpd_packed": (_I, [_I, _vp]),
        "or_wigner3j": (_D, [_I] * 6),
        "or_gaunt": (_D, [_I] * 6),
        "or_wfprod_mt": (None, [_I, _vp, _vp, _vp, _I, _vp, _vp, _vp, _I, _I, _I,
                                _I, _I, _I, _I, _vp, _L]),
        "or_wfprod_pw": (None, [_I, _vp, _vp, _vp, _vp, _vp, _I, _vp, _I, _I, _vp,
                                _D, _I, _vp, _I, _I, _I, _I, _I, _vp, _L]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a


def _p(a):
    return a.ctypes.data


# ----------------------------------------------------------------- weights
def weights_none(e_occ, e_unocc, omega, nkpt=None, nsym1=1, nkpts=None):
    """wghtSr (n, nw) Fortran-ordered, susceptibility.f:954-973."""
    e_occ = np.ascontiguousarray(e_occ, np.float64)
    e_unocc = np.ascontiguousarray(e_unocc, np.float64)
    nk, nocc = e_occ.shape
    nunocc = e_unocc.shape[1]
    omega = np.ascontiguousarray(omega, np.float64)
    nw = omega.size
    nkpt = nk if nkpt is None else nkpt
    nkpts = np.ones(nk, np.int32) if nkpts is None else np.ascontiguousarray(nkpts, np.int32)
    n = nk * nocc * nunocc
    w = np.zeros((nw, n), np.float64)  # C (nw, n) == Fortran (n, nw)
    lib().or_weights_none(nk, nocc, nunocc, _p(e_occ), _p(e_unocc), nkpt, nsym1,
                          _p(nkpts), nw, _p(omega), _p(w))
    return w.T  # (n, nw) view, Fortran-contiguous


# -------------------------------------------------------------------- chi0
def _fort_c128(a):
    return np.asfortranarray(a, dtype=np.complex128)


def chi0_pairsum(cprod, wght, pair_lo=-1, pair_hi=-1, nthreads=0, out=None):
    """Literal reference pair loop; returns packed (npack, nw) Fortran array."""
    cp = _fort_c128(cprod)
    n, nb = cp.shape
    w = np.asfortranarray(wght, np.float64)
    nw = w.shape[1]
    npack = nb * (nb + 1) // 2
    if out is None:
        out = np.zeros((nw, npack), np.complex128)
    lib().or_chi0_pairsum(n, nb, _p(cp), n, nw, _p(w), _p(out), pair_lo, pair_hi, nthreads)
    return out.T


def chi0_herk(cprod, wght, nthreads=0):
    cp = _fort_c128(cprod)
    n, nb = cp.shape
    w = np.asfortranarray(wght, np.float64)
    nw = w.shape[1]
    out = np.zeros((nw, nb * (nb + 1) // 2), np.complex128)
    lib().or_chi0_herk(n, nb, _p(cp), n, nw, _p(w), _p(out), nthreads)
    return out.T


# ---------------------------------------------------------------- packing
def pack_upper(full):
    """Column-major packed upper triangle of an (nb, nb) matrix."""
    nb = full.shape[0]
    iu = upper_indices(nb)
    return full[iu]


def upper_indices(nb):
    j = np.concatenate([np.full(k + 1, k) for k in range(nb)])
    i = np.concatenate([np.arange(k + 1) for k in range(nb)])
    return i, j


def unpack_herm(p, nb):
    full = np.zeros((nb, nb), np.complex128)
    i, j = upper_indices(nb)
    full[i, j] = p
    full[j, i] = np.conj(p)
    return full


# ----------------------------------------------------------- dielectric
def dielectric_diag(chi_packed, sqcoul):
    m = np.ascontiguousarray(chi_packed, np.complex128).copy()
    nb = sqcoul.size
    lib().or_dielectric_diag(nb, _p(np.ascontiguousarray(sqcoul, np.float64)), _p(m))
    return m


def dielectric_dense(chi_packed, vhalf):
    nb = vhalf.shape[0]
    v = np.asfortranarray(vhalf, np.complex128)
    c = np.ascontiguousarray(chi_packed, np.complex128)
    out = np.zeros(nb * (nb + 1) // 2, np.complex128)
    lib().or_dielectric_dense(nb, _p(v), _p(c), _p(out))
    return out


def inverse_hpd_packed(packed):
    m = np.ascontiguousarray(packed, np.complex128).copy()
    nb = int(round((np.sqrt(8 * m.size + 1) - 1) / 2))
    info = lib().or_inverse_hpd_packed(nb, _p(m))
    if info:
        raise np.linalg.LinAlgError(f"not positive definite (minor {info})")
    return m


def screened_w(einv_packed, vhalf):
    """W = v^1/2 eps^-1 v^1/2, packed upper (correlation.f:868-877:
    p_diagmatdiag(sqcoul, dielec) in the Coulomb eigenbasis).  vhalf: real
    vector sqrt(coul) or a dense nb x nb matrix."""
    vhalf = np.asarray(vhalf)
    nb = vhalf.shape[0]
    e = unpack_herm(np.asarray(einv_packed, np.complex128), nb)
    vh = np.diag(vhalf.astype(np.complex128)) if vhalf.ndim == 1 else vhalf.astype(np.complex128)
    return pack_upper(vh @ e @ vh)


def rpa_integrand(eps_packed, weight):
    """weight * sum(log(e) + 1 - e) over the eigenvalues e of eps, the RPA
    correlation-energy integrand of one frequency (correlation.f:686-690:
    Mdiagonalize, then sum(log(eig) + 1 - eig) * count(kptp==ikpt)/nkpt)."""
    e = np.asarray(eps_packed, np.complex128)
    nb = int(round((np.sqrt(8 * e.size + 1) - 1) / 2))
    eig = np.linalg.eigvalsh(unpack_herm(e, nb))
    return weight * float(np.sum(np.log(eig) + 1.0 - eig))


# ------------------------------------------------------------------ gaunt
def wigner3j(l1, l2, l3, m1, m2, m3):
    return lib().or_wigner3j(l1, l2, l3, m1, m2, m3)


def gaunt(l1, l2, l3, m1, m2, m3):
    return lib().or_gaunt(l1, l2, l3, m1, m2, m3)


# --------------------------------------------------------------- products
def wfprod_mt(types, cent, q, k1, k2, cmt, b1lo, nocc, b2lo, nunocc):
    """MT products (n, nbasp) Fortran array; wavefproducts.f:95-431.

    ``types``: list of dicts with lcut, nindx, lcutm, nindxm, neq, integral
    (3-d real array, Fortran order).  ``cmt``: complex (maxlm, ncent, nband,
    nkpt) Fortran-ordered.
    """
    keep = []
    arr = (MtType * len(types))()
    nbasp = 0
    for t, T in zip(arr, types):
        nindx = np.ascontiguousarray(T["nindx"], np.int32)
        nindxm = np.ascontiguousarray(T["nindxm"], np.int32)
        integ = np.asfortranarray(T["integral"], np.float64)
        keep += [nindx, nindxm, integ]
        t.lcut, t.lcutm, t.neq = T["lcut"], T["lcutm"], T["neq"]
        t.nindx, t.nindxm, t.integral = _p(nindx), _p(nindxm), _p(integ)
        t.d1, t.d2, t.d3 = integ.shape
        nbasp += T["neq"] * sum((2 * l + 1) * nindxm[l] for l in range(T["lcutm"] + 1))
    cmt = np.asfortranarray(cmt, np.complex128)
    maxlm, ncent, nband, _ = cmt.shape
    cent = np.ascontiguousarray(cent, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    k1 = np.ascontiguousarray(k1, np.int32)
    k2 = np.ascontiguousarray(k2, np.int32)
    nk = k1.size
    n = nk * nocc * nunocc
    out = np.zeros((nbasp, n), np.complex128)
    lib().or_wfprod_mt(len(types), C.cast(arr, _vp), _p(cent), _p(q), nk, _p(k1), _p(k2),
                       _p(cmt), maxlm, ncent, nband, b1lo, nocc, b2lo, nunocc, _p(out), n)
    return out.T


def wfprod_pw(k1, k2, g2, gpt, ngpt, cpw, gptm, vol, b1lo, nocc, b2lo, nunocc,
              cstep=None, sb=0):
    """IPW products (n, ngptm) Fortran array; wavefproducts.f:766-997."""
    gpt = np.ascontiguousarray(gpt, np.int32)  # (nkpt, maxgpt, 3)
    maxgpt = gpt.shape[1]
    ngpt = np.ascontiguousarray(ngpt, np.int32)
    cpw = np.asfortranarray(cpw, np.complex128)  # (maxgpt, nband, nkpt)
    nband = cpw.shape[1]
    gptm = np.ascontiguousarray(gptm, np.int32)
    ngptm = gptm.shape[0]
    k1 = np.ascontiguousarray(k1, np.int32)
    k2 = np.ascontiguousarray(k2, np.int32)
    g2 = np.ascontiguousarray(g2, np.int32)
    nk = k1.size
    n = nk * nocc * nunocc
    fullpw = cstep is not None
    cs = np.ascontiguousarray(cstep if fullpw else np.zeros(1), np.complex128)
    out = np.zeros((ngptm, n), np.complex128)
    lib().or_wfprod_pw(nk, _p(k1), _p(k2), _p(g2), _p(gpt), _p(ngpt), maxgpt, _p(cpw),
                       nband, ngptm, _p(gptm), float(vol), int(fullpw), _p(cs), int(sb),
                       b1lo, nocc, b2lo, nunocc, _p(out), n)
    return out.T
