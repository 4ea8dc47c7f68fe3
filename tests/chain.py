"""Oracle chain for a served / batched request: products -> chi0 (spin
factor 2) -> eps (dense v^1/2) -> eps^-1, from the device-resident wave
functions of its mesh."""
import numpy as np

import oracle


U64 = 2.0**-53


def einv_tol(cond, nb, chi_err=0.0):
    """Bound for the normwise difference of two eps^-1: backward-stable FP64
    inversions differ by up to ~c n u cond(eps), and a relative chi0
    difference chi_err propagates as ~cond(eps) chi_err."""
    return 1e-13 + 10 * cond * (nb * U64 + chi_err)


def eps_cond(chi, vh):
    """Largest 2-norm condition number of eps = 1 - v^1/2 chi0 v^1/2 over the
    frequencies (chi: nw x packed).  Two backward-stable FP64 inversions of
    eps differ by up to ~cond * 1e-16 relative, so eps^-1 tolerances scale
    with it."""
    nb = vh.shape[0]
    return max(np.linalg.cond(oracle.unpack_herm(oracle.dielectric_dense(c, vh), nb))
               for c in chi)


def oracle_for_request(lat, req):
    spec, d = req.spec, req.desc
    mesh = lat.mesh(spec.nmesh)
    natoms = lat.natoms_for(spec.nb)
    ngptm = spec.nb - natoms * lat.per_atom
    nk, nb_ = mesh.nkpt, lat.nband
    q, k2, g2 = mesh.qpoint(spec.iq)
    cmt = mesh.dcmt.download(np.complex128, lat.maxlm * lat.max_atoms * nb_ * nk)
    cmt = cmt.reshape((lat.maxlm, lat.max_atoms, nb_, nk), order="F")
    cpw = mesh.dcpw.download(np.complex128, mesh.maxgpt * nb_ * nk)
    cpw = cpw.reshape((mesh.maxgpt, nb_, nk), order="F")
    mt = oracle.wfprod_mt(lat.types[:natoms], lat.cent[:natoms], q, mesh.k1, k2,
                          np.asfortranarray(cmt[:, :natoms]), 0, d.nocc, d.nocc, d.nunocc)
    pw = oracle.wfprod_pw(mesh.k1, k2, g2, mesh.gpt, mesh.ngpt, cpw, lat.gsorted[:ngptm], 97.3,
                          0, d.nocc, d.nocc, d.nunocc)
    M = np.concatenate([mt, pw], axis=1)
    w = oracle.weights_none(req.e_occ, req.e_unocc, req.omega)
    chi = 2 * oracle.chi0_herk(M, w).T
    vh = lat.vhalf[: spec.nb, : spec.nb].astype(complex)
    einv = np.stack([oracle.inverse_hpd_packed(oracle.dielectric_dense(chi[f], vh))
                     for f in range(spec.nw)])
    return chi, einv, eps_cond(chi, vh)
