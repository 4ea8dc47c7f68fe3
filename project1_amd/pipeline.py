"""Host-side assembly of per-q pipelines over device-resident inputs.

A :class:`Crystal` holds the synthetic wave functions of one k mesh on the
device together with the muffin-tin and plane-wave product plans.  A
:class:`QJob` adds the q-point specifics (band window, energies,
frequencies, Coulomb matrix) and builds the C-ABI :class:`QPipe` that runs
products -> chi0 -> eps^-1 on its own stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import spexhip, synth

N_MT = 294  # 2 atoms x 49 lm x 3 radial functions


class Crystal:
    """Synthetic crystal on an nmesh^3 k mesh, resident in HBM."""

    def __init__(self, ctx: spexhip.Context, nmesh: int, nband: int, ngptm: int, seed: int = 7,
                 nfft=(36, 36, 36), gcut: float = 7.1):
        self.ctx = ctx
        self.sys = s = synth.product_system(
            nmesh=nmesh, nband=nband, gcut=gcut, ngptm=ngptm, nfft=nfft, ntype=2, lcut=6, nrad=3,
            lcutm=6, nradm=3, seed=seed)
        self.mt = ctx.mt_plan(s.types, s.cent, s.cmt.shape[0])
        self.pw = ctx.pw_plan(s.nfft, s.gpt, s.ngpt, s.gptm, s.vol)
        self.dcmt = ctx.to_dev(np.asfortranarray(s.cmt).ravel(order="F"))
        self.dcpw = ctx.to_dev(np.asfortranarray(s.cpw).ravel(order="F"))
        self.nb = s.nbasp + s.ngptm

    def free(self):
        for b in (self.dcmt, self.dcpw):
            b.free()
        self.mt.destroy()
        self.pw.destroy()


class QJob:
    """One q-point: energies, frequencies and v^1/2 on the device + a QPipe."""

    def __init__(self, ctx: spexhip.Context, crystal: Crystal, nocc: int, nunocc: int, nw: int,
                 precision: int, seed: int, dense_v: bool = True, graph: bool = False,
                 ozaki_group: int = 0):
        self.ctx, self.crystal, self.graph = ctx, crystal, graph
        s = crystal.sys
        nk = len(s.k1)
        nb = crystal.nb
        g = synth.rng(seed)
        eo, eu = synth.energies(g, nk, nocc, nunocc)
        om, _ = synth.gl_frequencies(nw)
        if dense_v:
            _, vh, _ = synth.coulomb(g, nb)
            self.vh = vh
            vbuf = np.asfortranarray(vh.astype(np.complex128))
        else:
            self.vh = np.sqrt(np.exp(g.uniform(np.log(1e-3), np.log(1e3), nb)))
            vbuf = self.vh
        self.e_occ, self.e_unocc, self.omega = eo, eu, om
        self.bufs = [ctx.to_dev(eo), ctx.to_dev(eu), ctx.to_dev(np.ones(nk, np.int32)),
                     ctx.to_dev(om), ctx.to_dev(vbuf)]
        d = spexhip.QDesc()
        d.nk, d.nocc, d.nunocc, d.nb, d.nw = nk, nocc, nunocc, nb, nw
        d.precision, d.vkind, d.alpha, d.ozaki_group = precision, 1 if dense_v else 0, 2.0, ozaki_group
        d.mt_plan, d.pw_plan = crystal.mt.h, crystal.pw.h
        d.cmt_dev, d.ncent, d.nband, d.nkpt = crystal.dcmt.ptr, s.cmt.shape[1], s.nband, s.nkpt
        d.q[:] = [float(v) for v in s.q]
        d.cpw_dev = crystal.dcpw.ptr
        k1 = np.ascontiguousarray(s.k1, np.int32)
        k2 = np.ascontiguousarray(s.k2, np.int32)
        g2 = np.ascontiguousarray(s.g2, np.int32)
        d.k1, d.k2, d.g2 = k1.ctypes.data, k2.ctypes.data, g2.ctypes.data
        d.b1lo, d.b2lo = 0, nocc
        d.e_occ_dev, d.e_unocc_dev, d.nkpts_dev = (self.bufs[0].ptr, self.bufs[1].ptr,
                                                   self.bufs[2].ptr)
        d.nkpt_w, d.nsym1 = nk, 1
        d.omega_dev, d.vhalf_dev = self.bufs[3].ptr, self.bufs[4].ptr
        self.desc = d
        self.footprint = spexhip.qpipe_footprint(d)
        self.pipe = spexhip.QPipe(ctx, d, [k1, k2, g2])
        self.n = nk * nocc * nunocc

    def launch(self):
        self.pipe.launch(self.graph)

    def wait(self) -> float:
        return self.pipe.wait()

    def free(self):
        self.pipe.destroy()
        for b in self.bufs:
            b.free()
