"""Serving workload: a Poisson stream of mixed-size chi0 / eps^-1 requests
through the continuous-batching server (``spexhip_server_*``).

Each request is one q-point pipeline (products -> HILBERT NONE weights ->
chi0 -> eps = 1 - v^1/2 chi0 v^1/2 -> eps^-1) of its own size:

* k mesh: Nk in {27, 64, 125, 216} (3^3 .. 6^3, simple cubic, q = k_1);
* mixed basis Nb ~ U[400, 2000]: ``natoms_for(Nb)`` muffin-tin atoms of 147
  functions each (lmax 6, 3 radial functions) plus Nb - 147*natoms
  interstitial plane waves (the shortest G, all inside the 16^3 product box
  of a 36^3 grid);
* N_omega in {8, 12, 16} Gauss-Legendre imaginary frequencies;
* band window Nocc = 16, Nunocc = 400; energies as in the single-q case;
* dense SPD v^1/2: the leading Nb x Nb block of one 2000 x 2000 SPD square
  root (a principal block of an SPD matrix is SPD, and it stays in the
  [1e-3, 1e3]^(1/2) eigenvalue range by interlacing).

Wave functions of every mesh (plane-wave coefficients damped by
exp(-|G|^2/8), muffin-tin coefficients for all atoms) are generated on the
device once and stay resident; the requests share them read-only.

Reference context: SPEX computes one q at a time per MPI rank and hands
q-points to idle ranks (correlation.f:316-402); the per-request memory test
is susceptibility.f:674-713.
"""
from __future__ import annotations

import argparse
import json
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import spexhip, synth

MESHES = (3, 4, 5, 6)
NW_CHOICES = (8, 12, 16)
NB_RANGE = (400, 2000)


def chi0_flops(nw: int, n: int, nb: int) -> float:
    """Complex rank-k update of the upper triangle: 8 flop per complex MAC."""
    return 8.0 * nw * n * nb * (nb + 1) / 2


def eps_flops(nw: int, nb: int) -> float:
    """Two dense complex GEMMs (v^1/2 chi0 v^1/2) + the HPD inverse."""
    return nw * (16.0 * nb**3 + 16.0 / 3 * nb**3)


@dataclass(frozen=True)
class Spec:
    nmesh: int
    nb: int
    nw: int
    seed: int
    nunocc: int = 0  # 0: the lattice's band window
    iq: int = 1      # q = k-point iq of the mesh


def request_mix(count: int, seed: int, meshes=MESHES, nb_range=NB_RANGE,
                nw_choices=NW_CHOICES) -> list[Spec]:
    g = synth.rng(seed)
    return [Spec(int(g.choice(meshes)), int(g.integers(nb_range[0], nb_range[1] + 1)),
                 int(g.choice(nw_choices)), int(seed * 1000 + i)) for i in range(count)]


class Lattice:
    """Atom types, centres, muffin-tin plans and the master v^1/2 shared by
    all meshes; meshes are created on first use and stay resident."""

    def __init__(self, ctx: spexhip.Context, *, nocc=16, nunocc=400, lcut=6, nrad=3,
                 gcut=7.1, nfft=(36, 36, 36), max_atoms=7, nb_max=NB_RANGE[1],
                 ipw_target=1000, seed=99):
        self.ctx, self.nocc, self.nunocc = ctx, nocc, nunocc
        self.nband = nocc + nunocc
        self.gcut, self.nfft, self.max_atoms = gcut, tuple(nfft), max_atoms
        self.ipw_target = ipw_target
        self.seed = seed
        g = synth.rng(seed)
        nindx = np.full(lcut + 1, nrad, np.int32)
        self.per_atom = int(sum((2 * l + 1) * nrad for l in range(lcut + 1)))
        self.maxlm = self.per_atom
        dims = synth.integral_dims(lcut, nindx, lcut, nindx)
        self.types = [dict(lcut=lcut, nindx=nindx, lcutm=lcut, nindxm=nindx, neq=1,
                           integral=np.asfortranarray(g.standard_normal(dims) / 10))
                      for _ in range(max_atoms)]
        self.cent = g.uniform(0, 1, (max_atoms, 3))
        self.mt_plans: dict[int, spexhip.Plan] = {}
        _, vh, _ = synth.coulomb(g, nb_max)
        self.vhalf = vh
        r = int(np.ceil(gcut)) + 1
        ax = np.arange(-r, r + 1)
        G = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
        n2 = np.einsum("ij,ij->i", G, G)
        self.gsorted = np.ascontiguousarray(G[np.lexsort((G[:, 2], G[:, 1], G[:, 0], n2))],
                                            np.int32)
        self.meshes: dict[int, Mesh] = {}
        self.lock = threading.Lock()

    def natoms_for(self, nb: int) -> int:
        """Atoms so that the plane-wave part stays near ipw_target columns."""
        a = -(-(nb - self.ipw_target) // self.per_atom)
        return int(min(self.max_atoms, max(2, a)))

    def mt_plan(self, natoms: int) -> spexhip.Plan:
        with self.lock:
            if natoms not in self.mt_plans:
                self.mt_plans[natoms] = self.ctx.mt_plan(self.types[:natoms],
                                                         self.cent[:natoms], self.maxlm)
            return self.mt_plans[natoms]

    def mesh(self, nmesh: int) -> "Mesh":
        with self.lock:
            if nmesh not in self.meshes:
                self.meshes[nmesh] = Mesh(self, nmesh)
            return self.meshes[nmesh]

    def free(self):
        for m in self.meshes.values():
            m.free()
        for p in self.mt_plans.values():
            p.destroy()
        self.meshes, self.mt_plans = {}, {}


class Mesh:
    """nmesh^3 k mesh with device-resident wave functions for max_atoms atoms."""

    def __init__(self, lat: Lattice, nmesh: int):
        ctx = lat.ctx
        self.lat, self.nmesh = lat, nmesh
        self.kpt = kpt = synth.kmesh(nmesh)
        self.nkpt = nk = kpt.shape[0]
        self.k1 = np.arange(nk, dtype=np.int32)
        self._q: dict[int, tuple] = {}
        r = int(np.ceil(lat.gcut)) + 1
        ax = np.arange(-r, r + 1)
        G = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
        lists = []
        for k in range(nk):
            kk = kpt[k] + G
            lists.append(G[np.einsum("ij,ij->i", kk, kk) <= lat.gcut**2])
        self.maxgpt = maxgpt = max(len(s) for s in lists)
        self.gpt = np.zeros((nk, maxgpt, 3), np.int32)
        self.ngpt = np.zeros(nk, np.int32)
        damp = np.zeros((maxgpt, nk))
        for k, s in enumerate(lists):
            self.ngpt[k] = len(s)
            self.gpt[k, : len(s)] = s
            damp[: len(s), k] = np.exp(-np.einsum("ij,ij->i", s, s) / 8.0)
        nb_, ma = lat.nband, lat.max_atoms
        seed = lat.seed * 100 + nmesh
        held = []  # freed again if a later allocation fails (nothing half-built stays)
        try:
            self.dcmt = ctx.alloc(16 * lat.maxlm * ma * nb_ * nk)
            held.append(self.dcmt)
            ctx.synth_normal(self.dcmt, 2 * lat.maxlm * ma * nb_ * nk, seed,
                             1.0 / np.sqrt(lat.maxlm))
            ddamp = ctx.to_dev(np.asfortranarray(damp).ravel(order="F"))
            held.append(ddamp)
            self.dcpw = ctx.alloc(16 * maxgpt * nb_ * nk)
        except spexhip.SpexError:
            ctx.sync()
            for b in held:
                b.free()
            raise
        ctx.synth_normal_rows(self.dcpw, maxgpt, nb_, nk, seed + 1, ddamp)
        ctx.sync()
        ddamp.free()
        self.pw_plans: dict[int, spexhip.Plan] = {}
        self.lock = threading.Lock()

    def qpoint(self, iq: int):
        """(q, k2, g2) for q = kpt[iq]: k2 = index of k+q folded into the mesh,
        g2 = its umklapp vector (vector.f:107-124, wavefproducts.f:543-552)."""
        iq %= self.nkpt
        if iq not in self._q:
            q = self.kpt[iq].copy()
            k2 = np.array([synth.kptsum_index(self.kpt, self.nmesh, iq, k) for k in self.k1],
                          np.int32)
            g2 = np.ascontiguousarray(np.rint(self.kpt[k2] - q - self.kpt[self.k1]), np.int32)
            self._q[iq] = (q, k2, g2)
        return self._q[iq]

    def pw_plan(self, ngptm: int) -> spexhip.Plan:
        with self.lock:
            if ngptm not in self.pw_plans:
                self.pw_plans[ngptm] = self.lat.ctx.pw_plan(
                    self.lat.nfft, self.gpt, self.ngpt, self.lat.gsorted[:ngptm], 97.3)
            return self.pw_plans[ngptm]

    def free(self):
        for p in self.pw_plans.values():
            p.destroy()
        self.pw_plans = {}
        self.dcmt.free()
        self.dcpw.free()


class Request:
    """Device inputs and the C descriptor of one q-point request."""

    def __init__(self, lat: Lattice, spec: Spec, precision: int, ozaki_group: int = 0,
                 packet_cap: int = 24 << 30):
        ctx = lat.ctx
        self.spec, self.precision = spec, precision
        mesh = lat.mesh(spec.nmesh)
        nk, nb, nw = mesh.nkpt, spec.nb, spec.nw
        natoms = lat.natoms_for(nb)
        ngptm = nb - natoms * lat.per_atom
        assert ngptm > 0, (nb, natoms)
        mt, pw = lat.mt_plan(natoms), mesh.pw_plan(ngptm)
        nunocc = spec.nunocc or lat.nunocc
        assert lat.nocc + nunocc <= lat.nband, (nunocc, lat.nband)
        q, k2, g2 = mesh.qpoint(spec.iq)
        g = synth.rng(spec.seed)
        eo, eu = synth.energies(g, nk, lat.nocc, nunocc)
        om, _ = synth.gl_frequencies(nw)
        vh = np.asfortranarray(lat.vhalf[:nb, :nb].astype(np.complex128))
        self.bufs = []
        try:
            for a in (eo, eu, np.ones(nk, np.int32), om, vh):
                self.bufs.append(ctx.to_dev(a))
        except spexhip.SpexError:
            for b in self.bufs:  # nothing half-staged stays allocated
                b.free()
            raise
        d = spexhip.QDesc()
        d.nk, d.nocc, d.nunocc, d.nb, d.nw = nk, lat.nocc, nunocc, nb, nw
        d.precision, d.vkind, d.alpha, d.ozaki_group = precision, 1, 2.0, ozaki_group
        d.mt_plan, d.pw_plan = mt.h, pw.h
        d.cmt_dev, d.ncent, d.nband, d.nkpt = mesh.dcmt.ptr, lat.max_atoms, lat.nband, nk
        d.q[:] = [float(v) for v in q]
        d.cpw_dev = mesh.dcpw.ptr
        d.k1, d.k2, d.g2 = mesh.k1.ctypes.data, k2.ctypes.data, g2.ctypes.data
        d.b1lo, d.b2lo = 0, lat.nocc
        d.e_occ_dev, d.e_unocc_dev, d.nkpts_dev = (self.bufs[0].ptr, self.bufs[1].ptr,
                                                   self.bufs[2].ptr)
        d.nkpt_w, d.nsym1 = nk, 1
        d.omega_dev, d.vhalf_dev = self.bufs[3].ptr, self.bufs[4].ptr
        # k packets: the largest packet whose pipeline fits packet_cap bytes
        # (the reference sizes its packets against MEM the same way)
        kp = nk
        d.kpacket = kp
        while kp > 1 and spexhip.qpipe_footprint(d) > packet_cap:
            kp = (kp + 1) // 2
            d.kpacket = kp
        self.desc = d
        self._keep = (mesh, k2, g2)  # k lists referenced by the descriptor
        self.e_occ, self.e_unocc, self.omega = eo, eu, om
        self.n = nk * lat.nocc * nunocc
        self.flops = chi0_flops(nw, self.n, nb) + eps_flops(nw, nb)
        self.footprint = spexhip.qpipe_footprint(d)

    def free(self):
        for b in self.bufs:
            b.free()
        self.bufs = []


def _pct(v, p):
    return float(np.percentile(np.asarray(v), p)) if len(v) else float("nan")


class Driver:
    """Submits requests (open or closed loop) and collects their completions."""

    def __init__(self, lat: Lattice, precision: int, max_inflight: int, hbm_budget: int,
                 backfill: bool = True, keep_results: bool = False, log=None):
        self.lat, self.precision = lat, precision
        self.server = spexhip.Server(lat.ctx, max_inflight, hbm_budget, backfill)
        self.keep_results = keep_results
        self.results: dict[int, tuple] = {}
        self.records: list[dict] = []
        self.log = log
        self._pending: list[tuple] = []  # (index, request, server id, submit time)
        self._finishing = 0  # taken off _pending, not yet recorded
        self._cv = threading.Condition()
        self._closed = False
        self._thr = threading.Thread(target=self._collect, daemon=True)
        self._thr.start()

    def submit(self, idx: int, req: Request) -> None:
        rid = self.server.submit(req.desc)
        with self._cv:
            self._pending.append((idx, req, rid, time.perf_counter()))
            self._cv.notify()

    def _collect(self):
        """Finish requests in completion order (release frees their HBM)."""
        while True:
            with self._cv:
                while not self._pending and not self._closed:
                    self._cv.wait()
                if not self._pending:
                    return
                pending = list(self._pending)
            done = [p for p in pending if self.server.state(p[2]) >= spexhip.Server.DONE]
            if not done:
                time.sleep(0.0005)
                continue
            with self._cv:
                for p in done:
                    self._pending.remove(p)
                self._finishing += len(done)
            for idx, req, rid, t_sub in done:
                self._finish(idx, req, rid, t_sub)
                with self._cv:
                    self._finishing -= 1

    def _finish(self, idx, req, rid, t_sub):
        err = None
        try:
            lat_ms, q_ms, dev_ms = self.server.wait(rid)
            if self.keep_results:
                self.results[idx] = self.server.result(rid)
        except spexhip.SpexError as e:
            lat_ms = q_ms = dev_ms = float("nan")
            err = str(e)
        t_done = time.perf_counter()
        self.server.release(rid)
        req.free()
        rec = dict(idx=idx, nk=req.desc.nk, nb=req.spec.nb, nw=req.spec.nw,
                   latency_ms=lat_ms, queue_ms=q_ms, device_ms=dev_ms, flops=req.flops,
                   t_submit=t_sub, t_done=t_done, error=err)
        with self._cv:
            self.records.append(rec)
        if self.log:
            self.log(rec)

    def outstanding(self) -> int:
        """Requests submitted and not yet finished."""
        with self._cv:
            return len(self._pending) + self._finishing

    def drain(self) -> None:
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        self._thr.join()

    def close(self) -> dict:
        self.drain()
        st = self.server.stats()
        self.server.destroy()
        return st


def summarize(records: list[dict], t0: float, stats: dict) -> dict:
    ok = [r for r in records if r["error"] is None]
    if not ok:
        return {"completed": 0, "errors": len(records)}
    t_end = max(r["t_done"] for r in ok)
    span = t_end - t0
    lat = [r["latency_ms"] for r in ok]
    return {
        "completed": len(ok), "errors": len(records) - len(ok),
        "span_s": round(span, 3),
        "qpoints_per_s": round(len(ok) / span, 4),
        "gflops": round(sum(r["flops"] for r in ok) / span / 1e9, 1),
        "p50_ms": round(_pct(lat, 50), 1), "p99_ms": round(_pct(lat, 99), 1),
        "mean_queue_ms": round(float(np.mean([r["queue_ms"] for r in ok])), 1),
        "mean_device_ms": round(float(np.mean([r["device_ms"] for r in ok])), 1),
        "peak_inflight": stats["peak_inflight"], "peak_hbm_bytes": stats["peak_hbm_bytes"],
    }


def _stage(lat: Lattice, spec: Spec, precision: int, patience_s: float = 120.0):
    """A request's inputs on the device.  When the device is momentarily full
    (pipelines in flight plus the allocator's cache), wait for completions
    to release memory and try again rather than fail the stream."""
    t_end = time.perf_counter() + patience_s
    while True:
        try:
            return Request(lat, spec, precision)
        except spexhip.SpexError as e:
            if e.code != spexhip.E_NOMEM or time.perf_counter() > t_end:
                raise
            time.sleep(0.05)


def run_closed(lat: Lattice, specs: list[Spec], precision: int, max_inflight: int,
               hbm_budget: int, keep_results=False, log=None):
    """All requests submitted at once (backlogged server): capacity."""
    reqs = [Request(lat, s, precision) for s in specs]
    lat.ctx.sync()
    drv = Driver(lat, precision, max_inflight, hbm_budget, keep_results=keep_results, log=log)
    t0 = time.perf_counter()
    try:
        for i, r in enumerate(reqs):
            drv.submit(i, r)
    finally:
        stats = drv.close()  # drains the server even when a submission fails
    return summarize(drv.records, t0, stats), drv.results


def run_sustained(lat: Lattice, specs: list[Spec], precision: int, max_inflight: int,
                  hbm_budget: int, log=None):
    """Capacity as the stream sees it: every request is staged by the
    submitting thread (as in the open loop) and submitted as soon as fewer
    than max_inflight + 1 are outstanding, so staging, admission and the
    device work overlap the way they do under load.  Completed requests
    per second from the first submission to the last completion."""
    drv = Driver(lat, precision, max_inflight, hbm_budget, log=log)
    t0 = time.perf_counter()
    try:
        for i, sp in enumerate(specs):
            req = _stage(lat, sp, precision)
            while drv.outstanding() > max_inflight:
                time.sleep(0.001)
            drv.submit(i, req)
    finally:
        stats = drv.close()
    return summarize(drv.records, t0, stats)


def run_poisson(lat: Lattice, specs: list[Spec], precision: int, rate: float, seconds: float,
                max_inflight: int, hbm_budget: int, seed: int = 5, log=None,
                patience_s: float = 120.0):
    """Open loop: Poisson arrivals at ``rate`` requests/s for ``seconds``
    (the spec list is cycled); each request's inputs are staged before its
    arrival time, and latency runs from submission to completion.  If a
    request cannot be staged within ``patience_s`` (the device stays full),
    the requests already submitted are drained and the error is raised
    with their summary attached as ``partial``."""
    g = synth.rng(seed)
    arrivals = []
    t = 0.0
    while True:
        t += g.exponential(1.0 / rate)
        if t > seconds:
            break
        arrivals.append(t)
    drv = Driver(lat, precision, max_inflight, hbm_budget, log=log)
    t0 = time.perf_counter()
    err = None
    try:
        for i, ta in enumerate(arrivals):
            req = _stage(lat, specs[i % len(specs)], precision, patience_s)
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            drv.submit(i, req)
    except spexhip.SpexError as e:
        err = e
    finally:
        stats = drv.close()  # drains the server even when staging fails
    out = summarize(drv.records, t0, stats)
    out.update(offered_rate=round(rate, 4), window_s=seconds, submitted=len(arrivals))
    if err is not None:
        err.partial = out
        raise err
    return out


def serve_benchmark(ctx: spexhip.Context, precisions: dict, seconds, cap_requests: int,
                    max_inflight: int = 4, hbm_budget: int = 200 << 30, load: float = 0.8,
                    seed: int = 2027, log=print) -> dict:
    """Sustained capacity (``cap_requests`` requests staged by the submitting
    thread, at most max_inflight + 1 outstanding) then an open-loop Poisson
    run at ``load`` x capacity, for each named precision.  ``seconds`` is the
    window of every precision, or a dict of windows by name."""
    lat = Lattice(ctx)
    specs = request_mix(4096, seed)
    for m in sorted({s.nmesh for s in specs}):
        lat.mesh(m)
    for na in sorted({lat.natoms_for(s.nb) for s in specs}):
        lat.mt_plan(na)
    out = {"mix": {"Nk": [m**3 for m in MESHES], "Nb": list(NB_RANGE), "Nw": list(NW_CHOICES),
                   "Nocc": lat.nocc, "Nunocc": lat.nunocc, "max_inflight": max_inflight,
                   "hbm_budget_gib": hbm_budget / 2**30, "load": load}}
    try:
        for name, prec in precisions.items():
            cap = run_sustained(lat, specs[:cap_requests], prec, max_inflight, hbm_budget,
                                log=log)
            if log:
                log(f"[serve] {name} sustained capacity {cap['qpoints_per_s']} q/s over "
                    f"{cap['completed']} requests")
            rate = load * cap["qpoints_per_s"]
            win = seconds[name] if isinstance(seconds, dict) else seconds
            res = run_poisson(lat, specs[cap_requests:], prec, rate, win, max_inflight,
                              hbm_budget, seed=seed + 1, log=log)
            res["capacity_qpoints_per_s"] = cap["qpoints_per_s"]
            res["capacity_requests"] = cap["completed"]
            # the queue kept up when the window's throughput is near the offered rate
            res["completed_over_offered"] = round(res["qpoints_per_s"] / rate, 3)
            if log:
                log(f"[serve] {name} " + json.dumps(res))
            out[name] = res
    finally:
        lat.free()
    return out


def main():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--seconds", type=float, default=600.0)
    p.add_argument("--cap-requests", type=int, default=24)
    p.add_argument("--max-inflight", type=int, default=4)
    p.add_argument("--budget-gib", type=float, default=200.0)
    p.add_argument("--load", type=float, default=0.8)
    p.add_argument("--precisions", default="fp64,split_fp32")
    args = p.parse_args()
    names = {"fp64": spexhip.PREC_FP64, "split_fp32": spexhip.PREC_SPLIT_FP32,
             "ozaki_int8": spexhip.PREC_OZAKI_INT8}
    precs = {n: names[n] for n in args.precisions.split(",")}
    done = [0]

    def log(x):
        if isinstance(x, dict):
            done[0] += 1
            if done[0] % 10 == 0:
                print(f"[serve] {done[0]} completed, last latency {x['latency_ms']:.0f} ms",
                      flush=True)
        else:
            print(x, flush=True)

    with spexhip.Context(0) as ctx:
        res = serve_benchmark(ctx, precs, args.seconds, args.cap_requests, args.max_inflight,
                              int(args.budget_gib * 2**30), args.load, log=log)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
