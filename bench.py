#!/usr/bin/env python3
"""Benchmark: chi0 + eps^-1 q-points/s on one MI355X, split forms vs FP64.

Workload (one "step" = one q-point of the single-q realistic case):
  Nk=64 k-points (4x4x4), Nocc=16, Nunocc=400 (n = 409,600 transitions),
  mixed basis Nb=1200 = 294 muffin-tin (2 atoms, lmax 6, 3 radial
  functions) + 906 interstitial plane waves, Nw=12 imaginary frequencies
  (Gauss-Legendre on [0,1] mapped to omega = x/(1-x)), 36^3 FFT grid,
  dense random SPD Coulomb matrix (eigenvalues log-uniform in [1e-3, 1e3]).
  Timed per step, inputs (wave functions, energies, v^1/2) resident in HBM:
    products M (MT Gaunt contraction + IPW FFT convolution, FP64)
    -> HILBERT NONE weights -> chi0 (rank-k, spin factor 2)
    -> eps = 1 - v^1/2 chi0 v^1/2 (FP64 MFMA GEMMs) -> eps^-1 (FP64).

Three chi0 forms run on the same inputs: plain FP64, split-FP32 (hi+lo,
3 FP32 MFMA products) and the Ozaki int8 form.  Prints ONE JSON line;
``value`` is q-points/s of the fastest split form whose chi0 and eps^-1
stay within 1e-8 (max-abs and Frobenius relative) of the FP64 form.  The
roofline of its chi0 kernel and the CPU baseline (the oracle's literal
pair loop on a bounded sample) ride along.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("χ0+ε⁻¹ q-points/s & GFLOP/s, split-FP32 vs FP64 at "
          "≤1e-8 rel; p50/p99 latency")
PEAK_F64 = 78.6  # TFLOP/s dense FP64 matrix, MI355X
WORKLOAD = ("single q, Nk=64 Nocc=16 Nunocc=400 Nb=1200 (294 MT + 906 IPW) Nw=12, 36^3 FFT "
            "grid, dense v^1/2; products->chi0->eps^-1")
PEAK_F32 = 157.3  # TFLOP/s dense FP32 matrix, MI355X


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nk", type=int, default=64)
    p.add_argument("--nocc", type=int, default=16)
    p.add_argument("--nunocc", type=int, default=400)
    p.add_argument("--nb", type=int, default=1200)
    p.add_argument("--nw", type=int, default=12)
    p.add_argument("--seed", type=int, default=2026)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-products", action="store_true",
                   help="use Gaussian products generated once instead of computing M each step")
    p.add_argument("--cpu-cols", type=int, default=128,
                   help="CPU baseline sample: all pairs among this many columns")
    p.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                   help="only run 5 plain device copies of MIB MiB (byte-counter calibration)")
    p.add_argument("--only", choices=[f for f, _ in FORMS] + ["theta"], default=None,
                   help="profiling aid: run one chi0 form (or the theta products leg) and "
                        "print its timings only")
    p.add_argument("--no-batch", action="store_true", help="skip the 32-q batched leg")
    p.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    p.add_argument("--serve-seconds", type=float, default=60.0,
                   help="open-loop window of the serving leg per precision")
    p.add_argument("--serve-cap", type=int, default=24,
                   help="requests of the sustained-capacity run per precision")
    return p.parse_args()


def chi0_flops(nw, n, nb):
    return 8.0 * nw * n * nb * (nb + 1) / 2


def eps_flops(nw, nb):
    return nw * (16.0 * nb**3 + 16.0 / 3 * nb**3)


class Pipeline:
    """Device-resident single-q pipeline: weights -> chi0 -> eps -> eps^-1."""

    def __init__(self, ctx, a):
        from project1_amd import synth

        self.ctx, self.a = ctx, a
        nk, no, nu, nb, nw = a.nk, a.nocc, a.nunocc, a.nb, a.nw
        self.n = n = nk * no * nu
        g = synth.rng(a.seed)
        eo, eu = synth.energies(g, nk, no, nu)
        self.omega, _ = synth.gl_frequencies(nw)
        _, vh, lam = synth.coulomb(g, nb)
        self.vh = vh
        self.dM = ctx.alloc(16 * n * nb)
        self.products = not a.no_products
        if self.products:
            # synthetic wave functions (cubic nk mesh, |k+G| <= 7.1 -> Ng ~ 1500, 906 IPWs on
            # a 36^3 grid; 2 atoms, lmax 6, 3 radial functions) -> M on the device each step
            nmesh = round(nk ** (1 / 3))
            assert nmesh ** 3 == nk, "--nk must be a cube for the product workload"
            self.sys = sy = synth.product_system(
                nmesh=nmesh, nband=no + nu, gcut=7.1, ngptm=nb - 294, nfft=(36, 36, 36), ntype=2,
                lcut=6, nrad=3, lcutm=6, nradm=3, seed=a.seed + 2)
            assert sy.nbasp + sy.ngptm == nb
            self.mtp = ctx.mt_plan(sy.types, sy.cent, sy.cmt.shape[0])
            self.pwp = ctx.pw_plan(sy.nfft, sy.gpt, sy.ngpt, sy.gptm, sy.vol)
            self.dcmt = ctx.to_dev(np.asfortranarray(sy.cmt).ravel(order="F"))
            self.dcpw = ctx.to_dev(np.asfortranarray(sy.cpw).ravel(order="F"))
        else:
            ctx.synth_normal(self.dM, 2 * n * nb, a.seed + 1, 1.0 / np.sqrt(nb))
        self.deo, self.deu = ctx.to_dev(eo), ctx.to_dev(eu)
        self.dnk = ctx.to_dev(np.ones(nk, np.int32))
        self.dom = ctx.to_dev(self.omega)
        self.dW = ctx.alloc(8 * n * nw)
        self.dV = ctx.to_dev(np.asfortranarray(vh.astype(np.complex128)))
        self.dX = ctx.alloc(16 * nb * nb * nw)
        self.dE = ctx.alloc(16 * nb * nb * nw)
        self.ev = [ctx.event() for _ in range(5)]
        ctx.sync()

    def make_products(self):
        ctx, a, sy = self.ctx, self.a, self.sys
        ctx.wfprod_mt_dev(self.mtp, self.dcmt, sy.cmt.shape[1], sy.nband, sy.nkpt, sy.q, sy.k1,
                          sy.k2, 0, a.nocc, a.nocc, a.nunocc, self.dM, a.nb, 0)
        ctx.wfprod_pw_dev(self.pwp, self.dcpw, sy.nband, sy.k1, sy.k2, sy.g2, 0, a.nocc, a.nocc,
                          a.nunocc, self.dM, a.nb, sy.nbasp)

    def step(self, prec):
        ctx, a, ev = self.ctx, self.a, self.ev
        ev[4].record()
        if self.products:
            self.make_products()
        ev[0].record()
        ctx.weights_none_dev(self.deo, self.deu, a.nk, a.nocc, a.nunocc, a.nk, 1, self.dnk,
                             self.dom, a.nw, self.dW)
        ev[1].record()
        ctx.chi0_dev(self.n, a.nb, self.dM, a.nb, False, a.nw, self.dW, prec, self.dX, a.nb,
                     alpha=2.0)
        ev[2].record()
        ctx.dielectric_dev(self.dX, a.nb, a.nw, a.nb, 1, self.dV, self.dE, a.nb)
        ctx.inverse_hpd_dev(self.dE, a.nb, a.nw, a.nb)
        ev[3].record()
        return (ev[2].ms_since(ev[1]), ev[3].ms_since(ev[2]), ev[3].ms_since(ev[4]),
                ev[0].ms_since(ev[4]))

    def free(self):
        bufs = [self.dM, self.deo, self.deu, self.dnk, self.dom, self.dW, self.dV, self.dX,
                self.dE]
        if self.products:
            bufs += [self.dcmt, self.dcpw]
            self.mtp.destroy()
            self.pwp.destroy()
        for b in bufs:
            b.free()

    def results(self):
        a = self.a
        chi = self.ctx.pack_upper(self.dX, a.nb, a.nw, a.nb)
        einv = self.ctx.pack_upper(self.dE, a.nb, a.nw, a.nb)
        return chi, einv


def run_form(pipe, prec, steps, warmup):
    for _ in range(warmup):
        pipe.step(prec)
    pipe.ctx.sync()
    t0 = time.perf_counter()
    chi_ms, eps_ms, tot_ms, prod_ms = [], [], [], []
    for _ in range(steps):
        c, e, t, p = pipe.step(prec)
        chi_ms.append(c)
        eps_ms.append(e)
        tot_ms.append(t)
        prod_ms.append(p)
    pipe.ctx.sync()
    wall = time.perf_counter() - t0
    return dict(wall_s=wall, ms_per_step=1e3 * wall / steps, chi0_ms=float(np.mean(chi_ms)),
                eps_ms=float(np.mean(eps_ms)), products_ms=float(np.mean(prod_ms)),
                step_ms_p50=float(np.percentile(tot_ms, 50)),
                step_ms_p99=float(np.percentile(tot_ms, 99)))


def errs(a, b):
    d = np.abs(a - b)
    return dict(max_abs=float(d.max()), max_rel=float(d.max() / np.abs(b).max()),
                fro_rel=float(np.linalg.norm(a - b) / np.linalg.norm(b)))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_herk(pipe, a, nthreads, tsample=16384):
    """Second CPU baseline: the oracle's blocked rank-k form (or_chi0_herk:
    per column j the weighted conj(M(:,j)) in blocks of 512 transitions,
    dotted against every column i <= j) over all Nb columns, on the first
    `tsample` transitions and one frequency, scaled linearly in n and Nw."""
    import oracle

    n, nb, nw = pipe.n, a.nb, a.nw
    ts = min(tsample, n)
    M = pipe.dM.download(np.complex128, (n, nb))
    Ms = np.asfortranarray(M[:ts])
    del M
    W = pipe.dW.download(np.float64, (nw, n)).T[:ts, :1].copy()
    t0 = time.perf_counter()
    oracle.chi0_herk(Ms, W, nthreads=nthreads)
    t = time.perf_counter() - t0
    scale = (n / ts) * nw
    return dict(value=1.0 / (t * scale), unit="q-points/s (chi0 only)", cores=nthreads,
                kind="port", sample=(f"or_chi0_herk over all {nb} columns, {ts} of {n} "
                                     f"transitions, 1 of {nw} frequencies ({t:.2f} s, "
                                     f"extrapolated x{scale:.0f}, linear in n and Nw)"),
                chi0_gflops=round(chi0_flops(1, ts, nb) / t / 1e9, 2))


def cpu_baseline(pipe, a):
    """The oracle's literal pair loop (susceptibility.f:1183-1240) on a sample."""
    import oracle

    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n, nb, nw = pipe.n, a.nb, a.nw
    cols = np.linspace(0, nb - 1, min(a.cpu_cols, nb)).astype(int)
    M = pipe.dM.download(np.complex128, (n, nb))
    Ms = np.asfortranarray(M[:, cols])
    del M
    W = pipe.dW.download(np.float64, (nw, n)).T
    nc = cols.size
    npairs = nc * (nc + 1) // 2
    t0 = time.perf_counter()
    oracle.chi0_pairsum(Ms, W, nthreads=nthreads)
    t_pairs = time.perf_counter() - t0
    total_pairs = nb * (nb + 1) / 2
    # one frequency of the dielectric step + inverse with the oracle
    chi_full = pipe.ctx.pack_upper(pipe.dX, nb, 1, nb)[0]
    t0 = time.perf_counter()
    eps = oracle.dielectric_dense(chi_full, pipe.vh.astype(np.complex128))
    oracle.inverse_hpd_packed(eps)
    t_eps1 = time.perf_counter() - t0
    t_q = t_pairs * total_pairs / npairs + t_eps1 * nw
    return dict(value=1.0 / t_q, unit="q-points/s", cores=nthreads, kind="port",
                cpu_model=cpu_model(), nproc=os.cpu_count(),
                sample=(f"literal pair loop over all {npairs} pairs of {nc} of {nb} columns "
                        f"({t_pairs:.2f} s, extrapolated x{total_pairs / npairs:.1f}) + eps/eps^-1 "
                        f"for 1 of {nw} frequencies ({t_eps1:.2f} s, x{nw}); chi0 scaled "
                        f"linearly in pairs (equal cost per pair)"),
                seconds_per_q_extrapolated=t_q,
                blocked_herk=cpu_herk(pipe, a, nthreads))


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` at `workload`, as measured offline with
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a copy-kernel
    calibration (profiles/traffic.json, written from those summaries); None
    when no measurement of this kernel at this workload is recorded."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    rec = json.loads(path.read_text()).get(kernel)
    if not rec or rec.get("workload") != workload:
        return None
    return rec["hbm_bytes_per_launch"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def hilbert_leg(pipe, a, nfrqs=64, reps=3):
    """Tetrahedron weights on the step's energies (4x4x4 mesh, insulator)
    over an exponential Hilbert mesh of `nfrqs` points, then their fold into
    per-transition weights of the Nw frequencies: HIP events on the context
    stream.  Not part of the timed step (HILBERT NONE there)."""
    ctx, n, nw = pipe.ctx, pipe.n, a.nw
    m = round(a.nk ** (1 / 3))
    frq = 0.05 + 0.01 * (1.1 ** np.arange(nfrqs) - 1)
    dWt = ctx.alloc(8 * n * nfrqs)
    dS = ctx.to_dev(np.ascontiguousarray(np.outer(1.0 / (1.0 + frq), 1.0 / (1.0 + pipe.omega)).T))
    dD = ctx.alloc(8 * n * nw)
    tw = lambda: ctx.tetra_weights_dev((m, m, m), np.eye(3), pipe.deo, pipe.deu, a.nocc,
                                       a.nunocc, 0.0, frq, dWt, n)
    fold = lambda: ctx.weights_hilbert_dev(n, nfrqs, dWt, n, 1e-10, nw, dS, nfrqs, dD, n)
    tw()
    fold()
    e0, e1, e2 = ctx.event(), ctx.event(), ctx.event()
    e0.record()
    for _ in range(reps):
        ctx.tetra_weights_dev((m, m, m), np.eye(3), pipe.deo, pipe.deu, a.nocc, a.nunocc, 0.0,
                              frq, dWt, n)
    e1.record()
    for _ in range(reps):
        fold()
    e2.record()
    ctx.sync()
    t_ms, f_ms = e1.ms_since(e0) / reps, e2.ms_since(e1) / reps
    for b in (dWt, dS, dD):
        b.free()
    for e in (e0, e1, e2):
        e.destroy()
    wbytes = 8 * n * nfrqs
    return {"kernels": "tetra_weights_kernel, weights_fold_kernel", "nfrqs": nfrqs,
            "tetra_ms": round(t_ms, 3), "fold_ms": round(f_ms, 3),
            "tetra_GBps_written": round(wbytes / (t_ms * 1e6), 1),
            "fold_GBps_read": round(wbytes / (f_ms * 1e6), 1),
            "note": "each tetra_weights call returns after its weights are written (it "
                    "reads a host frequency table); the events bracket the calls"}


def theta_leg(pipe, a, sb=3, reps=3):
    """The interstitial products of the step's workload with the step
    function theta(r) on the caller's 36^3 grid (SPEX's default FULLPW,
    wavefproducts.f:630-632): with theta the product is no longer a pure
    convolution, so the plan keeps the caller's grid instead of the 24^3
    alias-free one the timed step uses.  theta is synthetic, band-limited
    to |G_i| <= sb.  HIP events on the context stream around `reps` calls;
    not part of the timed step."""
    ctx, sy = pipe.ctx, pipe.sys
    g = np.random.default_rng(a.seed + 9)
    w = 2 * sb + 1
    cs = (g.standard_normal((w, w, w)) + 1j * g.standard_normal((w, w, w))) * 0.02
    cs = 0.5 * (cs + np.conj(cs[::-1, ::-1, ::-1]))
    cs[sb, sb, sb] = 0.75
    N = sy.nfft
    box = np.zeros(N[::-1], complex)
    for z in range(-sb, sb + 1):
        for y in range(-sb, sb + 1):
            for x in range(-sb, sb + 1):
                box[z % N[2], y % N[1], x % N[0]] = cs[z + sb, y + sb, x + sb]
    th = np.ascontiguousarray((np.fft.ifftn(box) * np.prod(N)).real).reshape(-1)
    plan = ctx.pw_plan(N, sy.gpt, sy.ngpt, sy.gptm, sy.vol, step_r=th)
    run = lambda: ctx.wfprod_pw_dev(plan, pipe.dcpw, sy.nband, sy.k1, sy.k2, sy.g2, 0, a.nocc,
                                    a.nocc, a.nunocc, pipe.dM, a.nb, sy.nbasp)
    run()
    e0, e1 = ctx.event(), ctx.event()
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    ctx.sync()
    ms = e1.ms_since(e0) / reps
    plan.destroy()
    for e in (e0, e1):
        e.destroy()
    pb = 16 * pipe.n * sy.ngptm + sy.cpw.nbytes  # IPW columns of M written, cpw read once
    gbs = pb / (ms * 1e6)
    return {"kernel": "pw_inverse_kernel + pw_pair_joint_kernel<36,36> (caller grid)",
            "grid": list(N), "ms": round(ms, 3), "achieved": round(gbs, 1), "peak": 8000.0,
            "unit": "GB/s", "frac": round(gbs / 8000.0, 4), "bytes_per_call": int(pb),
            "note": "IPW products with theta(r) on the 36^3 caller grid (SPEX's default FULLPW), "
                    "all %d transitions x %d IPWs; outside the timed step (which runs the "
                    "product without theta on the 24^3 alias-free grid)" % (pipe.n, sy.ngptm)}


def gamma_leg(pipe, a, reps=5):
    """The q = 0 head and wings (headwing_kernel) on this step's M and
    weights, with a synthetic momentum matrix, and the head/wing update of
    eps^-1: HIP events on the context stream around `reps` launches each.
    Not part of the timed step (the workload is q != 0)."""
    ctx, n, nb, nw = pipe.ctx, pipe.n, a.nb, a.nw
    dP = ctx.alloc(16 * 3 * n)
    ctx.synth_normal(dP, 6 * n, a.seed + 7, 1.0)
    dH, dWg = ctx.alloc(16 * 9 * nw), ctx.alloc(16 * 3 * nb * nw)
    hw = lambda: ctx.headwings_dev(n, nb, pipe.dM, nb, False, dP, n, nw, pipe.dW, n, dH, dWg,
                                   alpha=2.0)
    hw()
    e0, e1, e2 = ctx.event(), ctx.event(), ctx.event()
    e0.record()
    for _ in range(reps):
        hw()
    e1.record()
    for _ in range(reps):
        ctx.gamma_inverse_dev(nb, nw, pipe.dE, nb, dH, dWg)
    e2.record()
    ctx.sync()
    ms = e1.ms_since(e0) / reps
    inv_ms = e2.ms_since(e1) / reps
    for b in (dP, dH, dWg):
        b.free()
    for e in (e0, e1, e2):
        e.destroy()
    # algorithmic bytes per launch: M once per pass (16 B per element), the
    # momentum matrix (48 B per transition) and the weights (8 B each)
    passes = -(-nw // 16)
    byts = passes * (16 * n * nb + 48 * n) + 8 * n * nw
    gbs = byts / (ms * 1e6)
    return {"kernel": "headwing_kernel", "ms": round(ms, 3), "achieved": round(gbs, 1),
            "peak": 8000.0, "unit": "GB/s", "frac": round(gbs / 8000.0, 4),
            "bytes_per_launch": int(byts), "inverse_update_ms": round(inv_ms, 3),
            "note": "q = 0 head + wings of chi0 on this step's M (synthetic momentum matrix), "
                    "and the head/wing update of eps^-1 (gamma_wing + gamma_head kernels); "
                    "outside the timed step"}


def batch_leg(ctx, a, precision="split_fp32", budget=160 << 30):
    """32 uneven q-points (Nk=64, Nocc=16, Nunocc ~ U[200,600], Nb ~ U[600,1600],
    Nw=12).  split_fp32 / fp64: per k packet one grouped chi0 launch over all
    q with the autotuned tile configuration; ozaki_int8: per-q int8 rank-k
    launches on the q streams.  The fork/join pipeline is captured into one
    HIP graph.  Two q are checked against their FP64 per-q pipelines."""
    from project1_amd import serving, spexhip, synth

    prec = {"fp64": spexhip.PREC_FP64, "split_fp32": spexhip.PREC_SPLIT_FP32,
            "ozaki_int8": spexhip.PREC_OZAKI_INT8}[precision]
    lat = serving.Lattice(ctx, nocc=16, nunocc=600, seed=a.seed + 5)
    g = synth.rng(a.seed + 6)
    specs = [serving.Spec(nmesh=4, nb=int(g.integers(600, 1601)), nw=12, seed=a.seed * 10 + i,
                          nunocc=int(g.integers(200, 601)), iq=1 + i) for i in range(32)]
    reqs = [serving.Request(lat, sp, prec) for sp in specs]
    descs = [r.desc for r in reqs]
    kp = next(k for k in (64, 32, 16, 8, 4, 2, 1) if spexhip.Batch.footprint(descs, k) <= budget)
    t0 = time.perf_counter()
    b = spexhip.Batch(ctx, descs, kpacket=kp, config=-1, tune_k=16384)
    t_create = time.perf_counter() - t0
    cfg, npacket, tune = b.config()
    log(f"batch {precision}: {len(reqs)} q, kpacket {kp}, config {cfg}, created in {t_create:.1f} s")
    b.launch(graph=False)
    ms_eager = b.wait()
    log(f"batch {precision}: eager run {ms_eager:.0f} ms")
    b.launch(graph=True)  # captured once, then replayed
    ms_graph = b.wait()
    log(f"batch {precision}: graph run {ms_graph:.0f} ms")
    flops = sum(chi0_flops(r.spec.nw, r.n, r.spec.nb) + eps_flops(r.spec.nw, r.spec.nb)
                for r in reqs)
    check = {}
    for i in sorted({0, int(np.argmax([sp.nb for sp in specs]))}):
        ref = serving.Request(lat, specs[i], spexhip.PREC_FP64)
        p = spexhip.QPipe(ctx, ref.desc, [])
        p.launch()
        p.wait()
        rchi, reinv = p.result()
        chi, einv = b.result(i)
        check[f"q{i}_nb{specs[i].nb}"] = {"chi0": errs(chi, rchi), "eps_inv": errs(einv, reinv)}
        p.destroy()
        ref.free()
    b.destroy()
    for r in reqs:
        r.free()
    lat.free()
    out = {"workload": f"32 q, Nk=64, Nocc=16, Nunocc~U[200,600], Nb~U[600,1600], Nw=12, {precision}",
           "kpacket": kp, "packets": npacket, "create_s": round(t_create, 2),
           "ms_eager": round(ms_eager, 1), "ms_graph": round(ms_graph, 1),
           "qpts_per_s": round(32e3 / ms_graph, 3), "gflops": round(flops / ms_graph / 1e6, 1),
           "err_vs_fp64_per_q": check}
    if cfg >= 0:
        tm, kc, pc = spexhip.chi0_configs()[cfg]
        out["config"] = {"index": cfg, "tile": tm, "kc": kc, "wg_per_cu": pc}
        out["autotune_ms"] = [round(t, 3) for t in tune]
    return out


PEAK_I8 = 5000.0  # TOPS dense int8 matrix, MI355X
FORMS = (("fp64", "PREC_FP64"), ("split_fp32", "PREC_SPLIT_FP32"), ("ozaki_int8", "PREC_OZAKI_INT8"))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    from project1_amd import spexhip

    ctx = spexhip.Context(0)
    if a.copy_calib:
        nbytes = a.copy_calib << 20
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        ctx.synth_normal(src, nbytes // 8, 1)
        for _ in range(5):
            ctx.copy(dst, src, nbytes)
        ctx.sync()
        print(json.dumps({"copy_kernel_bytes_each_way": nbytes, "launches": 5}))
        ctx.close()
        return
    ctx.set_workspace_limit(64 << 30)  # one frequency group for the Ozaki form at this size
    pipe = Pipeline(ctx, a)
    if a.only == "theta":
        print(json.dumps({"only": "theta", "theta_products": theta_leg(pipe, a)}), flush=True)
        pipe.free()
        ctx.close()
        return
    res, out = {}, {}
    flops = chi0_flops(a.nw, pipe.n, a.nb)
    for name, pname in FORMS:
        if a.only and name != a.only:
            continue
        prec = getattr(spexhip, pname)
        r = run_form(pipe, prec, a.steps, a.warmup)
        log(f"{name}: {r['ms_per_step']:.1f} ms/step (chi0 {r['chi0_ms']:.1f} ms)")
        out[name] = pipe.results()
        r["qpts_per_s"] = 1e3 / r["ms_per_step"]
        r["gflops"] = (flops + eps_flops(a.nw, a.nb)) / (r["ms_per_step"] * 1e6)
        r["chi0_tflops"] = flops / (r["chi0_ms"] * 1e9)
        if name == "ozaki_int8":
            prep, rank, groups = ctx.chi0_phase_ms()
            r.update(prep_ms=prep, rank_ms=rank, groups=groups)
        res[name] = r
    if a.only:
        print(json.dumps({"only": a.only, "forms": res}), flush=True)
        pipe.free()
        ctx.close()
        return
    errs_ = {k: {"chi0": errs(out[k][0], out["fp64"][0]), "eps_inv": errs(out[k][1], out["fp64"][1])}
             for k in ("split_fp32", "ozaki_int8")}

    def ok(k):
        e = errs_[k]
        return all(e[q][m] <= 1e-8 for q in ("chi0", "eps_inv") for m in ("max_rel", "fro_rel"))

    head_name = "ozaki_int8" if ok("ozaki_int8") else "split_fp32"
    head = res[head_name]
    if head_name == "ozaki_int8":
        alg = flops / (head["rank_ms"] * 1e9)  # algorithmic chi0 TFLOP/s
        i8 = 10.5 * alg  # 84 int8 ops per complex MAC (10.5 x the 8 algorithmic flop)
        roof = {"bound": "mfma", "kernel": "ozk_herk_kernel", "achieved": round(alg, 2),
                "peak": PEAK_I8, "unit": "TFLOP/s", "frac": round(alg / PEAK_I8, 4),
                "traffic": measured_traffic("ozk_herk_kernel", WORKLOAD),
                "flop_per_launch": flops / 2,
                "note": "achieved = algorithmic chi0 flop (8 per complex multiply-add of the "
                        "upper triangle, 8*Nw*n*Nb(Nb+1)/2 = %.4e per step, half per launch) / "
                        "rank-k phase time (HIP events on the context stream around its two "
                        "launches, one per half of the transitions; the slicing of the second "
                        "half runs beside the first launch), against the dense int8 MFMA peak"
                        % flops,
                "issued_int8": {"achieved": round(i8, 1), "peak": PEAK_I8,
                                "unit": "TOP/s (int8 ops issued)", "frac": round(i8 / PEAK_I8, 4),
                                "note": "84 int8 ops issued per complex multiply-add (10 slice "
                                        "pairs x 4 real products + 2 for the (3,3) pair, 2 ops "
                                        "each): the MFMA pipe's own utilisation"},
                "fp64_equivalent": {"achieved": round(alg, 2), "peak": PEAK_F64,
                                    "frac": round(alg / PEAK_F64, 4)}}
    else:
        roof = {"bound": "mfma", "kernel": "chi0_split_kernel",
                "achieved": round(head["chi0_tflops"], 2), "peak": PEAK_F32, "unit": "TFLOP/s",
                "frac": round(head["chi0_tflops"] / PEAK_F32, 4), "traffic": None}
    roof["fp64_kernel"] = {"kernel": "chi0_f64_kernel", "achieved": round(res["fp64"]["chi0_tflops"], 2),
                           "peak": PEAK_F64, "unit": "TFLOP/s",
                           "frac": round(res["fp64"]["chi0_tflops"] / PEAK_F64, 4)}
    if pipe.products:
        # products in HBM terms: M written once plus the wave functions read once
        pb = 16 * pipe.n * a.nb + pipe.sys.cmt.nbytes + pipe.sys.cpw.nbytes
        gbs = pb / (head["products_ms"] * 1e6)
        roof["products"] = {
            "kernel": "wave-function products (MT GEMMs + pw_pair_joint_kernel)",
            "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(gbs / 8000.0, 4), "bytes_per_step": int(pb),
            "note": "algorithmic bytes over the products time; the step is bound by its FP64 "
                    "MFMA DFT passes and load latency, not by HBM"}
        # the same stage in FP64 flop: the separable pruned DFT of every state (backward, to
        # the grid) and pair (forward, to the 16^3 IPW box), N*16*(N^2+16N+256) complex
        # multiply-adds each on the plan's N^3 grid, and the two MT GEMMs per atom
        N = pipe.pwp.grid[0]
        dft = N * 16 * (N * N + 16 * N + 256)
        nstates = a.nk * (a.nocc + a.nunocc)
        mt = 0
        for t in pipe.sys.types:
            lm = int(sum((2 * l + 1) * int(t["nindx"][l]) for l in range(t["lcut"] + 1)))
            llm = int(sum((2 * l + 1) * int(t["nindxm"][l]) for l in range(t["lcutm"] + 1)))
            # stage A (T against all (k, b1)) and stage B (one GEMM per (k, b1)), per atom
            mt += int(t["neq"]) * (lm * llm * lm * a.nk * a.nocc + llm * a.nunocc * lm * a.nk * a.nocc)
        fl = 8.0 * (dft * (pipe.n + nstates) + mt)
        tf = fl / (head["products_ms"] * 1e9)
        roof["products"]["fp64"] = {
            "achieved": round(tf, 2), "peak": PEAK_F64, "unit": "TFLOP/s",
            "frac": round(tf / PEAK_F64, 4), "flop_per_step": fl, "grid": N,
            "note": "algorithmic FP64 flop (8 per complex multiply-add) of the DFT passes and the "
                    "MT GEMMs over the products time, against the FP64 MFMA peak"}
    line = {
        "metric": METRIC,
        "value": round(head["qpts_per_s"], 4),
        "unit": "q-points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("int8x4-slice/int32-exact/fp64-recombine (eps in f64)" if head_name == "ozaki_int8"
                  else "fp32x3-split/fp64-acc (eps in f64)"),
        "data": ("synthetic (seeded wave functions -> products on device each step, random SPD "
                 "Coulomb matrix)" if pipe.products else
                 "synthetic (seeded Gaussian products on device, random SPD Coulomb matrix)"),
        "config": {"workload": WORKLOAD,
                   "nk": a.nk, "nocc": a.nocc, "nunocc": a.nunocc, "nb": a.nb, "nw": a.nw,
                   "parallelism": "1 GPU", "headline_form": head_name},
        "gflops": round(head["gflops"], 1),
        "latency_ms": {"p50": round(head["step_ms_p50"], 3), "p99": round(head["step_ms_p99"], 3)},
        "forms": {k: {kk: round(v, 4) for kk, v in r.items()} for k, r in res.items()},
        "speedup_vs_fp64": {k: round(res["fp64"]["ms_per_step"] / res[k]["ms_per_step"], 3)
                            for k in ("split_fp32", "ozaki_int8")},
        # the same ratio over chi0 + eps^-1 alone (the products stage is the
        # same FP64 code in every form)
        "speedup_vs_fp64_chi0_eps": {
            k: round((res["fp64"]["chi0_ms"] + res["fp64"]["eps_ms"])
                     / (res[k]["chi0_ms"] + res[k]["eps_ms"]), 3)
            for k in ("split_fp32", "ozaki_int8")},
        "err_vs_fp64": errs_,
        "roofline": roof,
    }
    if pipe.products:
        line["theta_products"] = theta_leg(pipe, a)
        log(f"theta {json.dumps(line['theta_products'])}")
    line["gamma_headwings"] = gamma_leg(pipe, a)
    line["hilbert_weights"] = hilbert_leg(pipe, a)
    log(f"hilbert {json.dumps(line['hilbert_weights'])}")
    log(f"gamma {json.dumps(line['gamma_headwings'])}")
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pipe, a)
    pipe.free()
    if not a.no_batch:
        line["batch"] = {f: batch_leg(ctx, a, f) for f in ("split_fp32", "ozaki_int8")}
        log(f"batch {json.dumps(line['batch'])}")
    if not a.no_serve:
        from project1_amd import serving

        line["serving"] = serving.serve_benchmark(
            ctx, {"fp64": spexhip.PREC_FP64, "split_fp32": spexhip.PREC_SPLIT_FP32,
                  "ozaki_int8": spexhip.PREC_OZAKI_INT8},
            a.serve_seconds, a.serve_cap, log=lambda m: log(m) if isinstance(m, str) else None)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
