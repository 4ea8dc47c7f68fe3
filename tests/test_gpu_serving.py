"""Continuous-batching server: each result is the same alone and in the mix,
admission respects the HBM budget, and a served request matches the oracle
chain."""
import numpy as np
import pytest

from project1_amd import serving, spexhip, synth
from tests.chain import einv_tol, oracle_for_request
from tests.conftest import normwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lattice(ctx):
    lat = serving.Lattice(ctx, nocc=4, nunocc=12, lcut=2, nrad=2, gcut=3.1, nfft=(16, 16, 16),
                          max_atoms=3, nb_max=100, ipw_target=40, seed=5)
    yield lat
    lat.free()


def small_mix(count=8, seed=3):
    return serving.request_mix(count, seed, meshes=(2, 3), nb_range=(50, 100), nw_choices=(2, 3))


@pytest.mark.parametrize("precision", [spexhip.PREC_FP64, spexhip.PREC_SPLIT_FP32,
                                       spexhip.PREC_OZAKI_INT8])
def test_same_result_alone_and_in_mix(lattice, precision):
    """Tolerance 0: pipelines share no buffers and every kernel's reduction
    order is fixed, so concurrency cannot change a single bit."""
    specs = small_mix()
    alone, ra = serving.run_closed(lattice, specs, precision, max_inflight=1,
                                   hbm_budget=64 << 30, keep_results=True)
    mix, rm = serving.run_closed(lattice, specs, precision, max_inflight=4,
                                 hbm_budget=64 << 30, keep_results=True)
    print(f"prec={precision} alone {alone}\nmix {mix}")
    assert alone["completed"] == mix["completed"] == len(specs)
    assert alone["peak_inflight"] == 1
    for i in range(len(specs)):
        np.testing.assert_array_equal(rm[i][0], ra[i][0])
        np.testing.assert_array_equal(rm[i][1], ra[i][1])


def test_admission_by_footprint(lattice):
    """A budget that holds one pipeline at a time serialises the requests
    (and never exceeds the budget); a request larger than it is refused."""
    specs = [serving.Spec(nmesh=2, nb=80, nw=3, seed=100 + i) for i in range(4)]
    req = serving.Request(lattice, specs[0], spexhip.PREC_FP64)
    fp = req.footprint
    budget = 2 * fp - 1  # two never fit together
    res, _ = serving.run_closed(lattice, specs, spexhip.PREC_FP64, max_inflight=4,
                                hbm_budget=budget)
    assert res["completed"] == len(specs) and res["errors"] == 0
    assert res["peak_inflight"] == 1
    assert res["peak_hbm_bytes"] <= budget
    srv = spexhip.Server(lattice.ctx, 2, fp - 1)
    with pytest.raises(spexhip.SpexError) as e:
        srv.submit(req.desc)
    assert e.value.code == spexhip.E_NOMEM
    srv.destroy()
    req.free()


def test_served_request_matches_oracle(lattice):
    spec = serving.Spec(nmesh=2, nb=77, nw=3, seed=42, iq=3)
    _, res = serving.run_closed(lattice, [spec], spexhip.PREC_FP64, 2, 64 << 30,
                                keep_results=True)
    chi, einv = res[0]
    req = serving.Request(lattice, spec, spexhip.PREC_FP64)  # same seeded inputs
    rchi, reinv, cond = oracle_for_request(lattice, req)
    req.free()
    e1, e2 = normwise(chi, rchi), normwise(einv, reinv)
    print(f"served request vs oracle: chi0 {e1} eps^-1 {e2} cond(eps) {cond:.2e}")
    assert max(e1) < 1e-12 and max(e2) < einv_tol(cond, spec.nb)


def test_device_generator_matches_host_mirror(ctx):
    rows, cols, nmat = 37, 5, 3
    scale = np.random.default_rng(1).uniform(0, 2, (rows, nmat))
    scale[30:, 1] = 0.0
    ds = ctx.to_dev(np.asfortranarray(scale).ravel(order="F"))
    out = ctx.alloc(16 * rows * cols * nmat)
    ctx.synth_normal_rows(out, rows, cols, nmat, 77, ds)
    got = out.download(np.complex128, rows * cols * nmat).reshape((rows, cols, nmat), order="F")
    ref = synth.device_normal(rows * cols * nmat, 77).reshape((rows, cols, nmat), order="F")
    ref *= scale[:, None, :]
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13)
    plain = ctx.alloc(16 * 50)
    ctx.synth_normal(plain, 100, 77, 1.0)
    np.testing.assert_allclose(plain.download(np.complex128, 50), synth.device_normal(50, 77),
                               rtol=1e-13, atol=1e-13)
    for b in (ds, out, plain):
        b.free()


def test_context_close_drains_a_live_server():
    """Closing a context whose server still has requests queued and running
    stops the scheduler and drains the pipelines before any memory is freed
    (regression: the context used to be freed under a live scheduler
    thread).  Uses a context of its own."""
    ctx2 = spexhip.Context(0)
    lat = serving.Lattice(ctx2, nocc=4, nunocc=12, lcut=2, nrad=2, gcut=3.1, nfft=(16, 16, 16),
                          max_atoms=3, nb_max=100, ipw_target=40, seed=6)
    srv = spexhip.Server(ctx2, 2, 64 << 30)
    reqs = [serving.Request(lat, s, spexhip.PREC_FP64) for s in small_mix(6, seed=9)]
    ids = [srv.submit(r.desc) for r in reqs]
    assert srv.state(ids[0]) in (spexhip.Server.QUEUED, spexhip.Server.RUNNING,
                                 spexhip.Server.DONE)
    ctx2.close()
    assert srv.h is None
    srv.destroy()  # a no-op once the context has taken the server down


def test_staging_oom_drains_then_raises_and_context_closes():
    """A context whose HBM budget holds the lattice and two small pipelines
    but never the wave functions of an 8x8x8 k mesh: the open-loop driver
    stages and submits two small requests, then cannot stage the third (its
    mesh does not fit) within its patience.  It must drain the two submitted
    requests before raising SPEXHIP_E_NOMEM, leave nothing of theirs in use,
    and the context must then close cleanly."""
    kw = dict(nocc=4, nunocc=12, lcut=2, nrad=2, gcut=3.1, nfft=(16, 16, 16), max_atoms=3,
              nb_max=100, ipw_target=40, seed=6)
    small = [serving.Spec(nmesh=2, nb=80, nw=3, seed=200 + i) for i in range(2)]
    big = serving.Spec(nmesh=8, nb=80, nw=3, seed=210)
    with spexhip.Context(0) as probe:  # sizes, on an unlimited context
        lat = serving.Lattice(probe, **kw)
        lat.mesh(2)
        lat.mt_plan(lat.natoms_for(80))
        base = probe.mem_used()[0]
        req = serving.Request(lat, small[0], spexhip.PREC_FP64)
        per_req = probe.mem_used()[0] - base + req.footprint
        req.free()
        before = probe.mem_used()[0]
        lat.mesh(8)
        mesh8 = probe.mem_used()[0] - before
        lat.free()
    assert mesh8 > 2 * per_req + (1 << 20), (mesh8, per_req)
    budget = base + 2 * per_req + (mesh8 - 2 * per_req) // 2
    ctx2 = spexhip.Context(0, hbm_budget=budget)
    lat = serving.Lattice(ctx2, **kw)
    lat.mesh(2)
    lat.mt_plan(lat.natoms_for(80))
    used0 = ctx2.mem_used()[0]
    with pytest.raises(spexhip.SpexError) as e:
        serving.run_poisson(lat, small + [big], spexhip.PREC_FP64, rate=1e4, seconds=1.0,
                            max_inflight=2, hbm_budget=budget, patience_s=1.0)
    assert e.value.code == spexhip.E_NOMEM
    part = e.value.partial
    print("partial", part)
    assert part["completed"] == 2 and part["errors"] == 0
    # inputs and pipelines of the drained requests are released (freed blocks
    # go to the cache, which is not counted as used); the failed mesh left
    # nothing behind
    assert 8 not in lat.meshes
    assert ctx2.mem_used()[0] <= used0 + per_req
    lat.free()
    ctx2.close()
