"""GPU tests of the serving loop: results equal alone and in a mix (bitwise),
host-array requests match the CPU oracle, admission by HBM footprint under a
small budget (rejection and memory reuse), and the serving workload -- Poisson
arrivals, log-uniform integer edges in [32, 512] (primes included: those
lines go through Bluestein), half FP64 and half FP32 + 2 refinement steps, f = the benchmark field plus 16
random modes -- with every result checked against its exact potential."""
import math
import time

import numpy as np
import pytest

import fishfry_amd as ff
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)


def random_field(n, rng, nmodes=16, kmax=12):
    f = ff.field_default(n)
    modes, seen = [], set()
    while len(modes) < nmodes:
        k = tuple(int(v) for v in rng.integers(-kmax, kmax + 1, size=3))
        if k == (0, 0, 0) or not (1 <= math.sqrt(sum(v * v for v in k)) <= kmax):
            continue
        if k in seen or tuple(-v for v in k) in seen:
            continue
        seen.add(k)
        modes.append((k, float(rng.standard_normal()), float(rng.standard_normal())))
    return ff.field_set_modes(f, modes)


def test_alone_equals_in_mix_bitwise():
    rng = np.random.default_rng(11)
    srv = ff.Server(hbm_budget=8 << 30, streams=4, max_batch=8)
    targets = [((48, 40, 36), ff.FF_FP64, 0), ((40, 32, 50), ff.FF_FP32_REFINE, 2)]
    fields = [random_field(n, rng) for n, _, _ in targets]
    alone = []
    for (n, prec, steps), fld in zip(targets, fields):
        phi = np.empty(n)
        rid = srv.submit(n, prec, steps, field=fld, phi_host=phi)
        srv.wait(rid)
        alone.append(phi)
    # the same requests among 30 others of mixed shapes and precisions
    ids, outs = [], {}
    for k in range(32):
        if k in (7, 19):
            j = 0 if k == 7 else 1
            n, prec, steps = targets[j]
            phi = np.empty(n)
            ids.append(("t%d" % j, srv.submit(n, prec, steps, field=fields[j], phi_host=phi)))
            outs["t%d" % j] = phi
        else:
            n = tuple(int(v) for v in rng.integers(16, 96, size=3))
            prec = ff.FF_FP64 if k % 2 else ff.FF_FP32_REFINE
            ids.append(("o", srv.submit(n, prec, 2, field=random_field(n, rng))))
    for tag, rid in ids:
        srv.wait(rid)
    st = srv.stats()
    say("mix stats", st)
    assert np.array_equal(outs["t0"], alone[0])
    assert np.array_equal(outs["t1"], alone[1])
    assert st["completed"] == 34 and st["groups"] >= 2
    srv.close()


def test_large_alone_bitwise():
    """FF_SERVE_LARGE_ALONE: large requests (FP32 + 2 steps at 264^3, FP64 at
    400^3: above 3e8 point-passes) run with only small ones beside them, a
    medium one (FP32 + 2 at 200^3, 1.4e8 point-passes) waits while a large
    one runs, and every result equals the same request solved on a default
    server."""
    rng = np.random.default_rng(23)
    reqs = []
    for k in range(12):
        if k in (2, 3, 7):
            n, prec = {2: ((264, 264, 264), ff.FF_FP32_REFINE), 3: ((200, 200, 200), ff.FF_FP32_REFINE),
                       7: ((400, 400, 400), ff.FF_FP64)}[k]
        else:
            n = tuple(int(v) for v in rng.integers(24, 80, size=3))
            prec = ff.FF_FP64 if k % 2 else ff.FF_FP32_REFINE
        reqs.append((n, prec, random_field(n, rng)))
    outs = {}
    for flags in (0, ff.FF_SERVE_LARGE_ALONE):
        srv = ff.Server(hbm_budget=16 << 30, streams=4, max_batch=2, flags=flags)
        ids = []
        for n, prec, fld in reqs:
            phi = np.empty(n)
            ids.append((srv.submit(n, prec, 2, field=fld, phi_host=phi), phi))
        for rid, _ in ids:
            srv.wait(rid, timeout=120.0)
        say("flags", flags, srv.stats())
        outs[flags] = [phi for _, phi in ids]
        srv.close()
    for a, b in zip(outs[0], outs[ff.FF_SERVE_LARGE_ALONE]):
        assert np.array_equal(a, b)
    with pytest.raises(ff.FishFryError):
        ff.Server(flags=2)


def test_grouped_solves_bitwise():
    """Requests of one shape, precision and step count that meet in one
    dispatch group run each pass step as one launch for all of them (their
    buffers and domain constants from a per-instance table): the results are
    the same requests' results solved one at a time, bit for bit, FP64 and
    FP32 + 2 steps, with two domains in one group."""
    rng = np.random.default_rng(15)
    n = (40, 36, 32)
    reqs = []
    for k in range(10):
        prec, steps = (ff.FF_FP64, 0) if k < 5 else (ff.FF_FP32_REFINE, 2)
        fld = random_field(n, rng)
        if k % 2:
            for a in range(3):
                fld.d[a] *= 1.5  # another domain: other filter constants
        reqs.append((prec, steps, fld))
    outs = []
    for max_batch, streams in ((1, 1), (8, 1)):
        srv = ff.Server(hbm_budget=2 << 30, streams=streams, max_batch=max_batch, batch_window_us=5000.0)
        phis = [np.empty(n) for _ in reqs]
        ids = [srv.submit(n, prec, steps, field=fld, phi_host=phi) for (prec, steps, fld), phi in zip(reqs, phis)]
        for rid in ids:
            assert srv.wait(rid)["status"] == 0
        st = srv.stats()
        say("grouped", max_batch, st)
        if max_batch == 1:
            assert st["grouped_solves"] == 0
        else:
            assert st["grouped_solves"] >= 4
        outs.append(phis)
        srv.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_host_requests_match_oracle():
    rng = np.random.default_rng(12)
    srv = ff.Server(hbm_budget=2 << 30)
    shapes = [(32, 32, 32), (24, 20, 18), (30, 25, 27), (64, 16, 40)]
    fs = [rng.standard_normal(n) for n in shapes]
    phis = [np.empty(n) for n in shapes]
    ids = [srv.submit(n, f_host=f, phi_host=p) for n, f, p in zip(shapes, fs, phis)]
    for rid in ids:
        r = srv.wait(rid)
        assert r["status"] == 0 and r["latency_s"] >= r["queue_s"] >= 0
    for f, p in zip(fs, phis):
        want = oracle.solve(f)
        assert np.linalg.norm(p - want) / np.linalg.norm(want) < 1e-13
    srv.close()


def test_four_step_z_requests():
    """Grids whose z rows run in four steps get their z work array from the
    arena per request (the footprint counts it, after the FP32 refinement
    work area), so same-shape requests on different streams do not share
    one: every result matches the oracle, FP64 and FP32 + 2 steps, an odd nz
    with nx*ny odd included (its work array has a tail for the lone last
    row)."""
    rng = np.random.default_rng(14)
    n = (3, 2, 20000)
    srv = ff.Server(hbm_budget=1 << 30, streams=4, max_batch=1)
    nh = n[2] // 2 + 1
    block = lambda b: (b + 255) // 256 * 256
    assert srv.footprint(n) == 3 * block(16 * n[0] * n[1] * nh)
    refine = block(16 * n[0] * n[1] * nh + block(4 * n[0] * n[1] * n[2]) + 256)
    assert srv.footprint(n, ff.FF_FP32_REFINE, 2) == \
        2 * block(16 * n[0] * n[1] * nh) + block(refine + 16 * n[0] * n[1] * nh)
    fs = [rng.standard_normal(n) for _ in range(6)]
    phis = [np.empty(n) for _ in fs]
    ids = [srv.submit(n, f_host=f, phi_host=p) for f, p in zip(fs, phis)]
    for rid in ids:
        assert srv.wait(rid)["status"] == 0
    for f, p in zip(fs, phis):
        want = oracle.solve(f)
        assert np.linalg.norm(p - want) / np.linalg.norm(want) < 1e-13
    phis32 = [np.empty(n) for _ in range(3)]
    ids = [srv.submit(n, ff.FF_FP32_REFINE, 2, f_host=f, phi_host=p) for f, p in zip(fs, phis32)]
    for rid in ids:
        assert srv.wait(rid)["status"] == 0
    for f, p in zip(fs, phis32):
        want = oracle.solve(f)
        assert np.linalg.norm(p - want) / np.linalg.norm(want) < 1e-12
    # odd nz with nx*ny odd: the work array has a tail for the lone last row
    n3 = (3, 3, 16875)
    nh3 = n3[2] // 2 + 1
    Nh3 = n3[0] * n3[1] * nh3
    assert srv.footprint(n3) == 2 * block(16 * Nh3) + block(16 * Nh3) + block(16 * n3[2]) + block(32 * nh3)
    fs3 = [rng.standard_normal(n3) for _ in range(3)]
    phis3 = [np.empty(n3) for _ in fs3]
    ids = [srv.submit(n3, f_host=f, phi_host=p) for f, p in zip(fs3, phis3)]
    p32 = np.empty(n3)
    ids.append(srv.submit(n3, ff.FF_FP32_REFINE, 2, f_host=fs3[0], phi_host=p32))
    for rid in ids:
        assert srv.wait(rid)["status"] == 0
    for f, p in zip(fs3, phis3):
        want = oracle.solve(f)
        assert np.linalg.norm(p - want) / np.linalg.norm(want) < 1e-13
    want = oracle.solve(fs3[0])
    assert np.linalg.norm(p32 - want) / np.linalg.norm(want) < 1e-12
    srv.close()


def test_admission_by_footprint():
    rng = np.random.default_rng(13)
    n = (64, 64, 64)
    probe = ff.Server(hbm_budget=1 << 20)
    fp = probe.footprint(n, field=ff.field_default(n), want_norms=True)
    with pytest.raises(ff.FishFryError) as e:
        probe.submit(n, field=ff.field_default(n))
    assert e.value.status == ff.FF_ERR_REJECTED
    assert probe.stats()["rejected"] == 1
    probe.close()
    # room for two such requests at a time: twenty still complete, memory reused
    srv = ff.Server(hbm_budget=2 * fp + (1 << 16), streams=2, max_batch=4)
    ids = [srv.submit(n, field=random_field(n, rng), want_norms=True) for _ in range(20)]
    worst = max(srv.wait(rid)["norms"][2] for rid in ids)
    st = srv.stats()
    say("admission", fp, st, worst)
    assert st["completed"] == 20 and st["arena_peak_bytes"] <= st["arena_bytes"]
    assert worst < 1e-13
    srv.close()


def test_serving_workload():
    rng = np.random.default_rng(20261202)
    nreq, rate = 48, 400.0  # requests, arrivals per second
    srv = ff.Server(hbm_budget=24 << 30, streams=4, max_batch=8, batch_window_us=300.0)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs = []
    for k in range(nreq):
        if rng.random() < 0.5:
            e = int(round(math.exp(rng.uniform(math.log(32), math.log(512)))))
            n = (e, e, e)
        else:
            n = tuple(int(round(math.exp(rng.uniform(math.log(32), math.log(512))))) for _ in range(3))
        prec = ff.FF_FP64 if k % 2 == 0 else ff.FF_FP32_REFINE
        reqs.append((n, prec))
    t0 = time.perf_counter()
    ids = []
    for (n, prec), ta in zip(reqs, arrivals):
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append((n, prec, srv.submit(n, prec, 2, field=random_field(n, rng), want_norms=True)))
    worst = {ff.FF_FP64: 0.0, ff.FF_FP32_REFINE: 0.0}
    for n, prec, rid in ids:
        r = srv.wait(rid, timeout=300.0)
        worst[prec] = max(worst[prec], r["norms"][2])
    st = srv.stats()
    say("serving", st, "worst Linf", worst)
    assert st["completed"] == nreq
    assert worst[ff.FF_FP64] < 1e-13 and worst[ff.FF_FP32_REFINE] < 1e-13
    srv.close()


def test_plan_cache_eviction():
    # more distinct shapes than the server keeps plans for: the first shape's
    # plan is evicted and rebuilt, and its result is bitwise the same
    rng = np.random.default_rng(14)
    srv = ff.Server(hbm_budget=1 << 30, streams=2, max_batch=4)
    first = (20, 18, 22)
    fld = random_field(first, rng)
    a, b = np.empty(first), np.empty(first)
    srv.wait(srv.submit(first, ff.FF_FP32_REFINE, 2, field=fld, phi_host=a))
    ids = []
    for k in range(70):
        n = (32 + k % 7, 32 + (k // 7) % 10, 32 + k % 5)  # the benchmark f is at round-off from 32 up
        prec = ff.FF_FP32_REFINE if k % 2 else ff.FF_FP64
        ids.append(srv.submit(n, prec, 2, field=random_field(n, rng, kmax=3), want_norms=True))
    worst = max(srv.wait(rid)["norms"][2] for rid in ids)
    srv.wait(srv.submit(first, ff.FF_FP32_REFINE, 2, field=fld, phi_host=b))
    st = srv.stats()
    say("eviction", st, worst)
    assert st["completed"] == 72 and worst < 1e-13
    assert np.array_equal(a, b)
    srv.close()


def test_wait_semantics():
    srv = ff.Server(hbm_budget=4 << 30, streams=2)
    n = (256, 256, 256)
    rid = srv.submit(n, field=ff.field_default(n), want_norms=True)
    with pytest.raises(ff.FishFryError) as e:
        srv.wait(rid, timeout=1e-6)  # the request cannot complete within a microsecond
    assert e.value.status == ff.FF_ERR_TIMEOUT
    r = srv.wait(rid)
    assert r["status"] == 0 and r["norms"][2] < 1e-13 and r["queue_s"] >= 0
    with pytest.raises(ff.FishFryError) as e:
        srv.wait(rid)  # a result is handed out once
    assert e.value.status == ff.FF_ERR_NOT_FOUND
    with pytest.raises(ff.FishFryError) as e:
        srv.wait(123456789)
    assert e.value.status == ff.FF_ERR_NOT_FOUND
    srv.close()
