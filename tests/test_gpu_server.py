"""Request loop: results are identical alone and in a mixed batch (bit for bit,
the per-voxel arithmetic does not depend on batching), agree with the direct
field evaluation, and admission respects the HBM budget."""
import numpy as np
import pytest
import torch

from quijibo_amd import api, server as S, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, MAP_BLACK_HOLE

pytestmark = pytest.mark.gpu


def _mixed(n=12, seed=5):
    """Requests of all four FP32 variants in one stream: exponents 1 and 2 with
    escape 2000 (renormalise every 4 terms), integer exponents with escape 1e6
    (term bound over 30 bits: every 2 terms), general exponents, and exponents
    all 1 with escape 2000 (every 8 terms)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        e = int(rng.choice([24, 40, 64, 80]))
        nr = int(rng.integers(8, 64))
        nt = nr // 2
        kind = i % 4
        tp = rng.integers(1, 3, size=nt).astype(float)
        if kind == 2:
            tp = rng.choice([1.0, 1.5, 2.0, 0.75], size=nt)
        if kind == 3:
            tp = np.ones(nt)
        top = W.RootSet(rng.normal(scale=0.7, size=(nt, 4)), tp)
        bot = W.RootSet(rng.normal(scale=0.7, size=(nr - nt, 4)), np.ones(nr - nt))
        w = W.Workload("m%d" % i, (e, e + 3, e - 5), top, bot, max_iter=int(rng.integers(3, 6)),
                       escape=1e6 if kind == 1 else 2000.0)
        out.append(w)
    return out


def _direct(w, precision, flags=MAP_BLACK_HOLE):
    g = api.make_grid(w.res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, flags)
    return api.field_eval(g, r, p, precision, 0.05).cpu().numpy()


@pytest.mark.parametrize("precision", [FP32_BAND, FP64])
def test_alone_and_in_the_mix(gpu, precision):
    reqs = _mixed()
    srv = S.Server(hbm_budget_bytes=1 << 30, chunk_voxels=8192, chunks_per_launch=64, precision=precision)
    ids = [srv.submit(w, keep_field=True) for w in reqs]
    done = srv.wait_all(len(reqs), timeout=120)
    assert len(done) == len(reqs) and all(c.status == 0 for c in done)
    by_id = {c.id: c for c in done}
    mixed = {}
    for rid in ids:
        mixed[rid] = srv.fetch(by_id[rid], device=False)
        srv.release(rid)
    srv.close()
    for rid, w in zip(ids, reqs):
        alone_srv = S.Server(hbm_budget_bytes=256 << 20, chunk_voxels=4096, chunks_per_launch=16,
                             precision=precision)
        aid = alone_srv.submit(w, keep_field=True)
        c = alone_srv.wait_all(1, timeout=60)
        assert len(c) == 1 and c[0].status == 0
        alone = alone_srv.fetch(c[0], device=False)
        alone_srv.close()
        assert np.array_equal(alone, mixed[rid], equal_nan=True), w.name
        direct = _direct(w, precision)
        assert np.array_equal(direct, mixed[rid], equal_nan=True), w.name


def test_admission_by_memory_and_capacity(gpu):
    w = W.config1(64)  # 64^3 field = 2 MiB
    small = S.Server(hbm_budget_bytes=3 << 20, chunk_voxels=16384, chunks_per_launch=8)
    ids = [small.submit(w, keep_field=False) for _ in range(5)]
    done = small.wait_all(5, timeout=60)
    assert len(done) == 5 and all(c.status == 0 for c in done)
    # admissions are serialized by the 3 MiB budget: each admitted after the previous finished
    adm = sorted(c.admit_time for c in done)
    fin = sorted(c.done_time for c in done)
    for k in range(1, 5):
        assert adm[k] >= fin[k - 1] - 1e-6
    big = W.config1(128)  # 16 MiB > budget
    bid = small.submit(big)
    c = small.wait_all(1, timeout=30)
    assert c[0].id == bid and c[0].status == 4  # QJ_ERR_CAPACITY
    small.close()


def test_kept_field_blocks_until_released(gpu):
    w = W.config1(64)
    srv = S.Server(hbm_budget_bytes=3 << 20, chunk_voxels=16384, chunks_per_launch=8)
    a = srv.submit(w, keep_field=True)
    first = srv.wait_all(1, timeout=30)
    assert first[0].id == a and first[0].d_field
    b = srv.submit(w)
    assert srv.wait_all(1, timeout=0.5) == []  # no room while a's field is kept
    srv.release(a)
    second = srv.wait_all(1, timeout=30)
    assert second[0].id == b and second[0].status == 0
    srv.close()


def test_special_cases_match_the_direct_evaluation(gpu):
    """The request loop's band pass takes the same fast and guarded products
    as qj_field_eval: identical fields for roots on cell centres."""
    from test_gpu_field import _special_case_workload
    w = _special_case_workload()
    srv = S.Server(hbm_budget_bytes=64 << 20, chunk_voxels=256, chunks_per_launch=4, precision=FP32_BAND)
    rid = srv.submit(w, keep_field=True)
    c = srv.wait_all(1, timeout=60)
    assert len(c) == 1 and c[0].status == 0
    got = srv.fetch(c[0], device=False)
    srv.release(rid)
    srv.close()
    ref = _direct(w, FP32_BAND)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = np.isfinite(ref)
    assert np.array_equal(got[fin], ref[fin]) and np.array_equal(np.isinf(got), np.isinf(ref))


def test_lagrange_request_matches_the_direct_evaluation(gpu):
    """An FP64 request in the Lagrange order (QJ_MAP_LAGRANGE) through the
    request loop's FP64 kernel, against qj_field_eval with the same flags."""
    from quijibo_amd._lib import MAP_LAGRANGE
    w = W.config2(40)
    w.top.powers = w.top.powers * 1.1
    srv = S.Server(hbm_budget_bytes=64 << 20, chunk_voxels=4096, chunks_per_launch=8, precision=FP64)
    rid = srv.submit(w, keep_field=True, flags=MAP_BLACK_HOLE | MAP_LAGRANGE)
    c = srv.wait_all(1, timeout=60)
    assert len(c) == 1 and c[0].status == 0
    got = srv.fetch(c[0], device=False)
    srv.release(rid)
    srv.close()
    g = api.make_grid(w.res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    ref = api.field_eval(g, r, api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice,
                                               MAP_BLACK_HOLE | MAP_LAGRANGE), FP64).cpu().numpy()
    assert np.array_equal(got, ref, equal_nan=True)
    std = _direct(w, FP64)
    fin = np.isfinite(ref) & np.isfinite(std)
    assert np.max(np.abs(std[fin] - ref[fin])) > 1e-3  # the Lagrange order is a different map


def test_refusal_aging_and_per_result_check(gpu):
    """max_pending_work refuses a submission whose work queued ahead of it in
    shortest-first order (its own included) passes the limit (QJ_ERR_BUSY),
    but never one that arrives at an idle server, and not a smaller request
    that only the larger ones are ahead of; every completion carries its
    sampled FP64 cross-check."""
    from quijibo_amd._lib import QJError
    w = W.config2(64)
    work = 64 ** 3 * w.max_iter * (w.top.n + w.bottom.n)
    srv = S.Server(hbm_budget_bytes=256 << 20, chunk_voxels=16384, chunks_per_launch=8,
                   max_pending_work=1.5 * work, check_samples=700)
    big = W.config2(96)  # more work than the limit alone: accepted at an idle server
    a = srv.submit(big)
    with pytest.raises(QJError) as ei:
        srv.submit(big)  # the first one is ahead of it: 2 x 3.4 work
    assert ei.value.status == 7
    b = srv.submit(w)  # only its own work is ahead of it in shortest-first order
    done = srv.wait_all(2, timeout=60)
    assert sorted(c.id for c in done) == sorted([a, b]) and all(c.status == 0 for c in done)
    for c in done:
        assert c.checked == 700 and c.check_sign_flips == 0 and c.check_class_mismatch == 0
        assert 0 < c.check_max_abs_dF < 0.05 and c.check_max_rel_dF <= c.check_max_abs_dF
    d = srv.submit(big)  # idle again: accepted
    done = srv.wait_all(1, timeout=60)
    assert done[0].id == d and done[0].status == 0 and done[0].checked == 700
    srv.close()


def test_aged_request_is_admitted_first(gpu):
    """Once a waiting request is older than aging_s, requests of a quarter of
    its footprint or more are not admitted before it: every medium request
    admitted later than the big one's aging time (plus a margin for one
    admission pass) is admitted after it, and the big one is admitted while
    medium ones keep arriving (no starvation). Small requests are not held."""
    import time
    small = W.config1(16)   # 32 KiB field: below a quarter of the big one's footprint
    medium = W.config1(64)  # 2 MiB field: at least a quarter of it
    big = W.config1(96)     # 6.75 MiB field: fits the 8 MiB budget only beside no medium one
    aging = 0.05
    srv = S.Server(hbm_budget_bytes=8 << 20, chunk_voxels=16384, chunks_per_launch=4, aging_s=aging)
    ids_med = [srv.submit(medium) for _ in range(12)]  # 24 MiB: three budgets' worth
    bid = srv.submit(big)
    ids_small = []
    t_end = srv.now() + 0.3
    while srv.now() < t_end:
        ids_med.append(srv.submit(medium))
        ids_small.append(srv.submit(small))
        time.sleep(0.002)
    n = len(ids_med) + len(ids_small) + 1
    done = srv.wait_all(n, timeout=120)
    srv.close()
    assert len(done) == n and all(c.status == 0 for c in done)
    by = {c.id: c for c in done}
    b = by[bid]
    late = [by[i] for i in ids_med if by[i].admit_time > b.arrival_time + aging + 0.02]
    held_small = [by[i] for i in ids_small if b.arrival_time + aging + 0.02 < by[i].admit_time < b.admit_time]
    print("big waited %.1f ms; %d medium requests admitted after its aging time, %d small ones admitted while "
          "it waited" % (1e3 * (b.admit_time - b.arrival_time), len(late), len(held_small)), flush=True)
    assert all(c.admit_time >= b.admit_time for c in late)
    assert b.admit_time <= b.arrival_time + aging + 0.25


def test_config4_slice_against_the_oracle(gpu):
    """The first requests of the serving stream with their grids shrunk (edge /
    16) through the request loop, each result against the oracle: FP64 range
    against the FP64 oracle (zero sign flips, FP64 tolerance in the band) and
    the extended range against the long double oracle."""
    import oracle as O
    from quijibo_amd._lib import MAP_EXTENDED_RANGE
    reqs = [w for _, w in W.config4_requests(n=10)]
    for w in reqs:
        e = max(4, w.res[0] // 16)
        w.res = (e, e, e)
    srv = S.Server(hbm_budget_bytes=64 << 20, chunk_voxels=4096, chunks_per_launch=16, check_samples=256)
    sent = []
    for k, w in enumerate(reqs):
        flags = MAP_BLACK_HOLE | (MAP_EXTENDED_RANGE if k % 2 else 0)
        sent.append((srv.submit(w, keep_field=True, flags=flags), w, flags))
    done = {c.id: c for c in srv.wait_all(len(sent), timeout=120)}
    for rid, w, flags in sent:
        c = done[rid]
        assert c.status == 0
        got = srv.fetch(c, device=False)
        srv.release(rid)
        oflags = O.BLACK_HOLE | (O.EXTENDED if flags & MAP_EXTENDED_RANGE else 0)
        ref = O.field_map(O.make_grid(w.res, w.center, w.lengths), O.poly(w.top.roots, w.top.powers),
                          O.poly(w.bottom.roots, w.bottom.powers),
                          O.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, oflags))
        assert np.array_equal(np.isfinite(got), np.isfinite(ref)) and np.array_equal(np.isnan(got), np.isnan(ref))
        flips = int(np.count_nonzero((got.astype(np.float32) > 0) != (ref.astype(np.float32) > 0)))
        inband = np.abs(ref) < 0.04
        print("req %d (%d roots, %d^3, ext %d): flips %d, band %d, checked %d" %
              (rid, w.top.n + w.bottom.n, w.res[0], bool(flags & MAP_EXTENDED_RANGE), flips, inband.sum(),
               c.checked), flush=True)
        assert flips == 0
        assert np.all(np.abs(got[inband] - ref[inband]) <= 1e-8 * np.maximum(1.0, np.abs(ref[inband])))
        assert c.checked == min(256, got.size) and c.check_sign_flips == 0
    srv.close()


def test_escalation_recomputes_inaccurate_fp32_results_in_fp64(gpu):
    """A served FP32 result whose sampled cross-check misses escalate_abs_dF is
    computed again with every voxel in FP64 (exact integer powers): request 38
    of the config-4 stream (128^3, 109 + 109 roots, maxIter 5, extended range)
    has FP32 errors above 0.05 and sign flips; the escalated result equals the
    FP64-exact field and has none. An accurate request is not escalated."""
    from quijibo_amd._lib import FP64_EXACT, MAP_EXTENDED_RANGE
    flags = MAP_BLACK_HOLE | MAP_EXTENDED_RANGE
    hard = W.config4_requests()[38][1]  # the full 2000-request stream's draw
    assert hard.res == (128, 128, 128) and hard.max_iter == 5
    easy = W.config2(48)
    srv = S.Server(hbm_budget_bytes=256 << 20, chunk_voxels=16384, chunks_per_launch=16, check_samples=512,
                   escalate_abs_dF=0.025)
    a = srv.submit(hard, keep_field=True, flags=flags)
    b = srv.submit(easy, keep_field=True, flags=flags)
    done = {c.id: c for c in srv.wait_all(2, timeout=120)}
    assert done[a].status == 0 and done[b].status == 0
    print("hard: escalated %d, check max |dF| %.3g; easy: escalated %d, check max |dF| %.3g" % (
        done[a].escalated, done[a].check_max_abs_dF, done[b].escalated, done[b].check_max_abs_dF), flush=True)
    assert done[a].escalated == 1 and done[b].escalated == 0
    got = srv.fetch(done[a], device=False)
    srv.release(a)
    srv.release(b)
    srv.close()
    ref = _direct(hard, FP64_EXACT, flags=flags)
    # the second cross-check compares exact products with the polar form: FP64 rounding only
    assert done[a].check_sign_flips == 0 and done[a].check_max_rel_dF <= 1e-8
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    assert np.max(np.abs(got[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))) <= 1e-12
    assert np.count_nonzero((got.astype(np.float32) > 0) != (ref.astype(np.float32) > 0)) == 0
