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


def _direct(w, precision):
    g = api.make_grid(w.res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
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
