"""Python driver for the request loop (qj_server_*): submit field solves, poll
completions, fetch kept fields, and replay a timed request stream."""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from ._lib import FP32_BAND, MAP_BLACK_HOLE, check


class Server:
    def __init__(self, hbm_budget_bytes=8 << 30, chunk_voxels=65536, chunks_per_launch=512, max_roots=256,
                 inflight_launches=2, precision=FP32_BAND, band=0.05, device=0, launch_work=0.0,
                 max_pending_work=0.0, aging_s=0.0, check_samples=0, escalate_abs_dF=0.0):
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), int(chunk_voxels), int(chunks_per_launch), int(max_roots),
                                int(inflight_launches), int(precision), float(band), int(device), float(launch_work),
                                float(max_pending_work), float(aging_s), int(check_samples), float(escalate_abs_dF))
        h = ctypes.c_void_p()
        check(_lib.load().qj_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._keep = {}  # id -> arrays kept alive until the request is copied (submit copies synchronously)

    def now(self):
        return _lib.load().qj_server_now(self.handle)

    def submit(self, w, keep_field=False, arrival_time=None, flags=MAP_BLACK_HOLE):
        """Submit a workloads.Workload; returns the request id. Raises QJError
        with status 7 (QJ_ERR_BUSY) when the server refuses it."""
        tr = np.ascontiguousarray(w.top.roots, dtype=np.float64)
        tp = np.ascontiguousarray(w.top.powers, dtype=np.float64)
        br = np.ascontiguousarray(w.bottom.roots, dtype=np.float64).reshape(-1, 4)
        bp = np.ascontiguousarray(w.bottom.powers, dtype=np.float64)
        req = _lib.Request()
        req.grid.res[:] = [int(v) for v in w.res]
        req.grid.center[:] = list(w.center)
        req.grid.lengths[:] = list(w.lengths)
        req.grid.rotation[:] = [1.0, 0.0, 0.0, 0.0]
        req.top = _lib.Poly(tr.shape[0], tr.ctypes.data_as(_lib.c_double_p), tp.ctypes.data_as(_lib.c_double_p),
                            float(w.top.power_scalar))
        req.bottom = _lib.Poly(br.shape[0], br.ctypes.data_as(_lib.c_double_p) if br.shape[0] else None,
                               bp.ctypes.data_as(_lib.c_double_p) if br.shape[0] else None,
                               float(w.bottom.power_scalar))
        req.params = _lib.MapParams(int(w.max_iter), float(w.escape), float(w.exp_scaling), float(w.slice),
                                    int(flags))
        req.keep_field = 1 if keep_field else 0
        req.arrival_time = self.now() if arrival_time is None else float(arrival_time)
        rid = ctypes.c_uint64()
        check(_lib.load().qj_server_submit(self.handle, ctypes.byref(req), ctypes.byref(rid)))
        return rid.value

    def poll(self, max_n=256, timeout=0.0):
        buf = (_lib.Completion * max_n)()
        n = ctypes.c_int32()
        check(_lib.load().qj_server_poll(self.handle, buf, max_n, ctypes.byref(n), float(timeout)))
        return [buf[i] for i in range(n.value)]

    def fetch(self, completion, device=True):
        n = completion.voxels
        if device:
            out = torch.empty(n, dtype=torch.float64, device="cuda")
            ptr = ctypes.c_void_p(out.data_ptr())
        else:
            out = np.empty(n, dtype=np.float64)
            ptr = out.ctypes.data_as(ctypes.c_void_p)
        check(_lib.load().qj_server_fetch(self.handle, ctypes.byref(completion), ptr, None))
        return out

    def release(self, rid):
        check(_lib.load().qj_server_release(self.handle, ctypes.c_uint64(rid)))

    def wait_all(self, n, timeout=600.0):
        got = []
        t_end = time.time() + timeout
        while len(got) < n and time.time() < t_end:
            got.extend(self.poll(256, 0.05))
        return got

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().qj_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def replay(server, timed_requests, log=None, flags=MAP_BLACK_HOLE):
    """Submit (t, workload) pairs at their offsets on the server clock; return
    (completions, id -> workload, start, refused count).

    Latency of a request = done_time - arrival_time, both on the server clock;
    arrival_time is the scheduled arrival, so submission delay counts against it."""
    start = server.now() + 0.05
    ids = {}
    done = []
    refused = 0
    for i, (t, w) in enumerate(timed_requests):
        target = start + t
        while True:
            now = server.now()
            if now >= target:
                break
            done.extend(server.poll(256, 0.0))
            time.sleep(min(0.0005, target - now))
        try:
            rid = server.submit(w, keep_field=False, arrival_time=target, flags=flags)
        except _lib.QJError as e:
            if e.status != 7:
                raise
            refused += 1
            continue
        ids[rid] = w
        if log and i % 200 == 0:
            log("submitted %d/%d at t=%.2fs, completed %d" % (i, len(timed_requests), server.now() - start, len(done)))
    while len(done) < len(ids):
        got = server.poll(256, 0.1)
        done.extend(got)
        if log and got and len(done) % 200 < len(got):
            log("completed %d/%d at t=%.2fs" % (len(done), len(timed_requests), server.now() - start))
    return done, ids, start, refused


def latency_summary(done, ids):
    lat = {}
    allv = []
    bad = 0
    for c in done:
        if c.status != 0:
            bad += 1
            continue
        w = ids[c.id]
        l = c.done_time - c.arrival_time
        allv.append(l)
        lat.setdefault(int(w.res[0]), []).append(l)
    out = {"all": _pct(allv), "failed": bad}
    for e in sorted(lat):
        out["%d^3" % e] = _pct(lat[e])
    return out


def _pct(v):
    if not v:
        return None
    a = np.array(v) * 1e3
    return {"n": int(a.size), "p50_ms": round(float(np.percentile(a, 50)), 3),
            "p99_ms": round(float(np.percentile(a, 99)), 3), "max_ms": round(float(a.max()), 3)}
