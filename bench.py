"""Benchmark: rational-quaternion Julia field evaluation at 512^3 on one MI355X.

One step = one evaluation of the whole 512^3 field (config-2 roots: 32 top + 32
bottom Gaussian roots, powers in {1,2}, maxIter 4, escape 2000) with the FP32
bulk + FP64 band kernel pair. The plain all-FP64 kernel is timed beside it.
Inputs (roots) and the output field are resident in HBM during the timed region.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6    # MI355X published FP64 vector peak
# HBM bytes per k_grid32 launch of the 512^3 workload, measured with rocprofv3
# FETCH_SIZE and WRITE_SIZE (each scaled by the same counter on a 1 GiB copy):
# 0.160 MB fetched + 1113.24 MB written.
K_GRID32_TRAFFIC_BYTES = 1113401623
# Issued-instruction side of k_grid32 at 512^3 from rocprofv3 counters
# (SQ_INSTS_VALU, SQ_WAVES, GRBM_GUI_ACTIVE; profiles/k_grid32_counters_c2_512.md):
# VALU instructions per SIMD-cycle against the 0.48-0.5 a stream of independent
# FP32 FMAs issues on gfx950.
K_GRID32_VALU_BUSY = {"valu_insts_per_simd_cycle": 0.455, "fma_stream_insts_per_simd_cycle": 0.478,
                      "busy_frac": 0.95, "source": "profiles/k_grid32_counters_c2_512.md (three boxes: 0.45-0.46)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def basic_ops_per_voxel_iter(nt, nb):
    """Algorithmic IEEE basic ops per voxel-iteration of the reference method:
    per root term 4 (subtract) + 23 (pow) + 28 (Hamilton product); division 40,
    scaling 4, magnitude 8 (QUATERNION.cpp:154-200, 375-393)."""
    return 55 * (nt + nb) - 4 if nb > 0 else 55 * nt - 16


def cpu_baseline(w, gpu64, gpu32, slabs, threads):
    import oracle as O
    res = w.res
    g = O.make_grid(res, w.center, w.lengths)
    top = O.poly(w.top.roots, w.top.powers, w.top.power_scalar)
    bot = O.poly(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar)
    p = O.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, O.BLACK_HOLE)
    zs = np.linspace(0, res[2] - 1, slabs).round().astype(int)
    plane = res[0] * res[1]
    t0 = time.perf_counter()
    flips64 = flips32 = 0
    maxerr = 0.0
    for z in zs:
        f = O.field_map_zrange(g, top, bot, p, int(z), int(z) + 1, nthreads=threads)
        a = gpu64[z * plane:(z + 1) * plane]
        b = gpu32[z * plane:(z + 1) * plane]
        flips64 += int(np.count_nonzero((f.astype(np.float32) > 0) != (a.astype(np.float32) > 0)))
        flips32 += int(np.count_nonzero((f.astype(np.float32) > 0) != (b.astype(np.float32) > 0)))
        fin = np.isfinite(f) & np.isfinite(a)
        if fin.any():
            maxerr = max(maxerr, float(np.max(np.abs(f[fin] - a[fin]) / np.maximum(1.0, np.abs(f[fin])))))
    dt = time.perf_counter() - t0
    nominal = len(zs) * plane * w.max_iter
    return dict(value=nominal / dt / 1e9, seconds=dt, slabs=len(zs), flips_fp64=flips64, flips_fp32band=flips32,
                max_rel_err_fp64=maxerr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-slabs", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 1024^3 tiled section")
    ap.add_argument("--no-grad", action="store_true", help="skip the gradient section")
    ap.add_argument("--no-nlmc", action="store_true", help="skip the non-linear mesher section")
    ap.add_argument("--band", type=float, default=0.05)
    ap.add_argument("--serve", type=int, default=2000, help="requests in the serving stream (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--probe", action="store_true", help="counter probe: two field launches and a 1 GiB copy")
    ap.add_argument("--probe-precision", default="fp32_band", choices=["fp32_band", "fp64", "fp64_exact"])
    ap.add_argument("--mc-probe", action="store_true",
                    help="counter probe: marching cubes of one 256^3 torus tile (twice), then a 1 GiB copy")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving stream")
    ap.add_argument("--serve-launch-work", type=float, default=0.0,
                    help="root-term iterations per request-loop launch (0 = library default)")
    ap.add_argument("--serve-inflight", type=int, default=0, help="launches in flight (0 = library default)")
    ap.add_argument("--serve-pending-s", type=float, default=2.0,
                    help="refuse submissions beyond this many seconds of backlog (0 = never refuse)")
    ap.add_argument("--serve-aging-s", type=float, default=0.5, help="requests waiting longer go first")
    ap.add_argument("--serve-check", type=int, default=512, help="cross-check samples per served result")
    ap.add_argument("--serve-escalate", type=float, default=0.025,
                    help="second serving leg: results whose sampled max |dF| exceeds this are recomputed in "
                         "FP64 (0 = no second leg)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the field, mesher and gradient paths run on a single MI355X "
              "and have no multi-GPU launch; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from quijibo_amd import api, workloads as W
    from quijibo_amd._lib import FP32_BAND, FP64, FP64_EXACT, MAP_BLACK_HOLE

    torch.cuda.set_device(0)
    w = W.headline(args.res)
    nt, nb = w.top.n, w.bottom.n
    grid = api.make_grid(w.res, w.center, w.lengths)
    params = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    ws = api.Workspace()
    out32 = torch.empty(w.voxels, dtype=torch.float64, device="cuda")
    out64 = torch.empty(w.voxels, dtype=torch.float64, device="cuda")
    nominal = w.voxels * w.max_iter
    stream = torch.cuda.current_stream()

    if args.serve_only:
        out = {"serving": serve_stream(args, log, 0.0)}
        if args.serve_escalate > 0:
            out["serving_escalated"] = serve_stream(args, log, args.serve_escalate)
        print(json.dumps(out))
        return

    if args.mc_probe:
        print(json.dumps({"mc_probe": mc_probe(args, log)}))
        return

    if args.probe:
        # Counter probe for rocprofv3 --pmc: one FP32+band field after a warm-up, then a
        # plain 1 GiB device-to-device copy whose FETCH_SIZE / WRITE_SIZE readings give
        # the scale factor for the field kernel's readings.
        prec = {"fp32_band": FP32_BAND, "fp64": FP64, "fp64_exact": FP64_EXACT}[args.probe_precision]
        for i in range(2):
            log("probe: %s field %d" % (args.probe_precision, i))
            api.field_eval(grid, r, params, prec, args.band, out=out32 if prec == FP32_BAND else out64,
                           workspace=ws)
            torch.cuda.synchronize()
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        log("probe: copy 1 GiB")
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "%s field %d^3 x2, then copy of %d bytes" % (args.probe_precision, args.res,
                                                                                 8 << 27)}))
        return

    # plain FP64 kernel
    log("fp64: warmup")
    api.field_eval(grid, r, params, FP64, out=out64, workspace=ws)
    torch.cuda.synchronize()
    ws.reset_stats()
    api._lib.load().qj_workspace_timing(ws.handle, 1)
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        api.field_eval(grid, r, params, FP64, out=out64, workspace=ws)
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / args.fp64_steps
    st64 = ws.stats()
    k64 = _kernel_ms(api, ws)
    log("fp64: %.2f ms/step" % (t64 * 1e3))

    # FP64 with exponents 1 and 2 as exact products: the same precision as the
    # plain FP64 kernel, the algorithm of the FP32 path, so the two ratios
    # below separate the algorithm's share of the speedup from precision's
    out64x = torch.empty(w.voxels, dtype=torch.float64, device="cuda")
    log("fp64 exact: warmup")
    api.field_eval(grid, r, params, FP64_EXACT, out=out64x, workspace=ws)
    torch.cuda.synchronize()
    ws.reset_stats()
    api._lib.load().qj_workspace_timing(ws.handle, 1)
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        api.field_eval(grid, r, params, FP64_EXACT, out=out64x, workspace=ws)
    torch.cuda.synchronize()
    t64x = (time.perf_counter() - t0) / args.fp64_steps
    k64x = _kernel_ms(api, ws)
    api._lib.load().qj_workspace_timing(ws.handle, 0)
    log("fp64 exact: %.2f ms/step" % (t64x * 1e3))
    f64x = out64x.cpu().numpy()
    del out64x

    # FP32 bulk + FP64 band
    for _ in range(args.warmup):
        api.field_eval(grid, r, params, FP32_BAND, args.band, out=out32, workspace=ws)
    torch.cuda.synchronize()
    ws.reset_stats()
    api._lib.load().qj_workspace_timing(ws.handle, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        api.field_eval(grid, r, params, FP32_BAND, args.band, out=out32, workspace=ws)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.steps
    st32 = ws.stats()
    k32 = _kernel_ms(api, ws)
    api._lib.load().qj_workspace_timing(ws.handle, 0)
    log("fp32+band: %.2f ms/step" % (t32 * 1e3))

    f64 = out64.cpu().numpy()
    f32 = out32.cpu().numpy()
    flips_vs_fp64 = int(np.count_nonzero((f64.astype(np.float32) > 0) != (f32.astype(np.float32) > 0)))
    fin = np.isfinite(f64) & np.isfinite(f32)
    maxdf = float(np.max(np.abs(f64[fin] - f32[fin]))) if fin.any() else 0.0

    value = nominal / t32 / 1e9
    value64 = nominal / t64 / 1e9
    iters_per_launch = st32["iterations"] / args.steps
    bulk_ms = k32["bulk_ms"] / max(1, k32["launches"])
    band_ms = k32["band_ms"] / max(1, k32["launches"])
    flops = iters_per_launch * basic_ops_per_voxel_iter(nt, nb)
    achieved = flops / (bulk_ms * 1e-3) / 1e12
    traffic = K_GRID32_TRAFFIC_BYTES if args.res == 512 else None
    roofline = {"bound": "valu", "compute_unit": "FP32 vector ALU (no matrix op applies; published FP32 peak)",
                "kernel": "k_grid32", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                "frac_note": "delivered reference-method work: executed voxel-iterations x the reference's %d basic "
                             "ops each (the polar pow's 23 per term), over the published peak; k_grid32 takes "
                             "exponents 1 and 2 as exact products and issues fewer" % basic_ops_per_voxel_iter(nt, nb),
                "valu_busy": K_GRID32_VALU_BUSY,
                "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE passes scaled by a 1 GiB "
                                "copy (profiles/k_grid32_counters_c2_512.md); algorithmic bytes %d" % (w.voxels * 8),
                "algorithmic_flops_per_launch": flops, "kernel_ms": round(bulk_ms, 4),
                "executed_voxel_iters_per_launch": iters_per_launch}
    fp64_ms = k64["bulk_ms"] / max(1, k64["launches"])
    fp64_flops = st64["iterations"] / args.fp64_steps * basic_ops_per_voxel_iter(nt, nb)
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        log("cpu baseline: %d slabs on %d threads" % (args.cpu_slabs, threads))
        cb = cpu_baseline(w, f64, f32, args.cpu_slabs, threads)
        cpu = {"value": round(cb["value"], 6), "unit": "Gvoxel-iter/s", "cores": threads, "kind": "port",
               "sample": "%d of %d z-slabs (%dx%d voxels each) of the same 512^3 workload, oracle C restatement "
                         "(FP64, OpenMP), %.1f s" % (cb["slabs"], w.res[2], w.res[0], w.res[1], cb["seconds"]),
               "parity_on_sample": {"sign_flips_fp64_gpu": cb["flips_fp64"],
                                    "sign_flips_fp32band_gpu": cb["flips_fp32band"],
                                    "max_rel_err_fp64_gpu": cb["max_rel_err_fp64"]}}
    c3 = None if args.no_c3 else c3_tiles(args, log)
    grads = None if args.no_grad else gradients(args, log)
    nlmc = None if args.no_nlmc else nonlinear_mesher(args, log)
    serving = serving_esc = None
    if args.serve > 0:
        serving = serve_stream(args, log, 0.0)
        if args.serve_escalate > 0:
            serving_esc = serve_stream(args, log, args.serve_escalate)
    line = {
        "metric": "Gvoxel-iter/s at 512^3 (FP32+FP64 band vs FP64) + p50/p99 request latency",
        "value": round(value, 4),
        "unit": "Gvoxel-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded roots, numpy default_rng(1002))",
        "config": {"workload": "c2_rational_32x32_%d: %d^3 grid on [-1.5,1.5]^3, 32 top + 32 bottom roots, "
                               "powers U{1,2}, maxIter 4, escape 2000, band |F|<%g" % (args.res, args.res, args.band),
                   "res": args.res, "n_top": nt, "n_bottom": nb, "max_iter": w.max_iter, "band": args.band},
        "fp64_plain": {"value": round(value64, 4), "ms_per_step": round(t64 * 1e3, 3), "kernel_ms": round(fp64_ms, 3),
                       "achieved_tflops": round(fp64_flops / (fp64_ms * 1e-3) / 1e12, 3),
                       "peak": PEAK_FP64_TFLOPS},
        "speedup_vs_fp64": round(value / value64, 3),
        "fp64_exact": {"value": round(nominal / t64x / 1e9, 4), "ms_per_step": round(t64x * 1e3, 3),
                       "kernel_ms": round(k64x["bulk_ms"] / max(1, k64x["launches"]), 3),
                       "what": "all voxels in FP64, exponents 1 and 2 as exact products (QJ_FP64_EXACT)",
                       "sign_flips_vs_fp64_gpu": int(np.count_nonzero((f64.astype(np.float32) > 0) !=
                                                                      (f64x.astype(np.float32) > 0)))},
        "speedup_vs_fp64_exact": round(value / (nominal / t64x / 1e9), 3),
        "band": {"voxels": st32["band_voxels"] // args.steps, "fraction": st32["band_voxels"] / args.steps / w.voxels,
                 "band_kernel_ms": round(band_ms, 4)},
        "parity": {"sign_flips_vs_fp64_gpu": flips_vs_fp64, "max_abs_dF_vs_fp64_gpu": maxdf},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "c3_1024_tiles": c3,
        "gradients": grads,
        "nonlinear_mesher": nlmc,
        "serving": serving,
        "serving_escalated": serving_esc,
        "p50_ms_256": (serving or {}).get("latency", {}).get("256^3", {}).get("p50_ms") if serving else None,
        "p99_ms_256": (serving or {}).get("latency", {}).get("256^3", {}).get("p99_ms") if serving else None,
        "p99_ms_512": (serving or {}).get("latency", {}).get("512^3", {}).get("p99_ms") if serving else None,
        "serving_refused": (serving or {}).get("refused") if serving else None,
    }
    print(json.dumps(line), flush=True)


def serve_stream(args, log, escalate):
    """Config 4: Poisson arrivals at serve_rate, grid edge in {64,128,256,512}, 8-256 roots
    (half top, half bottom, Gaussian sigma 0.7), maxIter 3-5, through the request loop, in the
    reference's default-build range (QJ_MAP_EXTENDED_RANGE, Real = long double). A submission
    is refused when more than serve_pending_s of work (in seconds of the loop's measured
    root-term rate) is queued ahead of it in shortest-first order; requests waiting longer
    than serve_aging_s are admitted first and take up to a quarter of each launch; every result
    carries its sampled FP64 cross-check. escalate > 0: a result whose check shows a sign flip
    or a max |dF| above it is computed again with every voxel in FP64 before it completes."""
    from quijibo_amd import server as S, workloads as W
    from quijibo_amd._lib import FP32_BAND, MAP_BLACK_HOLE, MAP_EXTENDED_RANGE
    reqs = W.config4_requests(n=args.serve, rate=args.serve_rate)
    rate = 2.5e12  # root-term iterations per second the loop sustains on this stream
    srv = S.Server(hbm_budget_bytes=int(args.serve_budget_gib * (1 << 30)), precision=FP32_BAND, band=args.band,
                   launch_work=args.serve_launch_work, inflight_launches=args.serve_inflight,
                   max_pending_work=args.serve_pending_s * rate, aging_s=args.serve_aging_s,
                   check_samples=args.serve_check, escalate_abs_dF=escalate)
    flags = MAP_BLACK_HOLE | MAP_EXTENDED_RANGE
    # warm the graphs with one small request
    srv.submit(reqs[0][1], flags=flags)
    srv.wait_all(1, timeout=60)
    log("serving: %d requests at %.0f req/s" % (len(reqs), args.serve_rate))
    done, ids, start, refused = S.replay(srv, reqs, log, flags=flags)
    end = max(c.done_time for c in done) if done else start
    srv.close()
    vox_iters = sum(ids[c.id].voxels * ids[c.id].max_iter for c in done if c.status == 0)
    lat = S.latency_summary(done, ids)
    ok = [c for c in done if c.status == 0]
    check = {"checked_cells": int(sum(c.checked for c in ok)),
             "results_checked": int(sum(1 for c in ok if c.checked > 0)),
             "sign_flips": int(sum(c.check_sign_flips for c in ok)),
             "class_mismatch": int(sum(c.check_class_mismatch for c in ok)),
             "max_abs_dF": max((c.check_max_abs_dF for c in ok), default=0.0),
             "max_rel_dF": max((c.check_max_rel_dF for c in ok), default=0.0),
             "samples_per_result": args.serve_check,
             "escalated_to_fp64": int(sum(c.escalated for c in ok)),
             "escalate_above_abs_dF": escalate}
    return {"requests": len(reqs), "offered_rate_req_s": args.serve_rate, "completed": len(ok),
            "refused": refused, "wall_s": round(end - start, 3),
            "throughput_req_s": round(len(ok) / max(1e-9, end - start), 2),
            "throughput_gvoxel_iter_s": round(vox_iters / max(1e-9, end - start) / 1e9, 3), "latency": lat,
            "error_statement": check,
            "config": "hbm budget %.0f GiB, 65536-voxel chunks, launches of <= 512 chunks and ~%s root-term "
                      "iterations, %s launches in flight, fp32+fp64 band, extended range, shortest-remaining-work "
                      "first, aging %.2f s (aged requests take up to a quarter of each launch), refusal when more "
                      "than %.1f s of work (%.2g root-term iterations) is queued ahead in shortest-first order"
                      % (args.serve_budget_gib, ("%.1e" % args.serve_launch_work) if args.serve_launch_work else "3e9",
                         args.serve_inflight or 2, args.serve_aging_s, args.serve_pending_s,
                         args.serve_pending_s * rate)}


def nonlinear_mesher(args, log):
    """The low-memory non-linear mesher on the headline roots at 512^3: slabs of 64 z-slices
    evaluated FP32+band, bisection vertices (up to 6 steps, FP64 map evaluations), triangles;
    device memory O(512^2 x 64) instead of the 1 GiB field."""
    import torch
    from quijibo_amd import api, workloads as W
    from quijibo_amd._lib import FP32_BAND, MAP_BLACK_HOLE
    w = W.headline(512)
    g = api.make_grid(w.res, w.center, w.lengths)
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    ws = api.Workspace()
    small = api.make_grid((64, 64, 64), w.center, w.lengths)
    api.nonlinear_mesh(small, r, p, 0.0, FP32_BAND, args.band, 64, workspace=ws)  # warm-up
    v, t = api.nonlinear_mesh(g, r, p, 0.0, FP32_BAND, args.band, 64, workspace=ws, device_output=True)
    del v, t  # warm-up at full size: buffers grown once, output blocks cached by torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v, t = api.nonlinear_mesh(g, r, p, 0.0, FP32_BAND, args.band, 64, workspace=ws, device_output=True)
    torch.cuda.synchronize()
    ms_dev = (time.perf_counter() - t0) * 1e3
    del v, t
    t0 = time.perf_counter()
    v, t = api.nonlinear_mesh(g, r, p, 0.0, FP32_BAND, args.band, 64, workspace=ws)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    log("nonlinear mesher 512^3: %.1f ms (mesh left in HBM), %.1f ms with the copy to host memory, %d vertices, "
        "%d triangles" % (ms_dev, ms, len(v), len(t)))
    return {"workload": "512^3 headline roots (32+32, powers U{1,2}, maxIter 4), slabs of 64 z-slices, FP32+band "
                        "slab values, bisection vertices",
            "ms": round(ms_dev, 1), "ms_with_host_copy": round(ms, 1),
            "note": "ms: the mesh left in device tensors; ms_with_host_copy: copied into pageable host arrays "
                    "as the reference's TRIANGLE_MESH holds it",
            "vertices": int(len(v)), "triangles": int(len(t))}


def gradients(args, log):
    """Config 5 (40k points: 20k on a superquadric, 20k at distance 0.2 outside; 64+64 roots,
    maxIter 3): forward field and the reverse-mode gradient over all 512 root components and
    128 powers, FP64 and FP32. The AP pass (one iteration, escape 20, skip-origin and extended
    range) with its score and exponent gradients on a 97^3 grid of the 32+32 headline roots."""
    import torch
    from quijibo_amd import api, grad as G, workloads as W
    from quijibo_amd._lib import FP32_BAND, FP64, MAP_BLACK_HOLE, MAP_ADAPTIVE, MAP_SKIP_ORIGIN
    n = 20000
    w = W.config5(n_surface=n, n_exterior=n, max_iter=3)
    pts = torch.from_numpy(np.concatenate([w.extra["surface"], w.extra["exterior"]])).cuda()
    p = api.make_params(w.max_iter, 2000.0, 1.0, 0.0, MAP_BLACK_HOLE)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    ws = api.Workspace()
    reps = 10
    out = {}
    for name, prec in (("fp64", FP64), ("fp32", FP32_BAND)):
        F = api.points_eval(pts, r, p, prec, workspace=ws)
        seed = G.superquadric_objective_seed(F, n)
        api.points_adjoint(pts, r, p, seed, prec, workspace=ws)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            F = api.points_eval(pts, r, p, prec, workspace=ws)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            api.points_adjoint(pts, r, p, seed, prec, workspace=ws)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out["c5_" + name] = {"forward_ms": round((t1 - t0) / reps * 1e3, 3),
                             "adjoint_ms": round((t2 - t1) / reps * 1e3, 3)}
    log("gradients: c5 %s" % out)
    h = W.headline(97)
    res = (97, 97, 97)
    g = api.make_grid(res, h.center, h.lengths)
    ax = [-h.lengths[i] / 2 + (np.arange(res[i]) + 0.5) * h.lengths[i] / res[i] for i in range(3)]
    Zc, Yc, Xc = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    weight = torch.from_numpy((1.0 / (np.abs(np.sqrt(Xc ** 2 + Yc ** 2 + Zc ** 2) - 0.8) + 0.05)).reshape(-1)).cuda()
    pa = api.make_params(1, 20.0, h.exp_scaling, 0.0, MAP_SKIP_ORIGIN | MAP_ADAPTIVE)
    ra = api.Rational(api.POLYNOMIAL_4D(h.top.roots, h.top.powers), api.POLYNOMIAL_4D(h.bottom.roots, h.bottom.powers))
    api.ap_score_gradients(g, ra, pa, g, weight, workspace=ws)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        res_ap = api.ap_score_gradients(g, ra, pa, g, weight, workspace=ws)
    torch.cuda.synchronize()
    out["ap_97"] = {"ms": round((time.perf_counter() - t0) / reps * 1e3, 3), "score": res_ap["score"],
                    "adaptive_cells": res_ap["adaptive_cells"]}
    out["workload"] = ("c5: 40000 points, 64+64 roots, maxIter 3, gradient over 512 root components + 128 powers; "
                       "ap_97: 97^3 grid, 32+32 roots, one iteration, escape 20, score + exponent gradients")
    return out


def mc_probe(args, log):
    """Counter probe for marching cubes: the 256^3 tile (1, 1, 0) of the 1024^3 torus workload
    (it holds part of the surface) evaluated FP32+band in the extended range, meshed twice (a warm-up and the run the
    counters are read from), then a plain 1 GiB device-to-device copy whose FETCH_SIZE /
    WRITE_SIZE readings scale the others."""
    import ctypes
    import torch
    from quijibo_amd import api, workloads as W
    from quijibo_amd._lib import FP32_BAND, MAP_BLACK_HOLE, MAP_EXTENDED_RANGE
    w = W.config3(1024, 256)
    c, tl = W.tile_grid(w, 1, 1, 0, 256)
    g = api.make_grid((256, 256, 256), c, tl)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    params = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE | MAP_EXTENDED_RANGE)
    ws = api.Workspace()
    fld = api.field_eval(g, r, params, FP32_BAND, args.band, workspace=ws)
    torch.cuda.synchronize()
    L = api._lib.load()
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    runs = []
    for i in range(2):
        log("mc probe: marching cubes %d" % i)
        t0 = time.perf_counter()
        api.check(L.qj_mc_run(ws.handle, api._ptr(fld), ctypes.byref(g), 0.0, ctypes.byref(nv), ctypes.byref(nt),
                              api._stream_ptr(None)))
        torch.cuda.synchronize()
        runs.append({"ms": round((time.perf_counter() - t0) * 1e3, 3), "vertices": nv.value, "triangles": nt.value})
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    log("mc probe: copy 1 GiB")
    dst.copy_(src)
    torch.cuda.synchronize()
    return {"tile": [1, 1, 0], "points": 256 ** 3, "runs": runs, "copy_bytes": 8 << 27}


def c3_tiles(args, log):
    """1024^3 in 64 tiles of 256^3 (128+128 torus roots, maxIter 3), FP32+band in the reference's
    default-build range (QJ_MAP_EXTENDED_RANGE, Real = long double: products of 256 roots stay
    finite), then the reference's clean-up before meshing (clampInfsToMean, setBorders(totalCells),
    toothOptimizer.cpp:250-254) and marching cubes of all 64 tiles in one call. Every tile's error
    statement comes from a sampled FP64 cross-check (4096 cells per tile, untimed pass); two tiles
    are re-evaluated in full with the FP64 kernel."""
    import torch
    from quijibo_amd import api, workloads as W
    from quijibo_amd._lib import FP32_BAND, FP64, MAP_BLACK_HOLE, MAP_EXTENDED_RANGE
    w = W.config3(1024, 256)
    tile = 256
    per = tile ** 3
    nt_axis = w.res[0] // tile
    tiles = [(tz, ty, tx) for tz in range(nt_axis) for ty in range(nt_axis) for tx in range(nt_axis)]
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    params = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE | MAP_EXTENDED_RANGE)
    ws = api.Workspace()
    out = torch.empty(len(tiles) * per, dtype=torch.float64, device="cuda")
    grids, centers, lengths = [], [], []
    for (tz, ty, tx) in tiles:
        c, tl = W.tile_grid(w, tz, ty, tx, tile)
        grids.append(api.make_grid((tile, tile, tile), c, tl))
        centers.append(c)
        lengths.append(tl)
    log("c3: %d tiles of %d^3, %d+%d roots" % (len(tiles), tile, w.top.n, w.bottom.n))
    api.field_eval(grids[0], r, params, FP32_BAND, args.band, out=out[:per], workspace=ws)  # warm-up
    torch.cuda.synchronize()
    ws.reset_stats()
    t0 = time.perf_counter()
    for k, g in enumerate(grids):
        api.field_eval(g, r, params, FP32_BAND, args.band, out=out[k * per:(k + 1) * per], workspace=ws)
    torch.cuda.synchronize()
    t_eval = time.perf_counter() - t0
    st = ws.stats()
    log("c3: field %.1f ms" % (t_eval * 1e3))
    # error statement of every tile (untimed): sampled FP64 cross-check
    chk = api.Workspace()
    chk.check(4096)
    scratch = torch.empty(per, dtype=torch.float64, device="cuda")
    worst = {"max_abs_dF": 0.0, "max_rel_dF": 0.0, "sign_flips": 0, "class_mismatch": 0, "checked": 0}
    for k, g in enumerate(grids):
        chk.reset_stats()
        api.field_eval(g, r, params, FP32_BAND, args.band, out=scratch, workspace=chk)
        s_k = chk.stats()
        worst["max_abs_dF"] = max(worst["max_abs_dF"], s_k["check_max_abs_dF"])
        worst["max_rel_dF"] = max(worst["max_rel_dF"], s_k["check_max_rel_dF"])
        worst["sign_flips"] += s_k["check_sign_flips"]
        worst["class_mismatch"] += s_k["check_class_mismatch"]
        worst["checked"] += s_k["checked"]
    nonfinite = int((~torch.isfinite(out)).sum().item())
    log("c3: error statement over %d tiles %s, %d non-finite cells" % (len(tiles), worst, nonfinite))
    # clean-up before meshing, then all tiles in one marching-cubes call
    t0 = time.perf_counter()
    for k in range(len(tiles)):
        F = api.FIELD_3D(tile, tile, tile, center=centers[k], lengths=lengths[k], data=out[k * per:(k + 1) * per])
        F.clampInfsToMean(workspace=ws, return_mean=False)  # no host wait per tile
        F.setBorders(F.totalCells())
    torch.cuda.synchronize()
    t_clamp = time.perf_counter() - t0
    fields = [out[k * per:(k + 1) * per] for k in range(len(tiles))]
    # untimed call over all tiles: sizes the workspace's buffers and gives output tensors of
    # the meshes' size, which the timed call writes into directly (no allocation, no copy)
    vb, tb, _ = api.marching_cubes_tiles(fields, grids, workspace=ws)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v, t, counts = api.marching_cubes_tiles(fields, grids, workspace=ws, out=(vb, tb))
    torch.cuda.synchronize()
    t_mc = time.perf_counter() - t0
    tot_v, tot_t = int(v.shape[0]), int(t.shape[0])
    mc_bytes = len(tiles) * per * 8 + tot_v * 24 + tot_t * 12
    nan_vertices = int((~torch.isfinite(v)).any(dim=1).sum().item())
    log("c3: marching cubes %.1f ms, %d vertices, %d triangles, %d non-finite vertices" % (t_mc * 1e3, tot_v,
                                                                                          tot_t, nan_vertices))
    del v, t, vb, tb
    # full FP64 parity on two tiles: one holding the torus surface, one elsewhere (fields re-evaluated,
    # since the ones above were clamped)
    check = [tiles.index((1, 1, 0)), tiles.index((2, 2, 2))]
    parity = []
    f64 = torch.empty(per, dtype=torch.float64, device="cuda")
    for k in check:
        g = grids[k]
        api.field_eval(g, r, params, FP32_BAND, args.band, out=scratch, workspace=ws)
        host = scratch.cpu().numpy()
        api.field_eval(g, r, params, FP64, out=f64, workspace=ws)
        h64 = f64.cpu().numpy()
        flips = int(np.count_nonzero((h64.astype(np.float32) > 0) != (host.astype(np.float32) > 0)))
        fin = np.isfinite(h64) & np.isfinite(host)
        parity.append({"tile": list(tiles[k]), "sign_flips_vs_fp64_gpu": flips,
                       "max_abs_dF_vs_fp64_gpu": float(np.max(np.abs(h64[fin] - host[fin]))) if fin.any() else 0.0})
        log("c3: tile %s parity %s" % (tiles[k], parity[-1]))
    del out, f64, scratch, fields
    torch.cuda.empty_cache()
    vox_iters = len(tiles) * per * w.max_iter
    return {"workload": "1024^3 on [-1.5,1.5]^3 as 64 tiles of 256^3, 128+128 torus roots (R=1, r=0.35), powers 1, "
                        "maxIter 3, FP32+FP64 band, extended range",
            "gvoxel_iter_s": round(vox_iters / t_eval / 1e9, 4), "field_ms": round(t_eval * 1e3, 2),
            "band_voxels": st["band_voxels"], "nonfinite_cells": nonfinite, "error_statement": worst,
            "clamp_borders_ms": round(t_clamp * 1e3, 2), "mc_ms_all_tiles": round(t_mc * 1e3, 2),
            "mc_host_syncs": 1, "vertices": tot_v, "triangles": tot_t, "nonfinite_vertices": nan_vertices,
            "mc_roofline": {"bound": "hbm", "algorithmic_bytes": mc_bytes, "achieved": round(mc_bytes / t_mc / 1e9, 1),
                            "peak": 8000.0, "unit": "GB/s", "frac": round(mc_bytes / t_mc / 8e12, 4),
                            "note": "the fields read once (8 B per point) and the meshes written once (24 B per "
                                    "vertex, 12 B per triangle) over the call's time; the kernels are "
                                    "latency-bound (profiles/marching_cubes_c3_tile.md)"},
            "mc_note": "one qj_mc_run_tiles_into call over the 64 clamped tiles, the meshes written straight into "
                       "output tensors sized by an untimed call over the same tiles",
            "parity": parity}


def _kernel_ms(api, ws):
    import ctypes
    b = ctypes.c_double()
    d = ctypes.c_double()
    n = ctypes.c_int64()
    api.check(api._lib.load().qj_workspace_kernel_ms(ws.handle, ctypes.byref(b), ctypes.byref(d), ctypes.byref(n)))
    return {"bulk_ms": b.value, "band_ms": d.value, "launches": n.value}


if __name__ == "__main__":
    main()
