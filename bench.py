#!/usr/bin/env python3
"""Benchmark: DMC extraction throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload schwarz_1024]

A step is one full dualMC of the workload's field (already resident in
HBM): classification, compaction, vertices, quadrilaterals, 3X3Y (colour and
isolated), 3333, halfedges and triangles; the mesh stays in HBM.  `value` is
Gvoxels/s = nx*ny*nz / (time per step).

The roofline object is the cell stage as the survey's measurement section
defines it: algorithmic bytes B = 4 N + 24 V + 17 Q (the field read once,
FP32 position and normal per vertex, an int4 quadrilateral and its colour
byte) over the summed device time of the cell-stage kernels (k_sweep,
k_rows, k_vertices, k_vertices_slow, k_quads and the scans between them),
timed with HIP events on the engine's stream over the timed steps.
`kernels` lists the main kernels with their own byte models and fractions.
The CPU baseline is the reference cpp_dmc compiled from its sources
(oracle/_ref) -- or the C restatement if that is absent -- on one host
core: dualMC alone and estimateGradient + flip_gradient + dualMC, on a
512^3 corner block of the same field, and the whole 128^3 genus-2 field
(median of 5).  `--cpu-full` runs the survey's complete CPU plan instead
(medians of 5 on the 128^3 and 512^3 fields, one full 1024^3 run).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Bytes per extraction the cell-stage kernels move between L2 and the fabric
# (HBM and Infinity Cache), from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# scaled by the copy-kernel calibration
# (profiles/dmc_schwarz_1024_kernel_traffic.md): k_sweep 5.74 GB, k_rows
# 1.14 GB, k_vertices (its 3-wave form) 11.46 GB, k_quads 8.27 GB, the cell
# scans 0.47 GB.
TRAFFIC = {"schwarz_1024": 27_100_000_000}
CELL_KERNELS = ("k_sweep", "k_scan_cells", "k_vol_abase", "k_rows", "k_vol_bounds", "k_vertices", "k_vertices_slow",
                "k_quads")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _cpu_run(sub, g, iso):
    """One CPU run on a float field: (V, Q, gradient s, dualMC s, kind)."""
    import oracle
    if os.path.exists(oracle.REF_SO):
        nv, nq, tg, td = oracle.Reference().run_float(g, sub, iso, True, legs=True)
        return nv, nq, tg, td, "reference"
    orc = oracle.Oracle()
    s = sub.astype(np.float64).ravel()
    t = time.perf_counter()
    grad = orc.gradient(g, s)
    tg = time.perf_counter() - t
    t = time.perf_counter()
    m = orc.dualmc(g, s, grad, iso, True)
    td = time.perf_counter() - t
    return len(m["v"]), len(m["quads"]), tg, td, "port"


def _ogrid(field, grid, shape=None):
    import oracle
    nz, ny, nx = shape or field.shape
    return oracle.Grid(nx, ny, nz, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)


def cpu_case(field, grid, iso, reps, label):
    """`reps` CPU runs of the whole field (median), dualMC alone and with
    the gradient."""
    runs = []
    for r in range(reps):
        runs.append(_cpu_run(field, _ogrid(field, grid), iso))
        log(f"[cpu] {label}: run {r + 1}/{reps}: gradient {runs[-1][2]:.3f} s, dualMC {runs[-1][3]:.3f} s")
    tg = float(np.median([r[2] for r in runs]))
    td = float(np.median([r[3] for r in runs]))
    n = field.size
    log(f"[cpu] {label}: gradient {tg:.3f} s, dualMC {td:.3f} s (median of {reps})")
    return {"case": label, "points": int(n), "runs": reps, "kind": runs[0][4], "V": runs[0][0], "Q": runs[0][1],
            "gradient_s": round(tg, 4), "dualmc_s": round(td, 4),
            "dualmc_gvoxels_per_s": round(n / td / 1e9, 6),
            "gradient_plus_dualmc_gvoxels_per_s": round(n / (tg + td) / 1e9, 6)}


def cpu_baseline(field, grid, iso, block):
    """Reference cpp_dmc on one core: the block^3 corner of the field
    (estimateGradient + flip_gradient and dualMC with sFlag timed apart),
    and the whole 128^3 genus-2 field, median of 5."""
    from tmc_amd import workloads
    sub = np.ascontiguousarray(field[:block, :block, :block])
    nv, nq, tg, td, kind = _cpu_run(sub, _ogrid(sub, grid), iso)
    n = block ** 3
    f1, g1, i1 = workloads.genus2(128)
    c1 = cpu_case(f1, g1, i1, 5, "genus2_128")
    return {"value": round(n / (tg + td) / 1e9, 6), "unit": "Gvoxels/s", "cores": 1, "kind": kind,
            "sample": f"{block}^3 corner block of the same field (k,j,i < {block}); UniformGrid from the float "
                      f"field, estimateGradient + flip_gradient {tg:.2f} s + dualMC(sFlag=true) {td:.2f} s, "
                      f"V={nv} Q={nq}; host CPU: {cpu_model()} ({os.cpu_count()} logical)",
            "dualmc_only_gvoxels_per_s": round(n / td / 1e9, 6),
            "gradient_plus_dualmc_gvoxels_per_s": round(n / (tg + td) / 1e9, 6),
            "genus2_128": c1,
            "full_plan": "python bench.py --cpu-full (medians of 5 at 128^3 and 512^3, one 1024^3 run): "
                         "profiles/cpu_baseline_full.md"}


def cpu_full():
    """The survey's CPU plan on this host: 128^3 genus-2 and 512^3 blobs
    (median of 5 each) and the 1024^3 Schwarz-P field (one run), dualMC
    alone and with the gradient, one core."""
    from tmc_amd import workloads
    mem = next((l.split()[1] for l in open("/proc/meminfo") if l.startswith("MemTotal")), "?")
    out = {"host": f"{cpu_model()} ({os.cpu_count()} logical), MemTotal {mem} kB", "cores": 1, "cases": []}
    for name, reps in (("genus2_128", 5), ("blobs_512", 5), ("schwarz_1024", 1)):
        f, g, iso = workloads.CONFIGS[name]()
        out["cases"].append(cpu_case(f, g, iso, reps, name))
        del f
        print(json.dumps(out["cases"][-1]), flush=True)
    return out


def pack_device(eng, fields):
    """All fields in one device allocation at 256-B aligned offsets; returns
    the buffer and each field's device address."""
    offs, total = [], 0
    for f in fields:
        offs.append(total)
        total += (f.nbytes + 255) // 256 * 256
    host = np.zeros(total, np.uint8)
    for f, o in zip(fields, offs):
        host[o:o + f.nbytes] = np.frombuffer(f.tobytes(), np.uint8)
    buf = eng.alloc(total).upload(host)
    return buf, [buf.ptr.value + o for o in offs]


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else float("nan")


def bench_batch(count=4096, seed=4, steps=3, plain_sample=8):
    """Batched workload: `count` small volumes (16-96 per axis, log-uniform,
    mixed generators), fields resident in HBM.  Batched = one
    dmc_extract_batch over all of them; plain = one dmc_extract per volume
    (every `plain_sample`-th volume, extrapolated by volumes/s)."""
    import tmc_amd
    from tmc_amd import workloads
    t0 = time.perf_counter()
    specs = workloads.batch_specs(count, seed=seed)
    vols = workloads.make_many(specs)
    log(f"[batch] {count} volumes generated in {time.perf_counter() - t0:.1f} s")
    nvox = sum(f.size for f, _, _ in vols)
    eng = tmc_amd.Engine(0)
    buf, ptrs = pack_device(eng, [f for f, _, _ in vols])
    dv = [(p, iso, g) for p, (f, g, iso) in zip(ptrs, vols)]
    flags = tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES | tmc_amd.FIELD_ON_DEVICE
    eng.extract_batch(dv, flags=flags)
    # launch shape picked by the autotuner on this batch (untimed)
    t = time.perf_counter()
    tune = eng.autotune(dv, flags=flags, reps=2)
    tune["tune_s"] = round(time.perf_counter() - t, 2)
    tune["best"]["ms"] = round(tune["best"]["ms"], 3)
    tune["trials"] = [dict(x, ms=round(x["ms"], 3)) for x in tune["trials"]]
    log(f"[batch] autotune: {tune['best']} in {tune['tune_s']} s")
    # two untimed passes under the tuned shape (the second is captured into a
    # HIP graph), then the timed ones replay it
    for _ in range(2):
        eng.extract_batch(dv, flags=flags)
    t = time.perf_counter()
    for _ in range(steps):
        ms = eng.extract_batch(dv, flags=flags)
    batched_s = (time.perf_counter() - t) / steps
    batch_pass = ms[0].pass_mode
    nq = sum(m.nq for m in ms)
    log(f"[batch] batched: {batched_s * 1e3:.1f} ms per {count} volumes")
    sample = dv[::plain_sample]
    for p, iso, g in sample[:8]:
        eng.extract(p, iso, g, flags=flags)
    lat = []
    for p, iso, g in sample:
        t1 = time.perf_counter()
        eng.extract(p, iso, g, flags=flags)
        lat.append(time.perf_counter() - t1)
    plain_vps = len(sample) / sum(lat)
    log(f"[batch] plain: {len(sample)} volumes, {np.mean(lat) * 1e3:.2f} ms each")
    buf.free()
    eng.close()
    return {"workload": f"batch_{count}", "volumes": count, "voxels": int(nvox), "quads": int(nq),
            "spec": "axes round(exp(U[ln16,ln96])), generator U{genus2, schwarz_p, blobs}, iso U[-0.2,0.2] / "
                    "U[0.3,0.7] (blobs), seed 4, fields in HBM",
            "batched": {"volumes_per_s": round(count / batched_s, 1), "ms_per_batch": round(batched_s * 1e3, 3),
                        "pass": "HIP graph replay" if batch_pass == 2 else "launched kernel by kernel",
                        "gvoxels_per_s": round(nvox / batched_s / 1e9, 3)},
            "plain": {"volumes_per_s": round(plain_vps, 1), "sampled": len(sample),
                      "p50_ms": round(pct(lat, 50) * 1e3, 3), "p99_ms": round(pct(lat, 99) * 1e3, 3)},
            "speedup": round(count / batched_s / plain_vps, 2),
            "launch": dict(tune, note="both legs run with the tuned shape; results do not depend on it")}


def serve_pool(seed=5):
    """Request pool with the serving mix exactly: 14 volumes U{32..128}^3,
    5 of 256^3, 1 of 512^3 (70/25/5 % under uniform draws)."""
    from tmc_amd import workloads
    specs = []
    rng_small = [s for s in workloads.serve_specs(400, seed=seed) if s["shape"][0] < 256]
    specs += rng_small[:14]
    base = workloads.serve_specs(200, seed=seed + 1)
    specs += [dict(s, shape=(256, 256, 256)) for s in base[:5]]
    specs += [dict(base[5], shape=(512, 512, 512))]
    return specs


def serve_run(srv, pool, rate, n, seed, label, classes=None):
    """Poisson arrivals at `rate` per second (inf: all at once) drawing
    uniformly from `pool`; returns latencies and throughput (and p50/p99 per
    class when `classes` gives each pool entry's class)."""
    import queue
    import threading
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(pool), size=n)
    gaps = np.zeros(n) if not np.isfinite(rate) else rng.exponential(1.0 / rate, size=n)
    arrive = np.cumsum(gaps)
    q = queue.Queue()
    lat, batch_sizes, done_t, kind = [], [], [], []

    def collect():
        for _ in range(n):
            i, t = q.get()
            info = srv.wait(t)
            lat.append(info["latency"])
            batch_sizes.append(info["batch_size"])
            done_t.append(info["t_done"])
            kind.append(classes[pick[i]] if classes else None)
            srv.release(t)

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for i in range(n):
        wait = arrive[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        p, iso, g = pool[pick[i]]
        q.put((i, srv.submit(p, iso, g)))
    th.join()
    span = time.perf_counter() - t0
    out = {"rate_offered": None if not np.isfinite(rate) else round(rate, 1), "requests": n,
           "volumes_per_s": round(n / span, 1), "p50_ms": round(pct(lat, 50) * 1e3, 3),
           "p99_ms": round(pct(lat, 99) * 1e3, 3), "mean_batch": round(float(np.mean(batch_sizes)), 2)}
    if classes:
        per = {}
        for c in sorted(set(classes)):
            xs = [l for l, k in zip(lat, kind) if k == c]
            if xs:
                per[c] = {"n": len(xs), "p50_ms": round(pct(xs, 50) * 1e3, 3), "p99_ms": round(pct(xs, 99) * 1e3, 3)}
        out["per_class"] = per
    log(f"[serve] {label}: {out}")
    return out


def _alone_ms(eng, pool, flags, reps=3):
    """Median device-resident one-at-a-time extraction time of each pool
    entry (ms)."""
    out = []
    for p, iso, g in pool:
        eng.extract(p, iso, g, flags=flags)
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            eng.extract(p, iso, g, flags=flags)
            ts.append(time.perf_counter() - t)
        out.append(float(np.median(ts)) * 1e3)
    return out


def bench_serve(seed=5, secs=4.0, batch_points=32 << 20, workers=2, pool_kind="mix"):
    """Served workload: Poisson arrivals drawing uniformly from a pool of
    device-resident fields.  pool_kind "mix": the serving mix (70%
    U{32..128}^3, 25% 256^3, 5% 512^3); "small": 256 volumes of the batched
    workload (16-96 per axis).  One-at-a-time (max_batch 1, one worker) runs
    at 70% of its capacity and sets the latency bar (its p99).  The batched
    server (passes of at most `batch_points` grid points, `workers` passes
    running concurrently) is then run at increasing multiples of that rate;
    reported is the highest rate whose p99 stays at or below the bar, with
    its p50/p99 and volumes/s, per size class too.  `bound` is what the
    pool's one-at-a-time extraction times allow: a batched server cannot
    beat the GPU time of its large volumes, which batching does not
    shorten."""
    import tmc_amd
    from tmc_amd import workloads
    t0 = time.perf_counter()
    if pool_kind == "mix":
        specs = serve_pool(seed)
        budget = 96 << 30
    else:
        specs = workloads.batch_specs(256, seed=seed + 100)
        budget = 96 << 30
    vols = workloads.make_many(specs)
    classes = [("small" if max(s["shape"]) < 256 else f"{max(s['shape'])}^3") for s in specs]
    log(f"[serve] {pool_kind} pool of {len(vols)} volumes generated in {time.perf_counter() - t0:.1f} s")
    eng = tmc_amd.Engine(0)
    buf, ptrs = pack_device(eng, [f for f, _, _ in vols])
    pool = [(p, iso, g) for p, (f, g, iso) in zip(ptrs, vols)]
    flags = tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES | tmc_amd.FIELD_ON_DEVICE
    alone = _alone_ms(eng, pool, flags)
    per_class_alone = {c: round(float(np.median([t for t, k in zip(alone, classes) if k == c])), 3)
                       for c in sorted(set(classes))}
    big = sum(t for t, k in zip(alone, classes) if k != "small") / len(pool)
    mean_alone = float(np.mean(alone))
    bound = {"mean_alone_ms": round(mean_alone, 3), "alone_ms_per_class": per_class_alone,
             "large_gpu_ms_per_request": round(big, 3),
             "capacity_ratio_max": round(mean_alone / big, 2) if big > 0 else None,
             "rate_ratio_at_equal_p99_max": round(mean_alone / big / 0.7, 2) if big > 0 else None,
             "note": "a batched server spends at least the large volumes' one-at-a-time GPU time per request; "
                     "the plain server runs at 70% of its capacity"}
    plain_srv = tmc_amd.Server(0, max_batch=1, flags=flags, budget_bytes=budget)
    serve_run(plain_srv, pool, float("inf"), 2 * len(pool) if pool_kind == "mix" else 64, seed, "plain warm-up")
    cap_plain = serve_run(plain_srv, pool, float("inf"), 120 if pool_kind == "mix" else 400, seed + 1,
                          "plain capacity")["volumes_per_s"]
    rate = 0.7 * cap_plain
    plain = serve_run(plain_srv, pool, rate, max(100, int(rate * secs)), seed + 2, "plain at 70% capacity", classes)
    plain_srv.close()
    bsrv = tmc_amd.Server(0, flags=flags, budget_bytes=budget, max_batch_points=batch_points, workers=workers)
    serve_run(bsrv, pool, float("inf"), 4 * len(pool) if pool_kind == "mix" else 512, seed, "batched warm-up")
    cap_b = serve_run(bsrv, pool, float("inf"), 1200 if pool_kind == "mix" else 8192, seed + 3,
                      "batched capacity")["volumes_per_s"]
    best = None
    tried = []
    for mult in (1.1, 1.25, 1.5, 1.75, 2, 2.5, 3, 3.5, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32):
        r = mult * rate
        if r > 1.05 * cap_b:
            break
        res = serve_run(bsrv, pool, r, max(200, int(r * secs)), seed + 4 + int(4 * mult), f"batched at {mult}x",
                        classes)
        tried.append(dict(res, multiple=mult))
        if res["p99_ms"] <= plain["p99_ms"] and res["volumes_per_s"] >= 0.95 * r:
            best = dict(res, multiple=mult)
        else:
            break
    bsrv.close()
    buf.free()
    eng.close()
    mixdesc = ("70% U{32..128}^3, 25% 256^3, 5% 512^3 (pool of 20, uniform draws)" if pool_kind == "mix" else
               "256 volumes of the batched workload (axes round(exp(U[ln16,ln96])), mixed generators), uniform draws")
    out = {"workload": "serve_mix" if pool_kind == "mix" else "serve_small",
           "mix": mixdesc + ", Poisson arrivals, fields in HBM, budget 96 GiB",
           "plain": dict(plain, capacity_volumes_per_s=cap_plain),
           "batched_capacity_volumes_per_s": cap_b, "batch_points": batch_points, "workers": workers,
           "batched_at_equal_p99": best, "batched_runs": tried, "bound": bound}
    if best:
        out["rate_ratio_at_equal_p99"] = round(best["volumes_per_s"] / plain["volumes_per_s"], 2)
    return out


def bench_tmc(eng, buf, field, grid, iso, steps, warmup, cpu_block=None):
    """Topologically correct MC (tmc_extract) of the same field, resident in
    HBM: ms per step and Gvoxels/s, per-stage device times, the mesh counts.
    With cpu_block, the CPU baseline is the reference's t_mc compiled from
    mc/MarchingCubes.cpp (oracle/_ref) -- or the C restatement if that is
    absent -- on the cpu_block^3 corner block of the field, one core."""
    for _ in range(warmup):
        eng.tmc(buf, iso, grid, shape=field.shape, download=False)
    eng.synchronize()
    stage = {}
    t = time.perf_counter()
    for _ in range(steps):
        r = eng.tmc(buf, iso, grid, shape=field.shape, download=False, timing=True)
        for k, v in eng.stage_times().items():
            stage[k] = stage.get(k, 0.0) + v
    eng.synchronize()
    ms = (time.perf_counter() - t) / steps * 1e3
    # algorithmic bytes: the field read once (4 B per point) and the mesh
    # written once (48 B per vertex: FP64 position and normal; 12 B per triangle)
    alg = 4 * field.size + 48 * r["nv"] + 12 * r["nt"]
    out = {"algorithm": "topologically correct MC (mc/MarchingCubes.cpp t_mc)", "steps": steps,
           "ms_per_step": round(ms, 4), "gvoxels_per_s": round(field.size / (ms * 1e-3) / 1e9, 3),
           "algorithmic_bytes": alg, "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "mtris_per_s": round(r["nt"] / (ms * 1e-3) / 1e6, 2),
           "stage_ms": {k: round(v / steps, 4) for k, v in stage.items()},
           "V": r["nv"], "T": r["nt"], "active_cells": r["active_cells"], "ambiguous_cells": r["ambiguous_cells"],
           "face_failures": r["face_failures"]}
    log(f"[bench] tmc: {ms:.3f} ms/step")
    if cpu_block:
        import oracle
        b = min(cpu_block, *field.shape)
        sub = np.ascontiguousarray(field[:b, :b, :b], dtype=np.float64)
        g = oracle.Grid(b, b, b, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)
        if os.path.exists(oracle.TMC_REF_SO):
            kind, run = "reference", oracle.TmcReference().tmc
        else:
            kind, run = "port", oracle.Oracle().tmc
        t = time.perf_counter()
        res = run(g, sub, iso)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(sub.size / dt / 1e9, 6), "unit": "Gvoxels/s", "cores": 1, "kind": kind,
                               "sample": f"{b}^3 corner block of the same field; t_mc with gradient(), "
                                         f"{dt:.2f} s, V={len(res['v'])} T={len(res['tris'])}; host CPU: "
                                         f"{cpu_model()}"}
    return out


def calibrate_copy():
    """A plain copy kernel of known size: 1 GiB read + 1 GiB written per
    launch.  Run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, the ratio of
    the known bytes to the counter reading scales the counters of the DMC
    kernels.  Uses torch only (not the DMC library) in its own process."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3}), flush=True)


def bench_dropin(eng, field, grid, iso, reps=2):
    """The reference call surface: dualMC(i0, ugrid, v, n, tris, quads,
    sFlag) hands over a host grid and returns host arrays.  Timed: the
    field's upload, the extraction and the download of positions, normals,
    quads and triangles into host memory (the C++ drop-in then widens these
    into std::vector<Vector>)."""
    import tmc_amd
    flags = tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES
    eng.extract(field, iso, grid, flags=flags)
    eng.download(halfedges=False)
    t = time.perf_counter()
    for _ in range(reps):
        eng.extract(field, iso, grid, flags=flags)
        d = eng.download(halfedges=False)
    ms = (time.perf_counter() - t) / reps * 1e3
    mb = (d["v"].nbytes + d["n"].nbytes + d["quads"].nbytes + d["tris"].nbytes) / 1e6
    log(f"[bench] drop-in call with host buffers: {ms:.1f} ms")
    return {"ms_per_call": round(ms, 2), "gvoxels_per_s": round(field.size / (ms * 1e-3) / 1e9, 3),
            "field_upload_mb": round(field.nbytes / 1e6, 1), "mesh_download_mb": round(mb, 1),
            "what": "host field in, host v/n/quads/tris out (PCIe both ways), as dualMC's callers see it"}


def bench_small(reps=50):
    """One small volume at a time: the genus-2 field at 128^3 (resident in
    HBM), extracted again and again, as a served request of one shape is.
    Wall clock per call, from the call to the mesh's counts on the host (the
    pass's one wait); from the third call on, the pass replays a HIP
    graph."""
    import tmc_amd
    from tmc_amd import workloads
    f, g, iso = workloads.genus2(128)
    eng = tmc_amd.Engine(0)
    try:
        buf = eng.alloc(f.nbytes).upload(f)
        flags = tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES | tmc_amd.FIELD_ON_DEVICE
        for _ in range(4):
            m = eng.extract(buf.ptr.value, iso, g, flags=flags)
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            m = eng.extract(buf.ptr.value, iso, g, flags=flags)
            ts.append((time.perf_counter() - t) * 1e3)
        buf.free()
    finally:
        eng.close()
    out = {"workload": "genus2_128", "points": int(f.size), "calls": reps, "p50_ms": round(pct(ts, 50), 4),
           "p99_ms": round(pct(ts, 99), 4), "min_ms": round(min(ts), 4),
           "gvoxels_per_s_at_p50": round(f.size / (pct(ts, 50) * 1e-3) / 1e9, 3),
           "pass": {0: "launched kernel by kernel", 1: "captured", 2: "HIP graph replay"}[m.pass_mode],
           "host_syncs": m.host_syncs, "V": m.nv, "Q": m.nq}
    log(f"[bench] small volume {out}")
    return out


def bench_diagnostics(eng, reps=3):
    """Time the mesh diagnostics on the last mesh (wall clock per call,
    results copied to the host included): unique edges, valence
    distributions of the quad and triangle meshes, element quality."""
    calls = {"edges": lambda: eng.edges(),
             "valence_quads": lambda: eng.valence_histogram(),
             "valence_tris": lambda: eng.valence_histogram(triangles=True),
             "quality_quads": lambda: eng.element_quality(),
             "quality_tris": lambda: eng.element_quality(triangles=True)}
    out = {}
    for name, fn in calls.items():
        fn()
        t = time.perf_counter()
        for _ in range(reps):
            r = fn()
        out[name + "_ms"] = round((time.perf_counter() - t) / reps * 1e3, 3)
        if name == "edges":
            out["edges"] = int(len(r))
        if name.startswith("valence"):
            out[name + "_hist"] = [int(x) for x in r[0]]
    log(f"[bench] diagnostics {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="schwarz_1024", choices=["schwarz_1024", "blobs_512", "genus2_128"])
    ap.add_argument("--cpu-block", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (counter calibration)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched (4,096 volumes) leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served (Poisson mix) leg")
    ap.add_argument("--no-diagnostics", action="store_true", help="skip the mesh diagnostics timings")
    ap.add_argument("--no-dropin", action="store_true", help="skip the host-buffer drop-in call timing")
    ap.add_argument("--no-small", action="store_true", help="skip the 128^3 single-volume latency leg")
    ap.add_argument("--only", choices=["batch", "serve"], help="run only the batched or served leg")
    ap.add_argument("--tmc-only", action="store_true",
                    help="run only the topologically correct MC leg on the workload (for profiling)")
    ap.add_argument("--no-tmc", action="store_true", help="skip the topologically correct MC leg")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only run the full CPU plan (128^3 and 512^3 median of 5, one 1024^3 run; host only)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.cpu_full:
        print(json.dumps(cpu_full()), flush=True)
        return
    if args.only == "batch":
        print(json.dumps(bench_batch()), flush=True)
        return
    if args.only == "serve":
        print(json.dumps({"served": bench_serve(), "served_small": bench_serve(pool_kind="small")}), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}, WORLD_SIZE {world}); the multi-GPU "
              "sharding of this path is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    import tmc_amd
    from tmc_amd import workloads

    log(f"[bench] generating {args.workload}")
    t0 = time.perf_counter()
    if args.workload == "schwarz_1024":
        field, grid, iso = workloads.schwarz(1024, seed=3)
        desc = {"workload": "schwarz_p_1024", "field": "cos x + cos y + cos z + N(0,0.01^2), seed 3",
                "domain": "[-8pi,8pi]^3 node-aligned", "grid": [1024, 1024, 1024], "iso": 0.3}
    elif args.workload == "blobs_512":
        field, grid, iso = workloads.blobs(512, seed=2)
        desc = {"workload": "blobs_512", "field": "64 Gaussian blobs + N(0,0.05^2), seed 2",
                "domain": "[0,1]^3 cell-centred", "grid": [512, 512, 512], "iso": 0.5}
    else:
        field, grid, iso = workloads.genus2(128)
        desc = {"workload": "genus2_128", "field": "genus-2 implicit", "domain": "[-1.5,1.5]^3 cell-centred",
                "grid": [128, 128, 128], "iso": 0.0}
    log(f"[bench] field ready in {time.perf_counter() - t0:.1f} s")
    nvox = field.size

    eng = tmc_amd.Engine(0)
    buf = eng.alloc(field.nbytes).upload(field)
    if args.tmc_only:
        print(json.dumps(dict(desc, tmc=bench_tmc(eng, buf, field, grid, iso, args.steps, args.warmup))), flush=True)
        buf.free()
        eng.close()
        return
    flags = tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES
    for w in range(args.warmup):
        eng.extract(buf, iso, grid, flags=flags, shape=field.shape)
        log(f"[bench] warmup {w + 1}/{args.warmup}")
    eng.synchronize()
    # timed region: the steps as a caller runs them (after the first passes
    # of this shape, each step replays the captured pass)
    t = time.perf_counter()
    for s in range(args.steps):
        m = eng.extract(buf, iso, grid, flags=flags, shape=field.shape)
    eng.synchronize()
    dt = time.perf_counter() - t
    ms = dt / args.steps * 1e3
    pass_mode = m.pass_mode
    # the same steps launched kernel by kernel with HIP events around every
    # kernel on the engine's stream: stage and kernel device times
    stage_sum = {}
    t = time.perf_counter()
    for s in range(args.steps):
        m = eng.extract(buf, iso, grid, flags=flags | tmc_amd.STAGE_TIMING, shape=field.shape)
        for k, v in eng.stage_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    eng.synchronize()
    ms_timed = (time.perf_counter() - t) / args.steps * 1e3
    stages = {k: round(v / args.steps, 4) for k, v in stage_sum.items() if not k.startswith("k_")}
    kms = {k: v / args.steps for k, v in stage_sum.items() if k.startswith("k_")}
    log(f"[bench] {args.steps} steps, {ms:.3f} ms/step; stages {stages}")
    summ = m.summary()
    V, Q, A = summ["nv_cells"], summ["nq_cells"], summ["active_cells"]
    Vf, Qf = summ["nv"], summ["nq"]
    # cell stage: 4 N + 24 V + 17 Q over the summed device time of its kernels
    cell_t = sum(kms.get(k, 0.0) for k in CELL_KERNELS)
    cell_bytes = 4 * nvox + 24 * V + 17 * Q
    achieved = cell_bytes / (cell_t * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": TRAFFIC.get(args.workload),
                "kernel": "cell stage: " + " + ".join(CELL_KERNELS),
                "kernel_ms": round(cell_t, 4), "algorithmic_bytes": int(cell_bytes),
                "bytes_model": "4 N + 24 V + 17 Q: field read once, FP32 position + normal per vertex, int4 quad + "
                               "colour byte (V, Q before simplification)"}
    # per-kernel figures: device ms per extraction (HIP events), bytes by
    # each kernel's own model
    R = (grid.ny - 1) * (grid.nz - 1)
    models = {
        "k_sweep": (4 * nvox, "4 N (field read once)"),
        "k_rows": (4 * A + 20 * A + 8 * R, "4 A list entries read + 20 A per-cell words + 8 R row totals written"),
        "k_vertices": (24 * V, "24 V (FP32 position + normal; the engine writes FP64, 48 V)"),
        "k_quads": (17 * Q + 32 * Q, "17 Q (int4 quad + colour) + 32 Q (first halfedge rows)"),
        "k_he_twins": (16 * Q + 32 * Q + 20 * Q + 4 * V,
                       "16 Q quads + 32 Q halfedge rows read, 20 Q twins and flags + 4 V he_v written (the one "
                       "build through rows, on the cell stage's V, Q; later builds are carried over)"),
        "k_triangles": (16 * Qf + 24 * Qf + 24 * 4 * Qf, "16 Q quads + 96 Q positions read, 24 Q written"),
    }
    kernels = []
    for name, t in sorted(kms.items(), key=lambda kv: -kv[1]):
        rec = {"name": name, "ms": round(t, 4), "pct_of_step": round(100 * t / ms, 2)}
        if name in models:
            b, model = models[name]
            rec.update(bytes=int(b), bytes_model=model, achieved_gbs=round(b / (t * 1e-3) / 1e9, 1),
                       frac=round(b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
        kernels.append(rec)
    out = {
        "metric": "DMC Gvoxels/s at 1024^3 f32 (%HBM roofline)",
        "value": round(nvox / (ms * 1e-3) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "pass": {0: "launched kernel by kernel", 1: "captured into a HIP graph", 2: "HIP graph replay"}[pass_mode],
        "ms_per_step_kernel_by_kernel_with_events": round(ms_timed, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(desc, flags="3X3Y colour + 3X3Y isolated + 3333 + halfedges + triangles",
                       mesh_in="HBM", vertex_math="f64", V_cells=V, Q_cells=Q, V=Vf, Q=Qf,
                       active_cells=A, mquads_per_s=round(Q / (ms * 1e-3) / 1e6, 2),
                       host_syncs_per_step=m.host_syncs, stage_ms=stages,
                       hash_load_factor="none: vertex and quad ids come from prefix sums over the cells, "
                                        "there are no hash tables to load"),
        "roofline": roofline,
        "kernels": kernels,
    }
    out["diagnostics"] = None if args.no_diagnostics else bench_diagnostics(eng)
    out["small_volume"] = None if args.no_small else bench_small()
    if not args.no_dropin:
        out["dropin"] = bench_dropin(eng, field, grid, iso)
    if not args.no_tmc:
        log("[bench] topologically correct MC leg")
        out["tmc"] = bench_tmc(eng, buf, field, grid, iso, args.steps, 1,
                               None if args.no_cpu_baseline else args.cpu_block)
    buf.free()
    eng.close()
    if not args.no_batch:
        log("[bench] batched leg: 4,096 mixed volumes")
        out["batched"] = bench_batch()
    if not args.no_serve:
        log("[bench] served leg: Poisson serving mix")
        out["served"] = bench_serve()
        log("[bench] served leg: Poisson stream of small volumes")
        out["served_small"] = bench_serve(pool_kind="small")
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline on a {args.cpu_block}^3 block")
        out["cpu_baseline"] = cpu_baseline(field, grid, iso, min(args.cpu_block, min(field.shape)))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
