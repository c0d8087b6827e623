#!/usr/bin/env python3
"""Benchmark: DMC extraction throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload schwarz_1024]

A step is one full dualMC of the workload's field (already resident in
HBM): classification, compaction, vertices, quadrilaterals, 3X3Y (colour and
isolated), 3333, halfedges and triangles; the mesh stays in HBM.  `value` is
Gvoxels/s = nx*ny*nz / (time per step).  The roofline object is for the
field sweep (k_sweep), timed with HIP events on the engine's stream over
the timed steps.  The CPU baseline is the reference cpp_dmc compiled from its
sources (oracle/_ref) -- or the C restatement if that is absent -- timed on a
sub-block of the same field on one host core.
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
# HBM bytes per k_sweep launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE,
# scaled by the copy-kernel calibration (profiles/k_sweep_schwarz_1024_traffic.md)
TRAFFIC = {"schwarz_1024": 5375285233}


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


def cpu_baseline(field, grid, iso, block):
    """Reference cpp_dmc (estimateGradient + flip_gradient + dualMC with
    sFlag) on the block^3 corner of the field, one core."""
    import oracle
    sub = np.ascontiguousarray(field[:block, :block, :block])
    g = oracle.Grid(block, block, block, grid.x0, grid.y0, grid.z0, grid.dx, grid.dy, grid.dz)
    if os.path.exists(oracle.REF_SO):
        ref = oracle.Reference()
        kind = "reference"
        t = time.perf_counter()
        nv, nq = ref.run_float(g, sub, iso, True)
        dt = time.perf_counter() - t
    else:
        orc = oracle.Oracle()
        kind = "port"
        t = time.perf_counter()
        s = sub.astype(np.float64).ravel()
        m = orc.dualmc(g, s, orc.gradient(g, s), iso, True)
        dt = time.perf_counter() - t
        nv, nq = len(m["v"]), len(m["quads"])
    return {"value": round(block ** 3 / dt / 1e9, 6), "unit": "Gvoxels/s", "cores": 1, "kind": kind,
            "sample": f"{block}^3 corner block of the same field (k,j,i < {block}); UniformGrid from the float "
                      f"field, estimateGradient + flip_gradient + dualMC(sFlag=true); {dt:.2f} s, "
                      f"V={nv} Q={nq}; host CPU: {cpu_model()} ({os.cpu_count()} logical)"}


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
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
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
    flags = tmc_amd.SIMPLIFY_ALL | tmc_amd.TRIANGLES
    for w in range(args.warmup):
        eng.extract(buf, iso, grid, flags=flags, shape=field.shape)
        log(f"[bench] warmup {w + 1}/{args.warmup}")
    eng.synchronize()
    stage_sum = {}
    t = time.perf_counter()
    for s in range(args.steps):
        m = eng.extract(buf, iso, grid, flags=flags | tmc_amd.STAGE_TIMING, shape=field.shape)
        for k, v in eng.stage_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    eng.synchronize()
    dt = time.perf_counter() - t
    ms = dt / args.steps * 1e3
    stages = {k: round(v / args.steps, 4) for k, v in stage_sum.items()}
    log(f"[bench] {args.steps} steps, {ms:.3f} ms/step; stages {stages}")
    summ = m.summary()
    V, Q, A = summ["nv_cells"], summ["nq_cells"], summ["active_cells"]
    # field sweep: one read of the field (4 B per point) plus what it writes:
    # 4 B per active cell into the segment lists and 4 B per segment count
    sweep_ms = stages.get("sweep", float("nan"))
    R = (grid.ny - 1) * (grid.nz - 1)
    S = (grid.nx - 1 + 255) // 256
    sweep_bytes = 4 * nvox + 4 * A + 4 * R * S
    achieved = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": TRAFFIC.get(args.workload),
                "kernel": "k_sweep", "kernel_ms": sweep_ms, "algorithmic_bytes": sweep_bytes}
    # cell stage as a whole (sweep, rows, vertices, quads): field read once,
    # FP64 position + normal per vertex (48 B), 4 ids + 1 colour per quad
    ext_ms = sum(stages.get(k, 0.0) for k in ("sweep", "rows", "vertices", "quads"))
    ext_bytes = 4 * nvox + 48 * V + 17 * Q
    out = {
        "metric": "DMC Gvoxels/s at 1024^3 f32 (%HBM roofline)",
        "value": round(nvox / (ms * 1e-3) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(desc, flags="3X3Y colour + 3X3Y isolated + 3333 + halfedges + triangles",
                       mesh_in="HBM", vertex_math="f64", V_cells=V, Q_cells=Q, V=summ["nv"], Q=summ["nq"],
                       active_cells=A, mquads_per_s=round(Q / (ms * 1e-3) / 1e6, 2),
                       extraction_ms=round(ext_ms, 4),
                       extraction_roofline_frac=round(ext_bytes / (ext_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       stage_ms=stages),
        "roofline": roofline,
    }
    buf.free()
    eng.close()
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline on a {args.cpu_block}^3 block")
        out["cpu_baseline"] = cpu_baseline(field, grid, iso, min(args.cpu_block, min(field.shape)))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
