#!/usr/bin/env python3
"""Headline benchmark: exact orient3d + in_sphere throughput (bit-exact) on MI355X.

Workload (the degenerate-lattice configuration): the 32x32x32 integer lattice
and 2^24 mixed queries (half orient3d on lattice planes and unit-cube corners,
half in_sphere on exactly cospherical unit-cube corners and lattice spheres,
every in_sphere query positively oriented); see del3d_amd/workloads.py. A
step is one batched call over all 2^24 device-resident queries.

Prints one JSON line. value = whole-job Gpred/s = queries * steps / timed
seconds (HIP events on the work stream, device synchronized on both sides).
roofline: the dominant kernel (the fast path) with the FP64 work the
predicates' filters define per query (111 ops per in_sphere, 32 per
orient3d), over that kernel's average duration measured by HIP events around
it in the timed region. cpu_baseline: the reference (compiled from its own
sources into oracle/_ref) or, where that was not built, the oracle port, on
one core, over the first 2^23 queries of the same workload; its signs are
compared with the GPU's.
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

METRIC = "insphere+orient3d Gpred/s (bit-exact); Hilbert sort Mpts/s; served p99 ms"
FP64_PEAK_TFLOPS = 78.6     # MI355X published FP64 (FMA counted as 2)
HBM_PEAK_GBS = 8000.0
OPS_INSPHERE, OPS_ORIENT = 111, 32


def cpu_baseline(pts, op, idx, gpu_signs, sample: int):
    """Times the reference (or the oracle port) on one core over a sample of
    the same workload, and checks the GPU signs against it."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O

    kind = "reference" if O.have_ref() else "port"
    n = min(sample, len(op))
    t0 = time.perf_counter()
    ref = O.pred_batch(pts, op[:n], idx[:n], "ref" if kind == "reference" else "oracle")
    dt = time.perf_counter() - t0
    match = bool(np.array_equal(ref, gpu_signs[:n]))
    return {
        "value": n / dt / 1e9,
        "unit": "Gpred/s",
        "cores": 1,
        "kind": kind,
        "sample": f"first {n} of the {len(op)} queries of the same workload, single thread",
        "seconds": round(dt, 3),
        "signs_equal_gpu": match,
    }


def load_traffic():
    """HBM bytes per fast-path launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), if any."""
    f = ROOT / "profiles" / "pred_fast_c2_traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--queries", type=int, default=1 << 24)
    ap.add_argument("--cpu-sample", type=int, default=1 << 23)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run the calibration copy kernel (for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one GPU, so --gpus {args.gpus} "
              "is not run (the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        return 2

    import del3d_amd as D
    from del3d_amd import workloads as W

    if args.calib_copy_mib:
        import ctypes as C
        lib = C.CDLL(str(ROOT / "del3d_amd" / "libdel3d_calib.so"))
        lib.del3d_calib_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        nbytes = args.calib_copy_mib << 20
        a = D.DeviceArray((nbytes,), np.uint8)
        b = D.DeviceArray((nbytes,), np.uint8)
        a.zero()
        for _ in range(5):
            lib.del3d_calib_copy(a.ptr, b.ptr, nbytes, None)
        D.synchronize()
        print(json.dumps({"calib_copy_bytes": nbytes, "launches": 5}))
        return 0

    t_gen = time.perf_counter()
    pts, op, idx = W.c2(args.queries, seed=2)
    t_gen = time.perf_counter() - t_gen
    n_q = len(op)
    n_ins = int((op != 0).sum())
    stream = D.Stream()
    xyz = D.DeviceArray.from_host(pts)
    dop = D.DeviceArray.from_host(op)
    didx = D.DeviceArray.from_host(idx)
    out = D.DeviceArray((n_q,), np.int8)
    st = D.predicates_stats(xyz, dop, didx, out, stream)

    for _ in range(args.warmup):
        D.predicates_timed(xyz, dop, didx, out, stream)
    D.synchronize()

    e0, e1 = D.Event(), D.Event()
    fast_ms = []
    D.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        fast_ms.append(D.predicates_timed(xyz, dop, didx, out, stream)[0])
    e1.record(stream)
    e1.synchronize()
    D.synchronize()
    wall = time.perf_counter() - w0
    total_ms = e0.elapsed_ms(e1)
    signs = out.to_host()

    value = n_q * args.steps / (total_ms / 1e3) / 1e9
    fast_avg = float(np.mean(fast_ms))
    ops = OPS_INSPHERE * n_ins + OPS_ORIENT * (n_q - n_ins)
    achieved = ops / (fast_avg / 1e3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gpred/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "degenerate-lattice: 32^3 integer lattice, 2^24 mixed "
                               "orient3d/in_sphere queries, all in_sphere exactly cospherical",
                   "points": len(pts), "queries": n_q, "in_sphere": n_ins, "seed": 2},
        "roofline": {
            "bound": "mfma",
            "kernel": "pred_fast",
            "achieved": round(achieved, 4),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": load_traffic(),
            "algorithmic_ops_per_launch": ops,
            "kernel_ms_avg": round(fast_avg, 4),
            "note": "FP64 VALU (no MFMA on this path); filter op counts 111/32 per query",
        },
        "decided_by": st,
        "wall_s_timed": round(wall, 4),
        "input_gen_s": round(t_gen, 2),
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(pts, op, idx, signs, args.cpu_sample)
        line["cpu_baseline"] = cb
        line["bit_exact_vs_cpu_sample"] = cb.pop("signs_equal_gpu")
        line["gpu_over_cpu"] = round(value / cb["value"], 1)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
