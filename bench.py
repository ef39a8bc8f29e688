#!/usr/bin/env python3
"""Benchmark of the MI355X HEVC reconstruction engine.

Default workload: one 1920x1080 8-bit 4:2:0 B frame (seeded syntax: PUs 8..64
with AMP, 60% uni / 40% bi-prediction, quarter-pel MVs, 20% explicitly
weighted, residual TUs 4..32, deblocking and SAO on), reconstructed by the
plain form (one kernel per stage, launched in order) from syntax and reference
pictures already resident in HBM. A step is one full reconstruction of that
frame (expand, MC, residual, intra, deblock V/H, SAO). The output planes are
checked bit-exact against the reference decoder's golden SHA-256 before timing.

Prints one JSON line (see the keys below). --gpus above 1 exits: this engine
runs on one GPU per process and shards streams, not frames.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hevc.js_amd"))

import hevcr  # noqa: E402

METRIC = "1080p recon frames/s served vs plain GPU, bit-exact; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "b1080": dict(preset=1, seed=1, golden="b1080_s1",
                  desc="1920x1080 8-bit 4:2:0 B frame, plain form (one kernel per stage)"),
    "p1080": dict(preset=5, seed=1, golden="p1080_s1",
                  desc="1920x1080 8-bit 4:2:0 P frame, plain form (one kernel per stage)"),
    "intra416": dict(preset=0, seed=1, golden="intra416_s1",
                     desc="416x240 8-bit 4:2:0 intra frame, plain form (one kernel per stage)"),
    "intra1080": dict(preset=2, seed=1, golden="intra1080_s1",
                      desc="1920x1080 8-bit 4:2:0 intra frame, plain form (one kernel per stage)"),
    "b2160_10bit": dict(preset=3, seed=1, golden="b2160_10bit_s1",
                        desc="3840x2160 10-bit 4:2:0 B frame, plain form (one kernel per stage)"),
}

# Kernels behind each timed stage (names as rocprofv3 reports them).
STAGE_KERNELS = {
    "expand": ["k_expand_cu", "k_expand_tu", "k_expand_pu"],
    "mc": ["k_mc"],
    "residual": ["k_residual<2..5>"],
    "intra": ["k_intra (one launch per CTU diagonal)"],
    "deblock": ["k_deblock<vertical>", "k_deblock<horizontal>"],
    "sao": ["k_sao (one launch per plane)"],
}


def golden_hashes(name: str):
    path = os.path.join(ROOT, "tests", "golden", "frames.json")
    with open(path) as f:
        for c in json.load(f)["cases"]:
            if c["name"] == name:
                return c["stages"]["sao"]
    return None


def percentile(xs, q):
    xs = sorted(xs)
    k = (len(xs) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_baseline(frame: hevcr.Frame, budget_s: float):
    """The reference decoder's own reconstruction code (oracle/_ref, built from
    the reference sources) on the host, 1 thread, on the same frame; the CPU
    restatement (oracle/liboracle.so) when that build is absent."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libhevcref_harness.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if os.path.exists(ref_so):
        lib, fn, kind = C.CDLL(ref_so), "ref_recon_frame", "reference"
    elif os.path.exists(port_so):
        lib, fn, kind = C.CDLL(port_so), "oracle_recon_frame", "port"
    else:
        return None
    f = getattr(lib, fn)
    f.argtypes = [C.POINTER(hevcr.FrameSyntax), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    out = frame.new_picture()
    refs = frame.ref_plane_array()
    n, t0 = 0, time.perf_counter()
    while True:
        rc = f(C.byref(frame.syntax), refs, frame.num_ref_pics, None, None, hevcr.plane_out_array(out))
        if rc != 0:
            return None
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 2000:
            break
    return {"value": n / el, "unit": "frames/s", "cores": 1, "kind": kind,
            "sample": f"{n} reconstructions of the same frame ({el:.1f} s, 1 thread, {fn})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="b1080", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--latency-samples", type=int, default=200)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the counter-calibration copy kernel on BYTES bytes and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this engine runs on one GPU per process (asked for {max(args.gpus, world)}); "
              "streams shard across GPUs with no collective, so no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    eng = hevcr.Engine(0)
    if args.calibrate_copy:
        hevcr._check(hevcr.engine_lib().hevcr_bench_copy(eng.h, args.calibrate_copy, 5), "hevcr_bench_copy")
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy, "reps": 5}))
        return

    wl = WORKLOADS[args.workload]
    frame = hevcr.Frame(hevcr.preset(wl["preset"]), wl["seed"])
    k = frame.num_ref_pics
    sid = eng.stream(frame.width, frame.height, frame.bit_depth, k + 1)
    for slot in range(k):
        eng.upload(sid, slot, frame.refs[slot])
    pf = eng.prepare(sid, frame.syntax)

    # bit-exactness of the measured path against the reference's golden hashes
    pf.run(k)
    eng.sync()
    got = eng.download(sid, k, frame.width, frame.height, frame.bit_depth).sha256()
    want = golden_hashes(wl["golden"])
    bit_exact = want is not None and got == want
    if not bit_exact:
        print(f"bench.py: output differs from the reference golden ({got} vs {want})", file=sys.stderr)
        sys.exit(1)

    for _ in range(args.warmup):
        pf.run(k)
    eng.sync()
    eng.profiling(True)
    eng.reset_stage_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pf.run(k)
    eng.sync()
    t1 = time.perf_counter()
    stages = eng.stage_times()
    eng.profiling(False)
    fps = args.steps / (t1 - t0)

    lat = []
    for _ in range(args.latency_samples):
        a = time.perf_counter()
        pf.run(k)
        eng.sync()
        lat.append((time.perf_counter() - a) * 1e3)

    nbytes = pf.bytes()
    per_launch_ms = {s: (ms / n if n else 0.0) for s, (ms, n) in stages.items()}
    dom = max(per_launch_ms, key=lambda s: per_launch_ms[s])
    achieved = nbytes[dom] / (per_launch_ms[dom] * 1e-3) / 1e9 if per_launch_ms[dom] > 0 else 0.0

    base = None if args.no_cpu_baseline else cpu_baseline(frame, args.cpu_budget)
    s = frame.syntax
    line = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if frame.bit_depth == 8 else "u16",
        "data": "synthetic (seeded syntax and reference pictures, tests/golden preset %d seed %d)"
                % (wl["preset"], wl["seed"]),
        "config": {"workload": wl["desc"], "width": frame.width, "height": frame.height,
                   "bit_depth": frame.bit_depth, "form": "plain", "cus": s.num_cus, "pus": s.num_pus,
                   "tus": s.num_tus, "coeffs": s.num_coeffs},
        "bit_exact": bit_exact,
        "latency_ms": {"p50": round(percentile(lat, 0.5), 4), "p99": round(percentile(lat, 0.99), 4),
                       "samples": len(lat), "mode": "one frame submitted and waited for at a time"},
        "stages_ms_per_frame": {k2: round(v, 5) for k2, v in per_launch_ms.items()},
        "roofline": {"bound": "hbm", "stage": dom, "kernels": STAGE_KERNELS[dom],
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "algorithmic_bytes": round(nbytes[dom])},
        "cpu_baseline": base,
    }
    print(json.dumps(line), flush=True)
    pf.release()
    eng.destroy_stream(sid)
    eng.close()


if __name__ == "__main__":
    main()
