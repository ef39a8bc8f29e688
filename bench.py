#!/usr/bin/env python3
"""Benchmark of the MI355X HEVC reconstruction engine.

Default (--form served, --workload 1080p64): 64 concurrent 1920x1080 8-bit
4:2:0 streams, each a GOP of one intra frame and 15 P frames (every P frame
references the stream's previous <= 4 decoded frames; seeded syntax: CUs
8..64, PUs with AMP, quarter-pel MVs, 20 % explicitly weighted, TUs 4..32,
deblocking and SAO on), served by the continuous-batching server from
syntax already resident in HBM. A step is one serving pass over the whole
stream set (1024 frames) with every frame available at the start
(saturation); `value` is frames/s over the timed passes. The same frames run
through the plain form (one frame at a time, one kernel per stage) for
`plain` / `vs_plain`, and both forms are run again with frames arriving as a
Poisson process (240 frames/s aggregate) for p50/p99 per-frame latency.
Before timing, every frame's output of the served form is compared with the
plain form's (device digests of all three planes).

--workload cfg4 serves the mixed set instead: resolutions uniform over
360p/720p/1080p/2160p, 25 % 10-bit.

--form batched --workload cfg2: 32 independent 1080p frames (8 intra, 24 B)
in one batch replayed from a hipGraph. --form plain --workload b1080: one
1080p B frame, plain form. Both check their outputs against the reference
decoder's golden SHA-256 (tests/golden/frames.json) first.

Prints one JSON line. --gpus above 1 exits: this engine runs on one GPU per
process and shards streams, not frames.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import re
import subprocess
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hevc.js_amd"))

import hevcr  # noqa: E402
import serving  # noqa: E402
from workloads import BATCHES, batch_frames  # noqa: E402

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


def cpu_baseline(frames: list, budget_s: float):
    """The reference's reconstruction code on the host over the workload's
    independent frames: 1 thread cycling over them (about budget_s / 2 s), and
    nproc threads frame-parallel (thread i takes frames i, i + n, ...; about
    budget_s CPU seconds in all, wall clock over all threads)."""
    got = _recon_lib()
    if got is None:
        return None
    fn, fname, kind, build = got

    def run(frame):
        out = frame.new_picture()
        if fn(C.byref(frame.syntax), frame.ref_plane_array(), frame.num_ref_pics, None, None,
              hevcr.plane_out_array(out)) != 0:
            raise RuntimeError("CPU reconstruction failed")
    n, t0 = 0, time.perf_counter()
    while True:
        run(frames[n % len(frames)])
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s / 2 or n >= 4000:
            break
    single = n / el
    per_frame = el / n
    threads = max(1, min(os.cpu_count() or 1, 64))
    each = max(1, int(round(budget_s / per_frame / threads)))
    barrier = threading.Barrier(threads)
    spans = [None] * threads

    def work(i):
        barrier.wait()
        a = time.perf_counter()
        for j in range(each):
            run(frames[(i + j * threads) % len(frames)])
        spans[i] = (a, time.perf_counter())
    ts = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    wall = max(b for _, b in spans) - min(a for a, _ in spans)
    return {"value": round(threads * each / wall, 2), "unit": "frames/s", "cores": threads, "kind": kind,
            "sample": f"{threads} threads frame-parallel with independent contexts, {threads * each} frame "
                      f"reconstructions over the workload's {len(frames)} frames in {wall:.2f} s wall; {fname}",
            "single_thread": {"value": round(single, 2), "cores": 1,
                              "sample": f"{n} frame reconstructions cycling over the workload's frames, {el:.1f} s"},
            "scaling": round(threads * each / wall / single, 2), "cpu": _cpu_model(), "build": build}


def _recon_lib():
    """The reference decoder's reconstruction code (oracle/_ref, built by
    oracle/Makefile from the reference sources; the tree carries the build),
    else the CPU restatement (oracle/liboracle.so), saying which."""
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libhevcref_harness.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    info_path = os.path.join(ROOT, "oracle", "_ref", "BUILDINFO.txt")
    if os.path.exists(ref_so):
        build = open(info_path).read().strip().replace("\n", "; ") if os.path.exists(info_path) else "unrecorded"
        lib, fn, kind = C.CDLL(ref_so), "ref_recon_frame", "reference"
    elif os.path.exists(port_so):
        build = "oracle/_ref is absent from this tree: the oracle's C restatement of the reference is timed instead"
        lib, fn, kind = C.CDLL(port_so), "oracle_recon_frame", "port"
    else:
        return None
    f = getattr(lib, fn)
    f.argtypes = [C.POINTER(hevcr.FrameSyntax), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return f, fn, kind, build


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), "unknown")
    except (OSError, subprocess.SubprocessError):
        model = "unknown"
    return f"{model}, {os.cpu_count()} logical CPUs (lscpu / nproc)"


def _stream_items(spec: "serving.StreamSpec", nframes: int):
    """(syntax owner, syntax, slot) of a stream's first nframes frames (one
    closed GOP for the synthetic streams, the whole parsed bitstream otherwise)."""
    if spec.source:
        return [(a, a.syntax, slot) for a, slot in spec.source]
    items = []
    for t in range(nframes):
        fr = serving.gen_frame(spec, t)
        items.append((fr, fr.syntax, serving.out_slot(t)))
    return items


def _decode_chain(fn, spec, items, reps):
    """Decodes the stream's frames in order, reps times, with its own DPB."""
    dpb = [hevcr.Picture(spec.width, spec.height, spec.bit_depth) for _ in range(spec.dpb_slots)]
    arr = (C.c_void_p * (3 * len(dpb)))()
    n = 0
    for _ in range(reps):
        for _owner, syntax, slot in items:
            for k, r in enumerate(dpb):
                for c in range(3):
                    arr[3 * k + c] = r.planes[c].ctypes.data
            out = hevcr.Picture(spec.width, spec.height, spec.bit_depth)
            if fn(C.byref(syntax), arr, len(dpb), None, None, hevcr.plane_out_array(out)) != 0:
                raise RuntimeError("CPU reconstruction failed")
            dpb[slot] = out
            n += 1
    return n


def _chain_hashes(fn, spec, items):
    """SHA-256 per plane (write_yuv bytes) of each frame of the chain, decoded in order."""
    dpb = [hevcr.Picture(spec.width, spec.height, spec.bit_depth) for _ in range(spec.dpb_slots)]
    arr = (C.c_void_p * (3 * len(dpb)))()
    out_h = []
    for _owner, syntax, slot in items:
        for k, r in enumerate(dpb):
            for c in range(3):
                arr[3 * k + c] = r.planes[c].ctypes.data
        out = hevcr.Picture(spec.width, spec.height, spec.bit_depth)
        if fn(C.byref(syntax), arr, len(dpb), None, None, hevcr.plane_out_array(out)) != 0:
            raise RuntimeError("CPU reconstruction failed")
        dpb[slot] = out
        out_h.append(out.sha256())
    return out_h


def cpu_baseline_streams(specs: list, budget_s: float, served_sha: dict = None):
    """The reference's reconstruction code on the host: at 1 thread (one
    stream decoded in order, repeated to about budget_s / 2 seconds), and at
    nproc threads frame-parallel with independent contexts (thread i decodes
    stream i's chain with its own DPB, all threads started together, wall
    clock over all of them). Syntax generation is not timed. Then, untimed,
    the same threads hash every frame they reconstruct (SHA-256 per plane)
    and compare them with served_sha[(stream, frame)], the GPU's served frames
    hashed on the device."""
    got = _recon_lib()
    if got is None:
        return None
    fn, fname, kind, build = got
    threads = max(1, min(os.cpu_count() or 1, 64, len(specs)))
    gop = serving.GOP
    chains = [_stream_items(s, gop) for s in specs[:threads]]  # before any timing
    # 1 thread
    t0 = time.perf_counter()
    n1 = _decode_chain(fn, specs[0], chains[0], 1)
    once = time.perf_counter() - t0
    reps = max(1, int(budget_s / 2 / max(once, 1e-3)))
    t0 = time.perf_counter()
    n1 = _decode_chain(fn, specs[0], chains[0], reps)
    single = n1 / (time.perf_counter() - t0)
    # nproc threads, about budget_s CPU seconds in all
    reps_n = max(1, int(round(budget_s / (max(once, 1e-3) * threads))))
    barrier = threading.Barrier(threads)
    spans = [None] * threads
    counts = [0] * threads

    def work(i):
        barrier.wait()
        a = time.perf_counter()
        counts[i] = _decode_chain(fn, specs[i], chains[i], reps_n)
        spans[i] = (a, time.perf_counter())
    ts = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    wall = max(b for _, b in spans) - min(a for a, _ in spans)
    multi = sum(counts) / wall
    check = None
    if served_sha:
        hashes = [None] * threads
        vt = [threading.Thread(target=lambda i=i: hashes.__setitem__(i, _chain_hashes(fn, specs[i], chains[i])))
              for i in range(threads)]
        for t in vt:
            t.start()
        for t in vt:
            t.join()
        pairs = [(served_sha.get((i, t)), h) for i in range(threads) for t, h in enumerate(hashes[i])]
        equal = all(a == b for a, b in pairs)
        if not equal:
            print("bench.py: served frames differ from the reference's reconstruction", file=sys.stderr)
        check = {"frames": len(pairs), "equal": equal,
                 "method": "SHA-256 per plane of the write_yuv bytes: served frames hashed on the GPU "
                           "(hevcr_picture_sha256) against the reference's reconstruction of the same frames "
                           f"({fname}, hashlib)"}
    sp = specs[0]
    what = "pictures of the parsed bitstream" if sp.source else f"frames (closed GOPs of 1 intra + {gop - 1} P)"
    return {"value": round(multi, 2), "unit": "frames/s", "cores": threads, "kind": kind,
            "sample": f"{threads} threads, thread i decoding stream i of the served workload "
                      f"({sp.width}x{sp.height} {sp.bit_depth}-bit) in order with its own DPB: {sum(counts)} {what} "
                      f"in {wall:.2f} s wall; {fname}",
            "single_thread": {"value": round(single, 2), "cores": 1,
                              "sample": f"{n1} {what} of one stream in order, {n1 / single:.1f} s"},
            "scaling": round(multi / single, 2), "cpu": _cpu_model(), "build": build, "reference_check": check}


def reference_cli_context(path: str, runs: int = 3):
    """The reference decoder's CLI (oracle/_ref/hevc-decode, whole decoder:
    parsing + CABAC + reconstruction, its own WPP thread pool) on a bitstream,
    best of `runs` as tools/benchmark.sh does; its printed decode time."""
    exe = os.path.join(ROOT, "oracle", "_ref", "hevc-decode")
    if not os.path.exists(exe):
        return None
    best = None
    for _ in range(runs):
        try:
            out = subprocess.run([exe, path, "-o", "/dev/null"], capture_output=True, text=True, timeout=120).stdout
        except (OSError, subprocess.SubprocessError):
            return None
        m = re.search(r"Decoded (\d+) pictures in ([0-9.]+)ms", out)
        if not m:
            return None
        n, ms = int(m.group(1)), float(m.group(2))
        if best is None or ms < best[1]:
            best = (n, ms)
    return {"fps": round(best[0] * 1e3 / best[1], 2), "pictures": best[0], "ms": best[1], "runs": runs,
            "what": f"reference CLI (whole decoder incl. CABAC, its own thread pool) on "
                    f"{os.path.relpath(path, ROOT)}, best of {runs}", "cpu": _cpu_model()}


DROPIN_METRIC = "pictures/s through hevc_api.h (host parse + GPU reconstruction + frames to host), bit-exact"


def bench_dropin(args):
    """The drop-in decoder (hevc_api.h, libhevcr_decoder.so) on a whole
    bitstream: each of `threads` handles decodes the stream per step with
    hevc_decoder_decode and reads every frame (cropped uint16 planes on the
    host), all handles at once in their own threads (one shared engine
    underneath). The MD5 of one handle's frames is checked against the
    stream's golden; the reference CLI on the same stream is the baseline."""
    import hashlib
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import stream_goldens
    name = args.workload
    path = os.path.join(ROOT, "tests", "golden", "bitstreams", name + ".265")
    with open(path, "rb") as fh:
        data = fh.read()
    golden = next(s for s in stream_goldens.load() if s["name"] == name)
    L = hevcr.decoder_lib()
    threads = max(1, args.dropin_threads)
    decs = [L.hevc_decoder_create() for _ in range(threads)]

    def decode(d, check=False):
        if L.hevc_decoder_decode(d, data, len(data)) != 0:
            raise RuntimeError(L.hevcr_decoder_error(d))
        n = L.hevc_decoder_get_frame_count(d)
        f = hevcr.HEVCFrame()
        h = hashlib.md5() if check else None
        for i in range(n):
            if L.hevc_decoder_get_frame(d, i, C.byref(f)) != 0:
                raise RuntimeError("get_frame")
            if check:
                for pl in hevcr.frame_planes(f):
                    h.update(pl.astype("uint8").tobytes() if f.bit_depth <= 8 else pl.astype("<u2").tobytes())
        return n, (h.hexdigest() if check else None)

    try:
        n, md5 = decode(decs[0], check=True)
        with ThreadPoolExecutor(threads) as ex:
            for _ in range(args.warmup):
                list(ex.map(decode, decs))
            t0 = time.perf_counter()
            for _ in range(args.steps):
                counts = list(ex.map(decode, decs))
            wall = time.perf_counter() - t0
    finally:
        for d in decs:
            L.hevc_decoder_destroy(d)
    pics = sum(c for c, _ in counts) * args.steps
    cli = reference_cli_context(path)
    return {
        "metric": DROPIN_METRIC,
        "value": round(pics / wall, 2),
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "dtype": "u8" if golden["bit_depth"] <= 8 else "u16",
        "data": f"{os.path.relpath(path, ROOT)} (the reference's conformance bitstream)",
        "config": {"workload": f"{name}: {n} pictures {golden['width']}x{golden['height']} "
                               f"{golden['bit_depth']}-bit, whole-stream hevc_decoder_decode + get_frame",
                   "form": "dropin", "handles": threads},
        "bit_exact": {"equal": md5 == golden["md5_yuv_reference"], "method": "MD5 of the frames as write_yuv "
                      "writes them against the reference decoder's (and ffmpeg's) YUV MD5"},
        "roofline": None,
        "cpu_baseline": None if cli is None else {
            "value": cli["fps"], "unit": "pictures/s", "cores": os.cpu_count(), "kind": "reference",
            "sample": cli["what"], "cpu": cli["cpu"]},
    }


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` as measured by rocprofv3
    (FETCH_SIZE and WRITE_SIZE runs, scaled by the copy-kernel calibration) and
    recorded in profiles/*.md as 'traffic-per-launch: <workload> <kernel> <bytes>'."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(pdir, name)) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic-per-launch:" and parts[1] == workload and parts[2] == kernel:
                    return int(parts[3]), f"profiles/{name}"
    return None, None


def roofline_kernels(ktimes: dict, kbytes: dict, what: str, workload: str = ""):
    """Dominant kernel by GPU time (HIP events around each of its launches on the
    engine's stream), with achieved = its algorithmic bytes / its time."""
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    ms, launches = ktimes[dom]
    achieved = kbytes[dom] / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic, src = measured_traffic(workload, dom)
    out = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
           "algorithmic_bytes_per_launch": round(kbytes[dom] / max(launches, 1)),
           "avg_launch_ms": round(ms / max(launches, 1), 5), "launches": launches, "scope": what}
    if src:
        out["traffic_source"] = f"rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 per launch, {src}"
    return out


def time_plain(eng, runs, steps, warmup):
    """Plain form over (prepared, slot) pairs, one frame after another."""
    for _ in range(warmup):
        for pf, slot in runs:
            pf.run(slot)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for pf, slot in runs:
            pf.run(slot)
    eng.sync()
    return steps * len(runs) / (time.perf_counter() - t0)


def bench_plain(args, eng):
    """One frame, plain form; output checked against the reference golden first."""
    wl = WORKLOADS[args.workload]
    frame = hevcr.Frame(hevcr.preset(wl["preset"]), wl["seed"])
    k = frame.num_ref_pics
    sid = eng.stream(frame.width, frame.height, frame.bit_depth, k + 1)
    for slot in range(k):
        eng.upload(sid, slot, frame.refs[slot])
    pf = eng.prepare(sid, frame.syntax)
    pf.run(k)
    eng.sync()
    got = eng.download(sid, k, frame.width, frame.height, frame.bit_depth).sha256()
    if got != golden_hashes(wl["golden"]):
        print("bench.py: output differs from the reference golden", file=sys.stderr)
        sys.exit(1)
    for _ in range(args.warmup):
        pf.run(k)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pf.run(k)
    eng.sync()
    t1 = time.perf_counter()
    # per-stage and per-kernel GPU time over the same number of runs, HIP events
    # around each stage and kernel launch (outside the timed loop)
    eng.profiling(True)
    eng.reset_stage_times()
    for _ in range(args.steps):
        pf.run(k)
    eng.sync()
    stages = eng.stage_times()
    ktimes = eng.kernel_times()
    eng.profiling(False)
    lat = []
    for _ in range(args.latency_samples):
        a = time.perf_counter()
        pf.run(k)
        eng.sync()
        lat.append((time.perf_counter() - a) * 1e3)
    per_launch = {s2: (ms / args.steps) for s2, (ms, cnt) in stages.items()}
    kb = {k2: v * args.steps for k2, v in pf.kernel_bytes().items()}
    roof = roofline_kernels(ktimes, kb, f"{args.steps} plain runs of one frame", f"plain/{args.workload}")
    s = frame.syntax
    out = {
        "value": round(args.steps / (t1 - t0), 2),
        "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 4),
        "dtype": "u8" if frame.bit_depth == 8 else "u16",
        "data": f"synthetic (seeded syntax and reference pictures, preset {wl['preset']} seed {wl['seed']})",
        "config": {"workload": wl["desc"], "width": frame.width, "height": frame.height,
                   "bit_depth": frame.bit_depth, "form": "plain", "frames_per_step": 1, "cus": s.num_cus,
                   "pus": s.num_pus, "tus": s.num_tus, "coeffs": s.num_coeffs},
        "latency_ms": {"p50": round(percentile(lat, 0.5), 4), "p99": round(percentile(lat, 0.99), 4),
                       "samples": len(lat), "mode": "one frame submitted and waited for at a time"},
        "stages_ms_per_step": {k2: round(v, 5) for k2, v in per_launch.items()},
        "roofline": roof,
        "cpu_frames": [frame],
    }
    pf.release()
    eng.destroy_stream(sid)
    return out


def bench_batched(args, eng):
    wl = BATCHES[args.workload]
    ip, ni = wl["intra"]
    bp, nb = wl["inter"]
    frames, refs, members = batch_frames(args.workload)
    k = len(refs)
    n = len(frames)
    W, H, bd = frames[0].width, frames[0].height, frames[0].bit_depth
    sid = eng.stream(W, H, bd, k + n)
    for slot in range(k):
        eng.upload(sid, slot, refs[slot])
    slots = [k + i for i in range(n)]
    tuned = None
    if args.autotune:  # the fastest tuning for this size and bit depth, kept for the frames prepared below
        best, ms = eng.autotune([sid] * n, [f.syntax for f in frames], slots, reps=max(3, min(args.steps, 20)))
        tuned = {"chosen": best, "ms_per_replay": [round(x, 4) for x in ms],
                 "candidates": [hevcr.Tuning.candidate(c).as_dict() for c in range(hevcr.TUNE_CANDIDATES)]}
    prepared = [eng.prepare(sid, f.syntax) for f in frames]

    # plain form: reference outputs, golden checks, throughput
    for pf, slot in zip(prepared, slots):
        pf.run(slot)
    eng.sync()
    plain_hashes = [eng.download(sid, slot, W, H, bd).sha256() for slot in slots]
    for i, (p, s) in enumerate(members):
        g = wl["goldens"].get((p, s))
        if g and plain_hashes[i] != golden_hashes(g):
            print(f"bench.py: plain output of frame {i} differs from golden {g}", file=sys.stderr)
            sys.exit(1)
    plain_fps = time_plain(eng, list(zip(prepared, slots)), max(1, args.plain_steps), 1)
    plain_lat = []
    for i in range(min(args.latency_samples, 4 * n)):
        a = time.perf_counter()
        prepared[i % n].run(slots[i % n])
        eng.sync()
        plain_lat.append((time.perf_counter() - a) * 1e3)

    # batched form: bit-exact against the plain form, then timed graph replays
    batch = eng.batch(prepared, slots)
    for slot in slots:  # clear the outputs so the check sees the batch's own writes
        eng.upload(sid, slot, hevcr.Picture(W, H, bd))
    batch.run(True)
    eng.sync()
    for i, slot in enumerate(slots):
        if eng.download(sid, slot, W, H, bd).sha256() != plain_hashes[i]:
            print(f"bench.py: batched output of frame {i} differs from the plain form", file=sys.stderr)
            sys.exit(1)
    for _ in range(args.warmup):
        batch.run(True)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run(True)
    eng.sync()
    t1 = time.perf_counter()
    lat = []
    for _ in range(args.latency_samples):
        a = time.perf_counter()
        batch.run(True)
        eng.sync()
        lat.append((time.perf_counter() - a) * 1e3)
    # per-stage and per-kernel GPU time: the same launches without the graph, HIP
    # events around each stage and each kernel launch
    eng.profiling(True)
    eng.reset_stage_times()
    reps = max(1, min(args.steps, 10))
    for _ in range(reps):
        batch.run(False)
    eng.sync()
    stages = eng.stage_times()
    ktimes = eng.kernel_times()
    eng.profiling(False)
    per_launch = {s2: (ms / reps) for s2, (ms, cnt) in stages.items()}
    kb = {k2: v * reps for k2, v in batch.kernel_bytes().items()}
    roof = roofline_kernels(ktimes, kb, f"{reps} runs of the batch of {n} frames", f"batched/{args.workload}")
    fps = args.steps * n / (t1 - t0)
    out = {
        "value": round(fps, 2),
        "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 4),
        "dtype": "u8" if bd == 8 else "u16",
        "data": f"synthetic (seeded syntax and reference pictures; intra preset {ip} seeds 1..{ni}, "
                f"B preset {bp} seeds 1..{nb}, references of seed {wl['ref_seed']})",
        "config": {"workload": wl["desc"], "width": W, "height": H, "bit_depth": bd, "form": "batched",
                   "frames_per_step": n, "intra_frames": ni, "inter_frames": nb, "graph": True,
                   "tuning": eng.tuning(W, H, bd), "autotune": tuned},
        "latency_ms": {"p50": round(percentile(lat, 0.5), 4), "p99": round(percentile(lat, 0.99), 4),
                       "samples": len(lat),
                       "mode": "all frames of a step submitted at once; a frame completes with its step"},
        "plain": {"value": round(plain_fps, 2), "unit": "frames/s",
                  "latency_ms": {"p50": round(percentile(plain_lat, 0.5), 4),
                                 "p99": round(percentile(plain_lat, 0.99), 4)},
                  "form": "one frame at a time, one kernel per stage, no graph"},
        "vs_plain": round(fps / plain_fps, 3),
        "stages_ms_per_step": {k2: round(v, 5) for k2, v in per_launch.items()},
        "roofline": roof,
        "cpu_frames": frames,
    }
    batch.release()
    for pf in prepared:
        pf.release()
    eng.destroy_stream(sid)
    return out


def served_key(args) -> str:
    """Workload key for recorded profiler figures (resident serving passes of
    the default stream set, `--served-only --resident` under the profiler)."""
    default = args.streams in (0, SERVED[args.workload]["streams"]) and args.frames == serving.GOP
    return f"served/{args.workload}" if default else "served/custom"


def bench_served_only(args, eng, log, srv, specs, total, kbytes, wl):
    """Saturated serving passes only (frames submitted as host syntax, ingest
    included; no plain form, no Poisson pass): every k_mc / k_loopfilter / ...
    launch is a full serving step, so a profiler's per-kernel averages over
    this run compare with the roofline entry."""
    ingest = not args.resident
    for _ in range(args.warmup):
        serving.serve(srv, specs, True, close=False, ingest=ingest)
    walls = []
    for _ in range(args.steps):
        walls.append(serving.serve(srv, specs, True, close=False, ingest=ingest)[1])
    fps = total * len(walls) / sum(walls)
    eng.profiling(True)
    eng.reset_stage_times()
    serving.serve(srv, specs, True, close=False, ingest=ingest)
    ktimes = eng.kernel_times()
    eng.profiling(False)
    roof = roofline_kernels(ktimes, kbytes, f"one {'resident ' if not ingest else ''}serving pass over {total} frames",
                            served_key(args) if not ingest else "")
    log(f"served {fps:.1f} frames/s (served-only run, ingest={ingest})")
    res = {"value": round(fps, 2), "ms_per_step": round(sum(walls) / len(walls) * 1e3, 3), "dtype": "u8",
           "data": "synthetic", "config": {"workload": wl["desc"], "form": "served", "streams": len(specs),
                                           "ingest": ingest, "lanes": args.lanes, "max_inflight": args.max_inflight,
                                           "note": "served-only profiling run: no plain form, no Poisson pass"},
           "kernels": {k: {"ms_per_step": round(ms, 4), "launches": n} for k, (ms, n) in ktimes.items() if n},
           "roofline": roof}
    for s in specs:
        for pf in s.prepared:
            pf.release()
    srv.close()
    return res, None


SERVED = {
    "1080p64": dict(mix="1080p", streams=64,
                    desc="64 concurrent 1920x1080 8-bit 4:2:0 streams, GOP 1 intra + 15 P (P frames reference "
                         "the stream's previous <= 4 decoded frames), served with continuous batching"),
    "cfg4": dict(mix="cfg4", streams=64,
                 desc="64 concurrent streams, resolution uniform over 640x360/1280x720/1920x1080/3840x2160, "
                      "25 % 10-bit, GOP 1 intra + 15 P, served with continuous batching"),
    "bbb1080x64": dict(mix="bitstream", streams=64, bitstream=os.path.join("tests", "golden", "bitstreams",
                                                                          "bbb1080_50f.265"),
                       desc="64 concurrent streams each decoding the reference's 1080p Big Buck Bunny fixture "
                            "(50 pictures, P, wavefront, scaling lists) as parsed by the host front end, served "
                            "with continuous batching"),
}


def _new_server(args, eng, specs):
    srv = hevcr.Server(eng, int(args.budget_gb * 2**30), max_batch=args.max_batch, max_inflight=args.max_inflight,
                       lanes=args.lanes, ingest_threads=args.ingest_threads)
    for s in specs:
        s.sid = srv.open(s.width, s.height, s.bit_depth, s.dpb_slots)
    return srv


def _timed_passes(srv, specs, args, ingest):
    """Warm-up passes, then args.steps timed saturated passes: (frames/s,
    ms per pass, per-frame latencies in ms, frames per step, ingest bytes)."""
    total = sum(s.frames for s in specs)
    for _ in range(args.warmup):
        serving.serve(srv, specs, True, close=False, ingest=ingest)
    before = srv.info()
    walls, lat, sizes = [], [], []
    for _ in range(args.steps):
        recs, wall = serving.serve(srv, specs, True, close=False, ingest=ingest)
        walls.append(wall)
        lat += [(r["t_done"] - r["t_submit"]) * 1e3 for r in recs]
        sizes += [r["step_frames"] for r in recs]
    after = srv.info()
    phases = {k: after["ingest_" + k + "_s"] - before["ingest_" + k + "_s"]
              for k in ("validate", "lists", "stage", "upload")}
    return (total * len(walls) / sum(walls), sum(walls) / len(walls) * 1e3, lat, sizes,
            after["ingest_bytes"] - before["ingest_bytes"], after["ingest_busy_s"] - before["ingest_busy_s"],
            sum(walls), phases)


def bench_cfg4_sub(args, eng, log):
    """The cfg4 serving mix as a sub-result: 64 streams, resolution uniform over
    360p/720p/1080p/2160p, 25 % 10-bit, served from host syntax (ingest
    included) at saturation and under 240 frames/s Poisson arrivals, against
    the plain form from the same syntax; served == plain on every frame."""
    specs = serving.make_specs(64, "cfg4", seed=1)
    total = sum(s.frames for s in specs)
    serving.host_syntax(specs, compact=True, workers=16, log=log)
    srv = _new_server(args, eng, specs)
    plain_d, served_d = {}, {}
    order = [(si, t) for t in range(serving.GOP) for si in range(len(specs))]
    t0 = time.perf_counter()
    for si, t in order:
        s = specs[si]
        eng.decode(s.sid, s.host[t].syntax, serving.slot_of(s, t))
        plain_d[(si, t)] = eng.digests([(s.sid, serving.slot_of(s, t))])[0]
    plain_wall_verify = time.perf_counter() - t0

    def on_served(recs):
        for r, dg in zip(recs, eng.digests([(specs[r["tag"] // 100000].sid, r["out_slot"]) for r in recs])):
            served_d[divmod(r["tag"], 100000)] = dg
    serving.serve(srv, specs, True, on_done=on_served, close=False, ingest=True)
    equal = len(served_d) == total and all(served_d[k] == plain_d[k] for k in plain_d)
    if not equal:
        print("bench.py: cfg4 served output differs from the plain form", file=sys.stderr)
        sys.exit(1)
    plain_fps = total / min(serving.run_plain_ingest(eng, specs) for _ in range(2))
    fps, ms, lat_sat, _, ibytes, ibusy, twall, _ = _timed_passes(srv, specs, args, True)
    serving.poisson_arrivals(specs, args.rate, seed=7)
    recs, _ = serving.serve(srv, specs, False, close=False, ingest=True)
    lat = [(r["t_done"] - r["t_submit"]) * 1e3 for r in recs]
    plain_lat, _ = serving.serve_plain(eng, specs, ingest=True)
    plain_lat = [x * 1e3 for x in plain_lat]
    srv.close()
    log(f"cfg4: served {fps:.1f} frames/s, plain {plain_fps:.1f}, Poisson p99 {serving.percentile(lat, 99):.2f} ms")
    del plain_wall_verify
    return {"workload": SERVED["cfg4"]["desc"], "frames_per_step": total,
            "resolutions": sorted({f"{s.width}x{s.height}/{s.bit_depth}" for s in specs}),
            "value": round(fps, 2), "unit": "frames/s", "ms_per_step": round(ms, 3), "ingest": True,
            "ingest_gbps": round(ibytes / twall / 1e9, 2), "ingest_bytes_per_frame": round(ibytes / (total * args.steps)),
            "plain_ingest": round(plain_fps, 2), "vs_plain_ingest": round(fps / plain_fps, 3),
            "latency_ms": {"p50": round(serving.percentile(lat, 50), 3), "p99": round(serving.percentile(lat, 99), 3),
                           "offered_fps": args.rate, "samples": len(lat)},
            "plain_latency_ms": {"p50": round(serving.percentile(plain_lat, 50), 3),
                                 "p99": round(serving.percentile(plain_lat, 99), 3)},
            "latency_saturated_ms": {"p50": round(serving.percentile(lat_sat, 50), 3),
                                     "p99": round(serving.percentile(lat_sat, 99), 3)},
            "bit_exact": {"served_vs_plain_frames": total, "equal": True, "method": "device digests of Y, Cb, Cr"}}


def bench_served(args, eng, log):
    wl = SERVED[args.workload]
    if wl["mix"] == "bitstream":
        specs = serving.bitstream_specs(os.path.join(ROOT, wl["bitstream"]), args.streams or wl["streams"])
        log(f"parsed {wl['bitstream']}: {specs[0].frames} pictures")
    else:
        specs = serving.make_specs(args.streams or wl["streams"], wl["mix"], seed=1, frames=args.frames)
    total = sum(s.frames for s in specs)
    # every frame's syntax in host memory, as the streams' parsers hand it over (compact coefficients)
    serving.host_syntax(specs, compact=not args.dense, workers=16, log=log)
    srv = _new_server(args, eng, specs)
    # resident copies (syntax prepared in HBM ahead of time): the plain form's input and the resident figure
    serving.prepare_streams(eng, specs, workers=16, log=log)
    nbytes = {k: 0.0 for k in hevcr.STAGES}
    kbytes = {k: 0.0 for k in hevcr.KERNELS}
    for s in specs:
        for pf in s.prepared:
            for k, v in pf.bytes().items():
                nbytes[k] += v
            for k, v in pf.kernel_bytes().items():
                kbytes[k] += v

    if args.served_only:
        return bench_served_only(args, eng, log, srv, specs, total, kbytes, wl)

    # bit-exactness: every frame served from host syntax (ingest) equals the
    # plain form's (device digests); SHA-256 per plane of every served frame is
    # kept for the comparison with the reference's own reconstruction in the
    # CPU-baseline leg. A completed frame's slot is rewritten only 5 frames
    # later, and at most max_inflight - 1 later steps run when its completion
    # is seen.
    plain_d, served_d, served_sha = {}, {}, {}

    def after_plain(si, t):
        plain_d[(si, t)] = eng.digests([(specs[si].sid, serving.slot_of(specs[si], t))])[0]

    def on_served(recs):
        refs = [(specs[r["tag"] // 100000].sid, r["out_slot"]) for r in recs]
        for r, dg, sh in zip(recs, eng.digests(refs), eng.sha256(refs)):
            served_d[divmod(r["tag"], 100000)] = dg
            served_sha[divmod(r["tag"], 100000)] = sh
    serving.run_plain(eng, specs, after_frame=after_plain)
    serving.serve(srv, specs, True, on_done=on_served, close=False, ingest=True)
    bad = [k for k in plain_d if served_d.get(k) != plain_d[k]]
    if bad or len(served_d) != total:
        print(f"bench.py: served output differs from the plain form for {len(bad)} frames, e.g. {bad[:4]}",
              file=sys.stderr)
        sys.exit(1)
    log(f"served (ingest) == plain on all {total} frames")

    # plain form, all frames available (back to back, no waits between frames):
    # syntax resident in HBM, and from host syntax (each frame prepared and uploaded first)
    serving.run_plain(eng, specs)
    plain_fps = total / min(serving.run_plain(eng, specs) for _ in range(max(1, args.plain_steps)))
    plain_ingest_fps = total / min(serving.run_plain_ingest(eng, specs) for _ in range(2))
    log(f"plain {plain_fps:.1f} frames/s resident, {plain_ingest_fps:.1f} from host syntax")

    # served form at saturation: from host syntax (the headline), then resident
    fps, ms_step, lat_sat, step_sizes, ibytes, ibusy, twall, iphase = _timed_passes(srv, specs, args, True)
    log(f"served {fps:.1f} frames/s from host syntax ({ibytes / twall / 1e9:.1f} GB/s uploaded)")
    fps_res, ms_res, _, _, _, _, _, _ = _timed_passes(srv, specs, args, False)
    log(f"served {fps_res:.1f} frames/s from resident syntax")

    # per-stage and per-kernel GPU time (HIP events around each stage and each
    # kernel launch on the engine's stream) over one more saturated pass of the
    # resident form, whose steps are full and the same from run to run (the
    # kernels' efficiency; a profiler run of `--served-only --resident` sees the
    # same launches), and over one more pass from host syntax (where ingest
    # decides the steps)
    def profile_pass(ingest):
        eng.profiling(True)
        eng.reset_stage_times()
        serving.serve(srv, specs, True, close=False, ingest=ingest)
        st, kt = eng.stage_times(), eng.kernel_times()
        eng.profiling(False)
        ks = {k: {"ms_per_step": round(ms, 4), "launches": n,
                  "gbps": round(kbytes[k] / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
              for k, (ms, n) in kt.items() if n}
        return st, kt, ks
    stages, ktimes, kernels = profile_pass(False)
    per_pass = {k: ms for k, (ms, n) in stages.items()}
    roof = roofline_kernels(ktimes, kbytes, f"one resident serving pass over {total} frames (full steps)",
                            served_key(args))
    _, ktimes_in, kernels_in = profile_pass(True)
    roof_in = roofline_kernels(ktimes_in, kbytes, f"one serving pass over {total} frames from host syntax", "")

    # latency under a Poisson offered load, both forms from host syntax
    serving.poisson_arrivals(specs, args.rate, seed=7)
    recs, wall_p = serving.serve(srv, specs, False, close=False, ingest=True)
    lat = [(r["t_done"] - r["t_submit"]) * 1e3 for r in recs]
    plain_lat, _ = serving.serve_plain(eng, specs, ingest=True)
    plain_lat = [x * 1e3 for x in plain_lat]
    log(f"poisson {args.rate} fps: served p99 {serving.percentile(lat, 99):.2f} ms, "
        f"plain p99 {serving.percentile(plain_lat, 99):.2f} ms")
    for s in specs:
        for pf in s.prepared:
            pf.release()
        s.prepared = []
    srv.close()

    res = {
        "value": round(fps, 2),
        "ms_per_step": round(ms_step, 3),
        "dtype": "u8" if all(s.bit_depth == 8 for s in specs) else "u8/u16",
        "data": ("bitstream " + wl["bitstream"] + " parsed by the host front end (libhevcr_front.so)"
                 if wl["mix"] == "bitstream" else
                 "synthetic (seeded syntax per stream and frame; reference pictures are the streams' own decoded "
                 "frames)"),
        "config": {"workload": wl["desc"], "form": "served", "streams": len(specs),
                   "frames_per_stream": specs[0].frames, "frames_per_step": total,
                   "resolutions": sorted({f"{s.width}x{s.height}/{s.bit_depth}" for s in specs}),
                   "max_batch": args.max_batch, "max_inflight": args.max_inflight, "lanes": args.lanes,
                   "dpb_budget_gb": args.budget_gb, "ingest_threads": args.ingest_threads,
                   "coefficients": "dense" if args.dense else "compact",
                   "step": "one serving pass over every frame of every stream, all frames available at the start as "
                           "host syntax (the parsers' output); ingest (validation, work lists, upload over PCIe) is "
                           "inside the timed region"},
        "ingest": {"bytes_per_frame": round(ibytes / (total * args.steps)),
                   "gbps": round(ibytes / twall / 1e9, 2),
                   "host_busy_s_per_pass": round(ibusy / args.steps, 3),
                   "host_ms_per_frame": {k: round(v / (total * args.steps) * 1e3, 4) for k, v in iphase.items()},
                   "mode": "hevcr_server_submit_syntax: ingest threads validate, build work lists and upload each "
                           "frame (pinned staging, own device stream) while steps run"},
        "bit_exact": {"served_vs_plain_frames": total, "equal": True, "method": "device digests of Y, Cb, Cr"},
        "latency_ms": {"p50": round(serving.percentile(lat, 50), 3), "p99": round(serving.percentile(lat, 99), 3),
                       "samples": len(lat), "offered_fps": args.rate,
                       "completed_fps": round(total / wall_p, 1),
                       "arrival_span_s": round(max(max(s.arrivals[:s.frames]) for s in specs), 3),
                       "mode": "Poisson arrivals of host syntax; latency = host-observed completion - arrival"},
        "latency_saturated_ms": {"p50": round(serving.percentile(lat_sat, 50), 3),
                                 "p99": round(serving.percentile(lat_sat, 99), 3),
                                 "mode": "all frames submitted at the start"},
        "server_steps": {"mean_frames": round(sum(step_sizes) / len(step_sizes), 2),
                         "max_frames": max(step_sizes)},
        "served_resident": {"value": round(fps_res, 2), "unit": "frames/s", "ms_per_step": round(ms_res, 3),
                            "form": "served from syntax already resident in HBM (prepared before timing)"},
        "plain": {"value": round(plain_fps, 2), "unit": "frames/s",
                  "form": "one frame at a time, one kernel per stage, syntax resident in HBM"},
        "plain_ingest": {"value": round(plain_ingest_fps, 2), "unit": "frames/s",
                         "latency_ms": {"p50": round(serving.percentile(plain_lat, 50), 3),
                                        "p99": round(serving.percentile(plain_lat, 99), 3),
                                        "offered_fps": args.rate},
                         "form": "one frame at a time from host syntax (hevcr_decode_frame: prepare, upload, "
                                 "one kernel per stage)"},
        "vs_plain": round(fps / plain_fps, 3),
        "vs_plain_ingest": round(fps / plain_ingest_fps, 3),
        "stages_ms_per_step": {k: round(v, 4) for k, v in per_pass.items()},
        "kernels": kernels,
        "roofline": roof,
        "kernels_ingest_pass": kernels_in,
        "roofline_ingest_pass": roof_in,
    }
    if args.workload == "1080p64" and not args.no_cfg4:
        res["cfg4"] = bench_cfg4_sub(args, eng, log)
    cpu_specs = [s if s.source else serving.StreamSpec(s.width, s.height, s.bit_depth, s.seed, serving.GOP)
                 for s in specs]
    return res, (cpu_specs, served_sha)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (served: 3 passes; batched/plain: 40)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (served: 1; batched/plain: 5)")
    ap.add_argument("--form", default="served", choices=("served", "batched", "plain", "dropin"))
    ap.add_argument("--dropin-threads", type=int, default=1,
                    help="dropin: decoder handles decoding the stream at once, each in its own thread")
    ap.add_argument("--workload", default=None, help="served: " + ", ".join(sorted(SERVED)) + "; plain: " +
                    ", ".join(sorted(WORKLOADS)) + "; batched: " + ", ".join(sorted(BATCHES)))
    ap.add_argument("--streams", type=int, default=0, help="served: number of streams (default 64)")
    ap.add_argument("--frames", type=int, default=serving.GOP, help="served: frames per stream")
    ap.add_argument("--rate", type=float, default=240.0, help="served: Poisson offered load, frames/s")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="served: DPB budget")
    ap.add_argument("--max-batch", type=int, default=64)
    ap.add_argument("--max-inflight", type=int, default=2)
    ap.add_argument("--lanes", type=int, default=1, help="served: device streams steps run on concurrently")
    ap.add_argument("--autotune", action="store_true",
                    help="batched: time every tuning candidate on the batch's frames first and use the fastest")
    ap.add_argument("--served-only", action="store_true",
                    help="served: only the saturated passes (for profiler runs); no plain form or Poisson pass")
    ap.add_argument("--plain-steps", type=int, default=3, help="passes over the frames for the plain comparison")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--latency-samples", type=int, default=50)
    ap.add_argument("--ingest-threads", type=int, default=max(1, min(16, os.cpu_count() or 1) - 1),
                    help="served: host threads preparing frames submitted as syntax (default: one core "
                         "left to the thread that launches the steps)")
    ap.add_argument("--dense", action="store_true", help="served: dense coefficients instead of the compact form")
    ap.add_argument("--resident", action="store_true",
                    help="served-only: serve syntax prepared in HBM ahead of time (no ingest in the timed passes)")
    ap.add_argument("--no-cfg4", action="store_true", help="served 1080p64: skip the cfg4 sub-result")
    ap.add_argument("--tune", default="", metavar="K=V,...",
                    help="tuning for every picture size of the run, e.g. intra_persistent=0 (default: the engine's)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run the counter-calibration copy kernel on BYTES bytes and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this engine runs on one GPU per process (asked for {max(args.gpus, world)}); "
              "streams shard across GPUs with no collective, so no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)
    if args.form == "dropin":
        args.steps = args.steps or 5
        args.warmup = 1 if args.warmup is None else args.warmup
        args.workload = args.workload or "bbb1080_50f"
        res = bench_dropin(args)
        line = {"metric": res.pop("metric"), "value": res.pop("value"), "unit": "pictures/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None}
        line.update(res)
        print(json.dumps(line), flush=True)
        return
    if args.steps is None:
        args.steps = 3 if args.form == "served" else 40
    if args.warmup is None:
        args.warmup = 1 if args.form == "served" else 5
    if args.workload is None:
        args.workload = {"served": "1080p64", "batched": "cfg2", "plain": "b1080"}[args.form]
    table = {"served": SERVED, "batched": BATCHES, "plain": WORKLOADS}[args.form]
    if args.workload not in table:
        ap.error(f"unknown {args.form} workload {args.workload}")

    eng = hevcr.Engine(0)
    if args.tune:
        kv = dict(x.split("=") for x in args.tune.split(","))
        for w, h in serving.RESOLUTIONS + ((416, 240), (1920, 1088)):
            for bd in (8, 10):
                eng.set_tuning(w, h, bd, **{**eng.tuning(w, h, bd), **{k: int(v) for k, v in kv.items()}})
    if args.calibrate_copy:
        hevcr._check(hevcr.engine_lib().hevcr_bench_copy(eng.h, args.calibrate_copy, 5), "hevcr_bench_copy")
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy, "reps": 5}))
        return
    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    cpu_frames = cpu_specs = None
    if args.form == "served":
        res, cpu_specs = bench_served(args, eng, log)
        if cpu_specs is None:
            args.no_cpu_baseline = True
    else:
        res = bench_batched(args, eng) if args.form == "batched" else bench_plain(args, eng)
        cpu_frames = res.pop("cpu_frames")
    line = {
        "metric": METRIC,
        "value": res.pop("value"),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
    }
    line.update(res)
    line.setdefault("bit_exact", True)  # plain / batched: checked against the reference goldens above
    if args.no_cpu_baseline:
        line["cpu_baseline"] = None
    elif cpu_specs is not None:
        specs_cpu, served_sha = cpu_specs
        line["cpu_baseline"] = cpu_baseline_streams(specs_cpu, args.cpu_budget, served_sha)
        chk = line["cpu_baseline"].get("reference_check") if line["cpu_baseline"] else None
        if chk:
            line["bit_exact"]["vs_reference"] = chk
            line["bit_exact"]["equal"] = line["bit_exact"]["equal"] and chk["equal"]
        line["cpu_context"] = reference_cli_context(os.path.join(ROOT, "tests", "golden", "bitstreams",
                                                                 "bbb1080_50f.265"))
    else:
        line["cpu_baseline"] = cpu_baseline(cpu_frames, args.cpu_budget)
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
