"""Serving workload and drivers.

Synthetic streams: a GOP of one intra frame and 15 P frames; the P frame at
position k of its GOP references the stream's previous min(k, 4) decoded
frames (closed GOP), so each stream keeps a DPB of 5 pictures (4 references
plus the picture being reconstructed). Frame t of a stream is written to DPB
slot t % 5 and its list entry for the frame decoded j + 1 frames earlier points
at slot (t - 1 - j) % 5. Syntax follows the intra and P presets of the
generator, at the stream's resolution and bit depth (10-bit streams use the
Main10 distributions: levels x4, QP 17..37).

Bitstream streams: a stream can instead replay a real bitstream parsed by the
host front end (bitstream_specs), every stream with its own DPB.

Drivers:
  - serve(): frames submitted at their arrival times to an hevcr server,
    which batches whatever is ready each step (continuous batching);
  - run_plain(): the same frames one at a time through the plain form
    (one kernel per stage), in an order that respects each stream's chain.
"""
from __future__ import annotations

import random
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import ctypes as C

import numpy as np

import hevcr

GOP = 16
MAX_REFS = 4
DPB_SLOTS = MAX_REFS + 1
RESOLUTIONS = ((640, 360), (1280, 720), (1920, 1080), (3840, 2160))


@dataclass
class StreamSpec:
    width: int
    height: int
    bit_depth: int
    seed: int
    frames: int = GOP
    sid: int = -1                      # engine / server stream id
    prepared: list = field(default_factory=list)
    arrivals: list = field(default_factory=list)  # seconds from the start, per frame
    source: list | None = None         # parsed bitstream: [(SyntaxArrays, slot)] per frame
    dpb_slots: int = DPB_SLOTS
    host: list = field(default_factory=list)  # per frame: its syntax in host memory (an object with .syntax)


def frame_cfg(spec: StreamSpec, t: int) -> hevcr.SynthCfg:
    k = t % GOP
    cfg = hevcr.preset(2 if k == 0 else 5)
    cfg.width, cfg.height, cfg.bit_depth = spec.width, spec.height, spec.bit_depth
    if spec.bit_depth > 8:
        cfg.level_mul, cfg.qp_min = 4, 17
    if k:
        cfg.num_refs = cfg.num_refs_before = min(k, MAX_REFS)
        cfg.poc = t
    return cfg


def frame_seed(spec: StreamSpec, t: int) -> int:
    return (spec.seed * 0x9E3779B1 + t * 0x85EBCA77 + 1) & (2**63 - 1)


def out_slot(t: int) -> int:
    return t % DPB_SLOTS


def slot_of(spec: StreamSpec, t: int) -> int:
    """DPB slot frame t of a stream is written to."""
    return spec.source[t][1] if spec.source else out_slot(t)


def bitstream_source(path: str, slots: int = 8) -> tuple:
    """Parses an Annex-B bitstream with the host front end (hevcr.Parser) into
    [(SyntaxArrays, slot)] in decoding order, with picture k written to slot
    k % slots and the reference lists remapped to match (so that a picture's
    slot is rewritten only `slots` pictures later); returns (source, width,
    height, bit_depth)."""
    with open(path, "rb") as f:
        data = f.read()
    P = hevcr.Parser()
    P.feed(data)
    P.flush()
    holder = {}  # parser slot -> decoding index of the picture in it
    source, geo = [], None
    for ev in P.events():
        if ev.kind != hevcr.EVENT_DECODE:
            continue
        syn = ev.syntax.contents
        p = hevcr.FrameParams.from_buffer_copy(bytes(syn.p))
        k = ev.decode_index
        for l in range(2):
            for i in range(p.num_ref[l]):
                ref = holder[p.ref_slot[l][i]]
                if k - ref >= slots:
                    raise ValueError("a reference lies further back than the slot ring")
                p.ref_slot[l][i] = ref % slots
        holder[ev.slot] = k
        wc = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2
        hc = (p.height + (1 << p.ctb_log2) - 1) >> p.ctb_log2
        take = (lambda ptr, n, sz: C.string_at(ptr, n * sz) if n else b"")
        arrays = hevcr.SyntaxArrays(bytes(p), take(syn.cus, syn.num_cus, 16), take(syn.pus, syn.num_pus, 20),
                                    take(syn.tus, syn.num_tus, 28),
                                    np.frombuffer(take(syn.coeffs, syn.num_coeffs, 2), np.int16).copy(),
                                    take(syn.sao, wc * hc, 24), bytes(syn.scaling.contents) if syn.scaling else None)
        source.append((arrays, k % slots))
        geo = (ev.width, ev.height, ev.bit_depth)
    P.close()
    return (source,) + geo


def bitstream_specs(path: str, n: int, slots: int = 8) -> list[StreamSpec]:
    """n streams that each decode the same parsed bitstream (own DPB each)."""
    source, w, h, bd = bitstream_source(path, slots)
    return [StreamSpec(w, h, bd, s, len(source), source=source, dpb_slots=slots) for s in range(n)]


def gen_frame(spec: StreamSpec, t: int) -> hevcr.Frame:
    """Syntax of frame t with its list entries mapped to the stream's DPB slots."""
    f = hevcr.Frame(frame_cfg(spec, t), frame_seed(spec, t), with_refs=False)
    p = f.syntax.p
    for l in range(2):
        for i in range(p.num_ref[l]):
            j = p.ref_slot[l][i]            # the generator's picture index: POC t - 1 - j
            p.ref_slot[l][i] = (t - 1 - j) % DPB_SLOTS
            p.ref_poc[l][i] = t - 1 - j
    return f


def make_specs(n: int, mix: str, seed: int = 1, frames: int = GOP) -> list[StreamSpec]:
    """mix "1080p": n 1080p 8-bit streams; "cfg4": resolution uniform over
    RESOLUTIONS and 25 % 10-bit, drawn from a seeded RNG."""
    rng = random.Random(seed)
    specs = []
    for s in range(n):
        if mix == "1080p":
            w, h, bd = 1920, 1080, 8
        elif mix == "cfg4":
            w, h = RESOLUTIONS[rng.randrange(len(RESOLUTIONS))]
            bd = 10 if rng.random() < 0.25 else 8
        else:
            raise ValueError(mix)
        specs.append(StreamSpec(w, h, bd, seed * 1000 + s, frames))
    return specs


def dpb_bytes(spec: StreamSpec) -> int:
    """Device bytes of one stream's DPB as the engine lays it out (rows padded to 64 samples)."""
    b = 2 if spec.bit_depth > 8 else 1
    sy = (spec.width + 63) // 64 * 64
    sc = (spec.width // 2 + 63) // 64 * 64
    return spec.dpb_slots * b * (sy * spec.height + 2 * sc * (spec.height // 2))


def prepare_streams(engine: hevcr.Engine, specs: list[StreamSpec], workers: int = 16, log=None):
    """Generates and prepares every frame of every stream (syntax resident on
    the device); the streams must already have ids."""
    jobs = [(s, t) for s in specs for t in range(s.frames)]

    def work(job):
        s, t = job
        if s.host:  # the syntax already in host memory (host_syntax)
            return engine.prepare(s.sid, s.host[t].syntax)
        if s.source:
            return engine.prepare(s.sid, s.source[t][0].syntax)
        f = gen_frame(s, t)
        return engine.prepare(s.sid, f.syntax)

    t0 = time.time()
    with ThreadPoolExecutor(workers) as ex:
        done = list(ex.map(work, jobs))
    for (s, t), pf in zip(jobs, done):
        s.prepared.append(pf)
    if log:
        log(f"prepared {len(jobs)} frames in {time.time() - t0:.1f} s")


def host_syntax(specs: list[StreamSpec], compact: bool, workers: int = 16, log=None):
    """Every frame's syntax in host memory, as a parser hands it over (the
    generated or parsed records; with compact, coefficients in the compact
    form): the input of serving from syntax. Not timed."""
    jobs = [(s, t) for s in specs for t in range(s.frames)]

    def work(job):
        s, t = job
        src = s.source[t][0] if s.source else gen_frame(s, t)
        return hevcr.CompactSyntax(src) if compact else src

    t0 = time.time()
    with ThreadPoolExecutor(workers) as ex:
        done = list(ex.map(work, jobs))
    for s in specs:
        s.host = []
    for (s, t), h in zip(jobs, done):
        s.host.append(h)
    if log:
        log(f"host syntax of {len(jobs)} frames in {time.time() - t0:.1f} s (compact={compact})")


def poisson_arrivals(specs: list[StreamSpec], rate_fps: float, seed: int = 7):
    """Each stream's frames arrive as a Poisson process of rate rate_fps / n
    (together a Poisson process of rate rate_fps)."""
    rng = random.Random(seed)
    lam = rate_fps / len(specs)
    for s in specs:
        t = 0.0
        s.arrivals = []
        for _ in range(s.frames):
            t += rng.expovariate(lam)
            s.arrivals.append(t)


def serve(server: hevcr.Server, specs: list[StreamSpec], saturate: bool, on_done=None, close: bool = True,
          ingest: bool = False):
    """Runs every frame through the server. saturate: all frames arrive at
    the start; otherwise at spec.arrivals. With close, a stream is closed once
    its last frame has been delivered (on_done sees every completion first),
    which lets a budget-limited server admit the streams that wait. ingest:
    frames arrive as host syntax (spec.host) and the server's ingest threads
    prepare them; otherwise they arrive prepared (spec.prepared). Returns
    (completion records, wall seconds from the first arrival to the last
    completion)."""
    events = []
    for si, s in enumerate(specs):
        for t in range(s.frames):
            events.append((0.0 if saturate else s.arrivals[t], t, si))  # equal arrivals: frame t of every stream first
    events.sort()
    total = len(events)
    left = [s.frames for s in specs]
    records = []
    t0 = server.clock()
    nxt = 0
    idle = 0

    def deliver(got):
        if on_done and got:
            on_done(got)
        records.extend(got)
        for r in got:
            si = r["tag"] // 100000
            left[si] -= 1
            if left[si] == 0 and close:
                server.close_stream(specs[si].sid)

    while len(records) < total:
        now = server.clock() - t0
        while nxt < total and events[nxt][0] <= now:
            at, t, si = events[nxt]
            s = specs[si]
            if ingest:
                server.submit_syntax(s.sid, s.host[t], slot_of(s, t), tag=si * 100000 + t, t_arrival=t0 + at)
            else:
                server.submit(s.prepared[t], slot_of(s, t), tag=si * 100000 + t, t_arrival=t0 + at)
            nxt += 1
        launched = server.step()
        got = server.poll(wait=False)
        deliver(got)
        if launched or got:
            idle = 0
            continue
        if server.info()["inflight_steps"]:
            deliver(server.poll(wait=True))
        elif nxt < total:
            time.sleep(max(0.0, min(0.0005, events[nxt][0] - (server.clock() - t0))))
        else:  # every frame submitted, none in flight: waiting for ingest threads (or stalled)
            if idle == 0:
                t_idle = time.perf_counter()
            idle += 1
            if ingest:
                time.sleep(0.00005)
            if time.perf_counter() - t_idle > 30.0:
                raise RuntimeError(f"serving stalled: {server.info()}")
    t_end = max(r["t_done"] for r in records)
    return records, t_end - t0


def run_plain(engine: hevcr.Engine, specs: list[StreamSpec], after_frame=None):
    """Plain form over the same frames: one frame at a time (t-major, so each
    stream's chain is respected), launched back to back on the engine's
    stream. Returns wall seconds."""
    order = [(si, t) for t in range(max(s.frames for s in specs)) for si, s in enumerate(specs) if t < s.frames]
    engine.sync()
    t0 = time.perf_counter()
    for si, t in order:
        specs[si].prepared[t].run(slot_of(specs[si], t))
        if after_frame:
            after_frame(si, t)
    engine.sync()
    return time.perf_counter() - t0


def run_plain_ingest(engine: hevcr.Engine, specs: list[StreamSpec]):
    """The plain form from host syntax: each frame validated, prepared,
    uploaded and reconstructed by hevcr_decode_frame, one frame at a time
    (t-major order), launched back to back. Returns wall seconds."""
    order = [(si, t) for t in range(max(s.frames for s in specs)) for si, s in enumerate(specs) if t < s.frames]
    engine.sync()
    t0 = time.perf_counter()
    for si, t in order:
        s = specs[si]
        engine.decode(s.sid, s.host[t].syntax, slot_of(s, t))
    engine.sync()
    return time.perf_counter() - t0


def serve_plain(engine: hevcr.Engine, specs: list[StreamSpec], ingest: bool = False):
    """The plain form as a server: frames reconstructed one at a time, in
    arrival order (spec.arrivals), each waited for before the next starts
    (ingest: each prepared from its host syntax first, hevcr_decode_frame).
    Returns (latencies in seconds, wall seconds)."""
    events = sorted((s.arrivals[t], si, t) for si, s in enumerate(specs) for t in range(s.frames))
    engine.sync()
    t0 = time.perf_counter()
    lat = []
    for at, si, t in events:
        now = time.perf_counter() - t0
        if now < at:
            time.sleep(at - now)
        if ingest:
            engine.decode(specs[si].sid, specs[si].host[t].syntax, slot_of(specs[si], t))
        else:
            specs[si].prepared[t].run(slot_of(specs[si], t))
        engine.sync()
        lat.append(time.perf_counter() - t0 - at)
    return lat, time.perf_counter() - t0


def percentile(values: list, q: float) -> float:
    v = sorted(values)
    if not v:
        return float("nan")
    k = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[k]
