"""Compact coefficients and serving from host syntax, on the GPU.

The compact coefficient form (the residual kernels build each block in LDS
from its list of nonzero coefficients) gives the reference's golden planes
for every golden case and stage, every picture of the reference's
conformance streams, and batches that mix compact and dense frames. The
server's ingest threads prepare frames submitted as syntax while steps run:
every frame equals the oracle's decode of its stream, dense and compact
submissions give the same pictures, a frame whose syntax is invalid
completes with the error and leaves the others alone, and the ingest
counters account for every frame."""
import pytest

import hevcr
import oracle_lib
import serving
import stream_goldens

pytestmark = pytest.mark.gpu
CASES = oracle_lib.load_goldens()
STOPS = ((hevcr.RUN_STOP_RECON, "recon"), (hevcr.RUN_STOP_DEBLOCK, "deblock"), (hevcr.RUN_FULL, "sao"))
STREAMS = stream_goldens.supported()


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_compact_frames_match_goldens(engine, case):
    f = oracle_lib.make_frame(case)
    comp = hevcr.CompactSyntax(f)
    for stop, stage in STOPS:
        print(f"case {case['name']} stage {stage}, compact", flush=True)
        assert hevcr.decode_on_engine(engine, f, stop, comp.syntax).sha256() == case["stages"][stage], stage


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_compact_streams_match_reference(engine, stream):
    sid = engine.stream(stream["width"], stream["height"], stream["bit_depth"], stream["dpb_slots"])
    try:
        for k, (fr, arrays) in enumerate(stream_goldens.frames(stream)):
            print(f"stream {stream['name']} picture {k}, compact", flush=True)
            comp = hevcr.CompactSyntax(arrays)
            engine.decode(sid, comp.syntax, fr["slot"])
            assert engine.sha256([(sid, fr["slot"])])[0] == fr["sha256"]["sao"], f"picture {k}"
    finally:
        engine.destroy_stream(sid)


def test_batch_mixing_compact_and_dense_frames(engine):
    names = ["intra416_s1", "b640x360_10bit", "intra200x136_ctb16_10bit", "b1080_s1", "intra416_scaling"]
    by = {c["name"]: c for c in CASES}
    sids, prepared, slots = [], [], []
    try:
        for i, name in enumerate(names):
            f = oracle_lib.make_frame(by[name])
            k = f.num_ref_pics
            sid = engine.stream(f.width, f.height, f.bit_depth, k + 1)
            sids.append(sid)
            for r in range(k):
                engine.upload(sid, r, f.refs[r])
            syn = hevcr.CompactSyntax(f).syntax if i % 2 == 0 else f.syntax
            prepared.append((engine.prepare(sid, syn), f))
            slots.append(k)
        b = engine.batch([p for p, _ in prepared], slots)
        for rep in range(2):
            print(f"mixed batch replay {rep}", flush=True)
            b.run(True)
            engine.sync()
            got = engine.sha256(list(zip(sids, slots)))
            assert got == [by[n]["stages"]["sao"] for n in names]
        b.release()
    finally:
        for p, _ in prepared:
            p.release()
        for sid in sids:
            engine.destroy_stream(sid)


SMALL = [(416, 240, 8), (192, 128, 10), (256, 144, 8), (640, 360, 8)]


def _oracle_chain(spec):
    dpb = [hevcr.Picture(spec.width, spec.height, spec.bit_depth) for _ in range(serving.DPB_SLOTS)]
    out = []
    for t in range(spec.frames):
        f = serving.gen_frame(spec, t)
        f.refs, f.num_ref_pics = dpb, serving.DPB_SLOTS
        pic = oracle_lib.recon(f)[2]
        dpb[serving.out_slot(t)] = pic
        out.append(pic.sha256())
    return out


def _serve_syntax(engine, specs, compact, threads=4, max_inflight=2, poisson=None):
    got = {}
    with hevcr.Server(engine, 1 << 30, max_batch=64, max_inflight=max_inflight, ingest_threads=threads) as srv:
        for s in specs:
            s.sid = srv.open(s.width, s.height, s.bit_depth, serving.DPB_SLOTS)
        serving.host_syntax(specs, compact)
        if poisson:
            serving.poisson_arrivals(specs, poisson, seed=5)

        def on_done(recs):
            assert all(r["status"] == 0 for r in recs)
            hashes = engine.sha256([(specs[r["tag"] // 100000].sid, r["out_slot"]) for r in recs])
            for r, h in zip(recs, hashes):
                got[divmod(r["tag"], 100000)] = h
        serving.serve(srv, specs, poisson is None, on_done, ingest=True)
        info = srv.info()
    total = sum(s.frames for s in specs)
    assert info["ingested_frames"] == total and info["ingest_bytes"] > 0 and info["ingest_busy_s"] > 0
    return got


@pytest.mark.parametrize("compact", [False, True], ids=["dense", "compact"])
def test_served_from_syntax_matches_oracle(engine, compact):
    specs = [serving.StreamSpec(w, h, bd, seed=41 + k, frames=6) for k, (w, h, bd) in enumerate(SMALL)]
    want = {si: _oracle_chain(s) for si, s in enumerate(specs)}
    print(f"serving from syntax, compact={compact}", flush=True)
    got = _serve_syntax(engine, specs, compact)
    for si, s in enumerate(specs):
        for t in range(s.frames):
            assert got[(si, t)] == want[si][t], f"stream {si} frame {t}"


def test_served_from_syntax_poisson_equals_prepared(engine):
    mk = lambda: [serving.StreamSpec(w, h, bd, seed=61 + k, frames=8) for k, (w, h, bd) in enumerate(SMALL)]
    print("serving prepared frames", flush=True)
    ref = {}
    with hevcr.Server(engine, 1 << 30, max_batch=64, max_inflight=2) as srv:
        specs = mk()
        for s in specs:
            s.sid = srv.open(s.width, s.height, s.bit_depth, serving.DPB_SLOTS)
        serving.prepare_streams(engine, specs, workers=4)

        def on_done(recs):
            for r, h in zip(recs, engine.sha256([(specs[r["tag"] // 100000].sid, r["out_slot"]) for r in recs])):
                ref[divmod(r["tag"], 100000)] = h
        serving.serve(srv, specs, True, on_done)
        for s in specs:
            for p in s.prepared:
                p.release()
    print("serving syntax with 8 ingest threads, Poisson arrivals", flush=True)
    assert _serve_syntax(engine, mk(), True, threads=8, max_inflight=3, poisson=800.0) == ref


def test_invalid_syntax_completes_with_its_error(engine):
    specs = [serving.StreamSpec(256, 144, 8, seed=71 + k, frames=4) for k in range(2)]
    with hevcr.Server(engine, 1 << 30, max_batch=64, max_inflight=2, ingest_threads=2) as srv:
        for s in specs:
            s.sid = srv.open(s.width, s.height, s.bit_depth, serving.DPB_SLOTS)
        serving.host_syntax(specs, True)
        bad = specs[1].host[2]
        bad.syntax.p.ctb_log2 = 7  # rejected by validation on the ingest thread
        recs = []
        for si, s in enumerate(specs):
            for t in range(s.frames):
                srv.submit_syntax(s.sid, s.host[t], serving.out_slot(t), si * 100000 + t)
        while len(recs) < 8:
            srv.step()
            recs += srv.poll(wait=True)
        status = {divmod(r["tag"], 100000): r["status"] for r in recs}
        assert status[(1, 2)] == -3  # HEVCR_EINVAL
        assert all(v == 0 for k, v in status.items() if k != (1, 2))
        assert srv.info()["ingested_frames"] == 7


@pytest.mark.parametrize("name", ["intra416_s1", "b640x360_10bit", "intra416_slices_tiles_pcm"])
def test_compact_positions_taken_modulo_the_block(engine, name):
    """hevcr_syntax.h: the engine checks a compact list's count and extent,
    not its positions; a position is taken modulo the block size. Moving
    every position up by 4096 (a multiple of every block's N*N) and reversing
    each list's order give the same picture as the well-formed lists."""
    import numpy as np
    case = next(c for c in CASES if c["name"] == name)
    f = oracle_lib.make_frame(case)
    comp = hevcr.CompactSyntax(f)
    want = hevcr.decode_on_engine(engine, f, hevcr.RUN_FULL, comp.syntax).sha256()
    s = comp.syntax
    sig = comp.sig
    heads = set()
    offs = comp.tus[:, 16:28].copy().view(np.uint32).reshape(-1, 3)
    for i in range(s.num_tus):
        t = s.tus[i]
        for c in range(3):
            if (t.cbf >> c) & 1 and (c == 0 or t.log2_size > 2 or t.blk_idx == 3):
                heads.add(int(offs[i, c]))
    orig = sig.copy()
    try:
        for h in sorted(heads):
            k = int(orig[h])
            words = orig[h + 1:h + 1 + k].copy()
            sig[h + 1:h + 1 + k] = (words + 4096)[::-1]
        print(f"case {name}: {len(heads)} lists moved and reversed", flush=True)
        got = hevcr.decode_on_engine(engine, f, hevcr.RUN_FULL, comp.syntax).sha256()
    finally:
        sig[:] = orig
    assert got == want
