"""The HIP engine on real bitstreams: every picture of the reference's
conformance streams (x265 encodes, tests/golden/streams), reconstructed on
the GPU from the syntax the reference decoder's parser produced, with its
reference pictures resident in the stream's DPB slots in HBM, matches the
reference decoder's SHA-256 per plane; each whole stream matches the MD5 of
ffmpeg's decode given in the fixtures' README. The batched form runs picture
k of every stream together (mixed sizes and bit depths, hipGraph replay)."""
import numpy as np
import pytest

import hevcr
import oracle_lib
import stream_goldens

pytestmark = pytest.mark.gpu
STREAMS = stream_goldens.supported()
STOPS = ((hevcr.RUN_STOP_RECON, "recon"), (hevcr.RUN_STOP_DEBLOCK, "deblock"))


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _diagnose(engine, sid, stream, k, fr, arrays, dpb):
    """First stage whose hash differs, with the first differing sample against the oracle."""
    oracle = oracle_lib.recon_syntax(arrays.syntax, dpb)
    for idx, (stop, st) in enumerate(STOPS):
        pf = engine.prepare(sid, arrays.syntax)
        pf.run(fr["slot"], stop)
        pf.release()
        engine.sync()
        got = engine.download(sid, fr["slot"], stream["width"], stream["height"], stream["bit_depth"])
        if got.sha256() != fr["sha256"][st]:
            break
    else:
        st, idx, got = "sao", 2, None
    msg = f"picture {k} (POC {fr['poc']}) first differs after {st}"
    if got is not None:
        for c in range(3):
            d = np.argwhere(got.planes[c] != oracle[idx].planes[c])
            if len(d):
                y, x = d[0]
                msg += f"; plane {c}: {len(d)} samples, first (x={x}, y={y})"
    return msg


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_engine_decodes_reference_streams(engine, stream):
    w, h, bd = stream["width"], stream["height"], stream["bit_depth"]
    sid = engine.stream(w, h, bd, stream["dpb_slots"])
    dpb = [None] * stream["dpb_slots"]
    outputs = []
    try:
        for k, (fr, arrays) in enumerate(stream_goldens.frames(stream)):
            print(f"stream {stream['name']} picture {k} POC {fr['poc']}", flush=True)
            engine.decode(sid, arrays.syntax, fr["slot"])
            engine.sync()
            pic = engine.download(sid, fr["slot"], w, h, bd)
            if pic.sha256() != fr["sha256"]["sao"]:
                pytest.fail(_diagnose(engine, sid, stream, k, fr, arrays, dpb))
            dpb[fr["slot"]] = pic
            outputs.append(pic)
    finally:
        engine.destroy_stream(sid)
    md5 = stream_goldens.yuv_md5(stream, outputs)
    assert md5 == stream["md5_yuv_reference"]
    if stream["md5_yuv_readme"]:
        assert md5 == stream["md5_yuv_readme"]


def test_engine_batches_pictures_across_streams(engine):
    """Picture k of every stream in one batch, captured in a hipGraph and
    replayed; every picture matches its reference hash."""
    sids = [engine.stream(s["width"], s["height"], s["bit_depth"], s["dpb_slots"]) for s in STREAMS]
    iters = [stream_goldens.frames(s) for s in STREAMS]
    try:
        for k in range(max(len(s["frames"]) for s in STREAMS)):
            print(f"batch of picture {k}", flush=True)
            members, keep = [], []
            for s, sid, it in zip(STREAMS, sids, iters):
                if k < len(s["frames"]):
                    fr, arrays = next(it)
                    keep.append(arrays)
                    members.append((s, sid, fr, engine.prepare(sid, arrays.syntax)))
            b = engine.batch([m[3] for m in members], [m[2]["slot"] for m in members])
            b.run(graph=True)
            engine.sync()
            for s, sid, fr, pf in members:
                pic = engine.download(sid, fr["slot"], s["width"], s["height"], s["bit_depth"])
                assert pic.sha256() == fr["sha256"]["sao"], f"{s['name']} picture {k}"
            b.release()
            for m in members:
                m[3].release()
    finally:
        for sid in sids:
            engine.destroy_stream(sid)
