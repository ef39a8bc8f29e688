"""The coded-stream golden cases (tests/golden/coded/: slice segments,
dependent segments, tiles, PCM, transquant bypass, weighted prediction per
segment and the other picture and slice options) on the GPU:
  - through the reference's decoder API (hevc_api.h, libhevcr_decoder.so:
    host front end + HIP engine), each stream decodes to the YUV whose MD5
    the reference decoder's CLI gives;
  - through the engine directly, each picture the front end parses is
    reconstructed equal to the oracle's picture from the same syntax, in the
    dense and the compact coefficient form, and picture k of every stream
    together in one batch replayed from a hipGraph."""
import ctypes as C
import hashlib

import pytest

import coded_streams as cs
import hevcr
import oracle_lib

pytestmark = pytest.mark.gpu
STREAMS = cs.load()


def frame_bytes(f: hevcr.HEVCFrame) -> bytes:
    out = b""
    for p in hevcr.frame_planes(f):
        out += p.astype("uint8").tobytes() if f.bit_depth <= 8 else p.astype("<u2").tobytes()
    return out


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_decoder_api_reproduces_reference_md5(stream):
    L = hevcr.decoder_lib()
    print(f"stream {stream['name']}: hevc_decoder_decode", flush=True)
    d = L.hevc_decoder_create()
    try:
        data = cs.read(stream["name"])
        rc = L.hevc_decoder_decode(d, data, len(data))
        assert rc == 0, L.hevcr_decoder_error(d)
        n = L.hevc_decoder_get_frame_count(d)
        assert n == stream["pictures"]
        h = hashlib.md5()
        f = hevcr.HEVCFrame()
        for i in range(n):
            assert L.hevc_decoder_get_frame(d, i, C.byref(f)) == 0
            h.update(frame_bytes(f))
        assert h.hexdigest() == stream["md5_yuv_reference"]
    finally:
        L.hevc_decoder_destroy(d)


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def parsed_pictures(stream):
    """[(slot, SyntaxArrays, oracle picture)] in decoding order."""
    evs = cs.parse(cs.read(stream["name"]))
    slots = max(e["dpb_slots"] for e, _ in evs)
    dpb = [None] * slots
    out = []
    for e, syn in evs:
        if e["kind"] == hevcr.EVENT_DECODE:
            pic = oracle_lib.recon_syntax(syn.syntax, dpb)[2]
            dpb[e["slot"]] = pic
            out.append((e["slot"], syn, pic))
    return slots, out


@pytest.mark.parametrize("compact", [False, True], ids=["dense", "compact"])
@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_engine_equals_oracle_on_parsed_pictures(engine, stream, compact):
    w, h, bd = stream["width"], stream["height"], stream["bit_depth"]
    slots, pics = parsed_pictures(stream)
    sid = engine.stream(w, h, bd, slots)
    try:
        for k, (slot, syn, want) in enumerate(pics):
            print(f"stream {stream['name']} picture {k} {'compact' if compact else 'dense'}", flush=True)
            src = hevcr.CompactSyntax(syn) if compact else syn
            engine.decode(sid, src.syntax, slot)
            engine.sync()
            got = engine.download(sid, slot, w, h, bd)
            assert got.sha256() == want.sha256(), f"picture {k}"
    finally:
        engine.destroy_stream(sid)


def test_engine_batches_parsed_pictures_across_streams(engine):
    parsed = [parsed_pictures(s) for s in STREAMS]
    sids = [engine.stream(s["width"], s["height"], s["bit_depth"], p[0]) for s, p in zip(STREAMS, parsed)]
    try:
        for k in range(max(len(p[1]) for p in parsed)):
            print(f"batch of picture {k}", flush=True)
            members = []
            for s, sid, (_, pics) in zip(STREAMS, sids, parsed):
                if k < len(pics):
                    slot, syn, want = pics[k]
                    members.append((s, sid, slot, want, engine.prepare(sid, syn.syntax)))
            b = engine.batch([m[4] for m in members], [m[2] for m in members])
            b.run(graph=True)
            engine.sync()
            for s, sid, slot, want, pf in members:
                got = engine.download(sid, slot, s["width"], s["height"], s["bit_depth"])
                assert got.sha256() == want.sha256(), f"{s['name']} picture {k}"
            b.release()
            for m in members:
                m[4].release()
    finally:
        for sid in sids:
            engine.destroy_stream(sid)
