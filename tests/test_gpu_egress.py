"""Egress on the GPU (hevcr_picture_egress): every conformance bitstream of
the reference, parsed by the host front end and reconstructed by the engine,
serialised on the device in write_yuv layout (picture.cpp:42-99: conformance
window, 8-bit samples as bytes, deeper ones as 16-bit little-endian) gives the
reference decoder's YUV MD5, whether the bytes land in ordinary host memory
(written to device memory, then copied) or in pinned host memory (written by
the GPU directly); the uint16 plane layout equals hevcr_picture_download; one
call over many pictures equals one call per picture."""
import hashlib
import os

import numpy as np
import pytest

import hevcr
import stream_goldens

pytestmark = pytest.mark.gpu
STREAMS = stream_goldens.supported()
BITSTREAMS = os.path.join(stream_goldens.GOLDEN_DIR, "bitstreams")


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def bitstream(name):
    with open(os.path.join(BITSTREAMS, name + ".265"), "rb") as f:
        return f.read()


def decode_outputs(engine, data, on_output):
    """Parser events into the engine; on_output(sid, event) at each OUTPUT
    event, while its slot still holds the picture."""
    P = hevcr.Parser()
    sid, seq = None, None
    try:
        P.feed(data)
        P.flush()
        for ev in P.events():
            if ev.sequence != seq:
                if sid is not None:
                    engine.destroy_stream(sid)
                sid, seq = engine.stream(ev.width, ev.height, ev.bit_depth, ev.dpb_slots), ev.sequence
            if ev.kind == hevcr.EVENT_DECODE:
                engine.decode(sid, ev.syntax.contents, ev.slot)
            else:
                on_output(sid, ev)
    finally:
        P.close()
        if sid is not None:
            engine.destroy_stream(sid)


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_egress_yuv_gives_reference_md5(engine, stream):
    print(f"stream {stream['name']}: egress to pageable and pinned memory", flush=True)
    md5 = {k: hashlib.md5() for k in ("pageable", "pinned")}
    pinned = {}

    def on_output(sid, ev):
        pic = [(sid, ev.slot, ev.width, ev.height, ev.bit_depth)]
        crop = [list(ev.crop)]
        n = hevcr.egress_size(ev.width, ev.height, ev.bit_depth, hevcr.EGRESS_YUV, crop[0])
        md5["pageable"].update(engine.egress(pic, hevcr.EGRESS_YUV, crop).tobytes())
        if pinned.get("buf") is None or pinned["buf"].nbytes < n:
            if pinned.get("buf") is not None:
                pinned["buf"].close()
            pinned["buf"] = hevcr.PinnedBuffer(n)
        md5["pinned"].update(engine.egress(pic, hevcr.EGRESS_YUV, crop, out=pinned["buf"]).tobytes())

    try:
        decode_outputs(engine, bitstream(stream["name"]), on_output)
    finally:
        if pinned.get("buf") is not None:
            pinned["buf"].close()
    for where, h in md5.items():
        assert h.hexdigest() == stream["md5_yuv_reference"], where
        if stream["md5_yuv_readme"]:
            assert h.hexdigest() == stream["md5_yuv_readme"], where


@pytest.mark.parametrize("name", ["full_qcif_10f_10bit", "b_qcif_10f"])
def test_egress_u16_equals_download(engine, name):
    print(f"stream {name}: uint16 planes", flush=True)
    seen = []

    def on_output(sid, ev):
        down = engine.download(sid, ev.slot, ev.width, ev.height, ev.bit_depth)
        raw = engine.egress([(sid, ev.slot, ev.width, ev.height, ev.bit_depth)], hevcr.EGRESS_U16)
        u16 = raw.view("<u2")
        ny, nc = ev.width * ev.height, (ev.width // 2) * (ev.height // 2)
        assert np.array_equal(u16[:ny].reshape(ev.height, ev.width), down.planes[0])
        assert np.array_equal(u16[ny:ny + nc].reshape(ev.height // 2, ev.width // 2), down.planes[1])
        assert np.array_equal(u16[ny + nc:].reshape(ev.height // 2, ev.width // 2), down.planes[2])
        seen.append(ev.poc)

    decode_outputs(engine, bitstream(name), on_output)
    assert len(seen) == 10


def test_egress_many_pictures_in_one_call(engine):
    """Three streams of different sizes and bit depths, their slots egressed
    together (crops differ per picture) equal one call per picture."""
    pics, crops = [], []
    sids = []
    for w, h, bd in ((416, 240, 8), (176, 144, 10), (64, 64, 8)):
        sid = engine.stream(w, h, bd, 2)
        sids.append(sid)
        rng = np.random.default_rng(w + bd)
        for slot in range(2):
            p = hevcr.Picture(w, h, bd)
            for c in range(3):
                p.planes[c][:] = rng.integers(0, 1 << bd, p.planes[c].shape, dtype=np.uint16)
            engine.upload(sid, slot, p)
            pics.append((sid, slot, w, h, bd))
            crops.append([2 * slot, 2, 0, 4 * slot])
    try:
        for fmt in (hevcr.EGRESS_YUV, hevcr.EGRESS_U16):
            together = engine.egress(pics, fmt, crops).tobytes()
            one_by_one = b"".join(engine.egress([p], fmt, [c]).tobytes() for p, c in zip(pics, crops))
            assert together == one_by_one
        # the write_yuv bytes of one uploaded picture, computed on the host
        sid, slot, w, h, bd = pics[3]
        ref = engine.download(sid, slot, w, h, bd)
        l, r, t, b = crops[3]
        want = b"".join(
            ref.planes[c][t // s:ref.planes[c].shape[0] - b // s, l // s:ref.planes[c].shape[1] - r // s]
            .astype("<u2" if bd > 8 else np.uint8).tobytes() for c, s in ((0, 1), (1, 2), (2, 2)))
        assert engine.egress([pics[3]], hevcr.EGRESS_YUV, [crops[3]]).tobytes() == want
    finally:
        for sid in sids:
            engine.destroy_stream(sid)
