"""SHA-256 per plane on the GPU (hevcr_picture_sha256) against Python's
hashlib (FIPS 180-4) over the same write_yuv bytes (picture.cpp:42-99):
random pictures of 8 and 10 bits with random conformance-window crops, sizes
whose plane lengths end at every padding case (a partial last block with
room for the length, one without room, none), and the reference decoder's
own per-plane hashes of every picture of its conformance bitstreams."""
import hashlib

import numpy as np
import pytest

import hevcr
import stream_goldens
from test_gpu_egress import bitstream, decode_outputs

pytestmark = pytest.mark.gpu
STREAMS = stream_goldens.supported()


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _yuv_plane_bytes(pic: hevcr.Picture, c: int, crop) -> bytes:
    l, r, t, b = crop
    s = 2 if c else 1
    p = pic.planes[c][t // s:pic.planes[c].shape[0] - b // s, l // s:pic.planes[c].shape[1] - r // s]
    return p.astype(np.uint8).tobytes() if pic.bit_depth <= 8 else p.astype("<u2").tobytes()


def test_random_pictures_and_crops_match_hashlib(engine):
    rng = np.random.default_rng(20261204)
    sizes = [(8, 8), (16, 8), (24, 16), (64, 64), (200, 136), (176, 144), (416, 240), (1920, 1080), (8, 1000)]
    cases = []
    for w, h in sizes:
        for bd in (8, 10):
            for _ in range(3):
                crop = [0, 0, 0, 0]
                if rng.random() < 0.7:
                    crop = [2 * int(rng.integers(0, max(1, w // 8))), 2 * int(rng.integers(0, max(1, w // 8))),
                            2 * int(rng.integers(0, max(1, h // 8))), 2 * int(rng.integers(0, max(1, h // 8)))]
                cases.append((w, h, bd, crop))
    lengths = set()
    sids, pics, want, refs, crops = [], [], [], [], []
    try:
        for w, h, bd, crop in cases:
            sid = engine.stream(w, h, bd, 1)
            sids.append(sid)
            pic = hevcr.Picture(w, h, bd)
            for p in pic.planes:
                p[...] = rng.integers(0, 1 << bd, p.shape)
            engine.upload(sid, 0, pic)
            planes = [_yuv_plane_bytes(pic, c, crop) for c in range(3)]
            lengths |= {len(b) % 64 for b in planes}
            want.append([hashlib.sha256(b).hexdigest() for b in planes])
            refs.append((sid, 0))
            crops.append(crop)
        print(f"{len(refs)} pictures in one call", flush=True)
        assert engine.sha256(refs, crops) == want
        assert 0 in lengths and any(0 < x < 56 for x in lengths) and any(x >= 56 for x in lengths)
        print("one call per picture, no crop", flush=True)
        for (sid, slot), (w, h, bd, _) in zip(refs[:12], cases[:12]):
            pic = engine.download(sid, slot, w, h, bd)
            assert engine.sha256([(sid, slot)]) == [pic.sha256()]
    finally:
        for sid in sids:
            engine.destroy_stream(sid)


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_stream_pictures_match_reference_hashes(engine, stream):
    """Every picture of the bitstream, decoded by the engine, hashed on the
    device: the uncropped planes give the reference decoder's per-plane
    hashes; the cropped write_yuv bytes, chained, give its YUV MD5 (checked
    through hashlib on the downloaded bytes of the same crops)."""
    print(f"stream {stream['name']}", flush=True)
    frames = stream["frames"]
    got, md5 = [], hashlib.md5()

    def on_output(sid, ev):
        crop = list(ev.crop)
        dev = engine.sha256([(sid, ev.slot), (sid, ev.slot)], [[0, 0, 0, 0], crop])
        got.append((ev.poc, dev[0]))
        yuv = engine.egress([(sid, ev.slot, ev.width, ev.height, ev.bit_depth)], crops=[crop])
        md5.update(yuv.tobytes())
        off = 0
        for c in range(3):
            s = 2 if c else 1
            w = (ev.width - crop[0] - crop[1]) // s
            h = (ev.height - crop[2] - crop[3]) // s
            nb = w * h * (1 if ev.bit_depth <= 8 else 2)
            assert dev[1][c] == hashlib.sha256(yuv[off:off + nb].tobytes()).hexdigest()
            off += nb
    decode_outputs(engine, bitstream(stream["name"]), on_output)
    assert md5.hexdigest() == stream["md5_yuv_reference"]
    assert len(got) == len(frames)
    order = sorted(frames, key=lambda f: (f["cvs"], f["poc"]))  # output order: POC within each sequence
    for (poc, planes), f in zip(got, order):
        assert poc == f["poc"] and planes == f["sha256"]["sao"]
