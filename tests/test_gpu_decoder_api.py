"""End to end on the GPU through the reference's own decoder API
(hevc_api.h, served by libhevcr_decoder.so: host front end + HIP engine):
every conformance bitstream of the reference decodes to the YUV whose MD5
the fixtures' README gives for ffmpeg's decode (and the reference decoder's
own), whole-stream and incrementally (feed in chunks, drain, flush)."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import hevcr
import stream_goldens

pytestmark = pytest.mark.gpu
STREAMS = stream_goldens.supported()
BITSTREAMS = os.path.join(stream_goldens.GOLDEN_DIR, "bitstreams")


def bitstream(name):
    with open(os.path.join(BITSTREAMS, name + ".265"), "rb") as f:
        return f.read()


def frame_bytes(f: hevcr.HEVCFrame) -> bytes:
    """write_yuv serialisation (picture.cpp:42-99): 1 byte per sample up to 8 bits, uint16 LE above."""
    out = b""
    for p in hevcr.frame_planes(f):
        out += p.astype(np.uint8).tobytes() if f.bit_depth <= 8 else p.astype("<u2").tobytes()
    return out


def check_md5(stream, md5):
    assert md5 == stream["md5_yuv_reference"]
    if stream["md5_yuv_readme"]:
        assert md5 == stream["md5_yuv_readme"]


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_decode_whole_bitstream(stream):
    L = hevcr.decoder_lib()
    print(f"stream {stream['name']}: hevc_decoder_decode", flush=True)
    d = L.hevc_decoder_create()
    try:
        data = bitstream(stream["name"])
        rc = L.hevc_decoder_decode(d, data, len(data))
        assert rc == 0, L.hevcr_decoder_error(d)
        n = L.hevc_decoder_get_frame_count(d)
        assert n == len(stream["frames"])
        h = hashlib.md5()
        f = hevcr.HEVCFrame()
        for i in range(n):
            assert L.hevc_decoder_get_frame(d, i, C.byref(f)) == 0
            h.update(frame_bytes(f))
        check_md5(stream, h.hexdigest())
        info = hevcr.HEVCStreamInfo()
        assert L.hevc_decoder_get_info(d, C.byref(info)) == 0
        cl, cr, ct, cb = stream["crop"]
        assert (info.width, info.height, info.bit_depth, info.chroma_format) == (
            stream["width"] - cl - cr, stream["height"] - ct - cb, stream["bit_depth"], 1)
    finally:
        L.hevc_decoder_destroy(d)


@pytest.mark.parametrize("name", ["b_qcif_10f", "full_qcif_10f_10bit", "bbb1080_50f"])
def test_decode_incrementally(name):
    stream = next(s for s in STREAMS if s["name"] == name)
    L = hevcr.decoder_lib()
    print(f"stream {name}: feed / drain / flush", flush=True)
    d = L.hevc_decoder_create()
    try:
        data = bitstream(name)
        h = hashlib.md5()
        f = hevcr.HEVCFrame()
        cnt = C.c_int()
        total = 0
        for i in range(0, len(data), 4096):
            chunk = data[i:i + 4096]
            assert L.hevc_decoder_feed(d, chunk, len(chunk)) == 0, L.hevcr_decoder_error(d)
            assert L.hevc_decoder_drain(d, C.byref(cnt)) == 0
            for k in range(cnt.value):
                assert L.hevc_decoder_get_drained_frame(d, k, C.byref(f)) == 0
                h.update(frame_bytes(f))
            total += cnt.value
        assert L.hevc_decoder_flush(d) == 0, L.hevcr_decoder_error(d)
        # the flushed pictures are readable as drained frames right after the flush (hevc_api.cpp:143-151)
        k = 0
        while L.hevc_decoder_get_drained_frame(d, k, C.byref(f)) == 0:
            h.update(frame_bytes(f))
            k += 1
        total += k
        assert total == len(stream["frames"])
        check_md5(stream, h.hexdigest())
    finally:
        L.hevc_decoder_destroy(d)


def _decode_md5(L, d, name):
    data = bitstream(name)
    rc = L.hevc_decoder_decode(d, data, len(data))
    assert rc == 0, L.hevcr_decoder_error(d)
    h = hashlib.md5()
    f = hevcr.HEVCFrame()
    for i in range(L.hevc_decoder_get_frame_count(d)):
        assert L.hevc_decoder_get_frame(d, i, C.byref(f)) == 0
        h.update(frame_bytes(f))
    return h.hexdigest()


def test_get_info_follows_the_last_decode():
    """hevc_api.cpp:153-166: stream info comes from the first picture of the
    last hevc_decoder_decode call (HEVC_ERROR before any), with profile and
    level 0; a second stream of another size on the same handle reports its
    own size."""
    L = hevcr.decoder_lib()
    info = hevcr.HEVCStreamInfo()
    d = L.hevc_decoder_create()
    try:  # incremental use fills no batch output
        data = bitstream("i_64x64_full")
        assert L.hevc_decoder_feed(d, data, len(data)) == 0
        assert L.hevc_decoder_flush(d) == 0
        assert L.hevc_decoder_get_info(d, C.byref(info)) == -1
    finally:
        L.hevc_decoder_destroy(d)
    d = L.hevc_decoder_create()
    try:
        assert L.hevc_decoder_get_info(d, C.byref(info)) == -1
        for name in ("i_64x64_full", "i_qcif_176x144", "bbb1080_50f"):
            stream = next(s for s in STREAMS if s["name"] == name)
            check_md5(stream, _decode_md5(L, d, name))
            assert L.hevc_decoder_get_info(d, C.byref(info)) == 0
            cl, cr, ct, cb = stream["crop"]
            assert (info.width, info.height, info.profile, info.level) == (
                stream["width"] - cl - cr, stream["height"] - ct - cb, 0, 0)
    finally:
        L.hevc_decoder_destroy(d)


def test_decoders_share_one_engine():
    """Several live decoders (one engine underneath) interleave their calls
    and each still decodes its own stream exactly; destroying one leaves the
    others working."""
    L = hevcr.decoder_lib()
    names = ["b_qcif_10f", "full_qcif_10f_10bit", "p_qcif_10f"]
    decs = [L.hevc_decoder_create() for _ in names]
    try:
        for rnd in range(2):
            for d, name in zip(decs, names):
                print(f"round {rnd}: {name}", flush=True)
                stream = next(s for s in STREAMS if s["name"] == name)
                check_md5(stream, _decode_md5(L, d, name))
            L.hevc_decoder_destroy(decs.pop(0))
            names.pop(0)
    finally:
        for d in decs:
            L.hevc_decoder_destroy(d)
