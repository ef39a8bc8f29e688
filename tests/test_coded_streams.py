"""The host front end on the coded-stream golden cases (tests/golden/coded/,
written by tests/golden/stream_writer.cpp): pictures of several slice
segments (dependent ones included), tiles with and without loop filtering
across them, wavefront entry points, PCM with and without loop filtering,
transquant bypass, weighted prediction per segment, scaling lists, reference
list modification, parallel merge levels, a conformance window. Bitstream ->
front end -> CPU oracle gives the YUV whose MD5 the reference decoder's CLI
gives for the same stream; feeding the stream in small pieces changes
nothing; a picture that misses a slice segment is an error."""
import ctypes as C
import hashlib

import numpy as np
import pytest

import coded_streams as cs
import hevcr
import oracle_lib

STREAMS = cs.load()


def oracle_md5(data: bytes) -> tuple[str, int]:
    evs = cs.parse(data)
    dpb = [None] * max(e["dpb_slots"] for e, _ in evs)
    outs, crop, bd = [], None, 8
    for e, syn in evs:
        if e["kind"] == hevcr.EVENT_DECODE:
            dpb[e["slot"]] = oracle_lib.recon_syntax(syn.syntax, dpb)[2]
        else:
            outs.append(dpb[e["slot"]].planes)
            crop, bd = e["crop"], e["bit_depth"]
    return cs.md5_of_outputs(outs, crop, bd), len(outs)


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_front_end_and_oracle_reproduce_reference_md5(stream):
    md5, n = oracle_md5(cs.read(stream["name"]))
    assert n == stream["pictures"]
    assert md5 == stream["md5_yuv_reference"]


def test_multi_segment_pictures_are_described_per_segment():
    """A picture of several segments carries one hevcr_slice per segment and
    a segment index per CTB, in the order the segments cover the picture;
    tiles and the PCM loop-filter flag are handed on."""
    seen_slices = seen_tiles = 0
    for name in ("slices_dep_p", "tiles4x2_explicit_nolf", "pcm_nolf_tiles"):
        for e, syn in cs.parse(cs.read(name)):
            if syn is None:
                continue
            s = syn.syntax
            p = s.p
            wc = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2
            hc = (p.height + (1 << p.ctb_log2) - 1) >> p.ctb_log2
            if s.num_slices:
                seen_slices += 1
                seg = np.ctypeslib.as_array(s.ctb_slice, shape=(wc * hc,))
                assert seg.max() == s.num_slices - 1 and set(seg.tolist()) == set(range(s.num_slices))
            if s.tiles:
                seen_tiles += 1
                t = s.tiles.contents
                assert sum(t.col_width[:t.cols]) == wc and sum(t.row_height[:t.rows]) == hc
                assert t.loop_filter_across_tiles == (0 if name == "tiles4x2_explicit_nolf" else 1)
            if name == "pcm_nolf_tiles":
                assert s.pcm_loop_filter_disabled == 1
    assert seen_slices and seen_tiles


@pytest.mark.parametrize("name", ["slices_types_wp", "tiles_slices_832x480"])
def test_chunked_feed_is_identical(name):
    data = cs.read(name)

    def digest(chunk):
        P = hevcr.Parser()
        h = hashlib.sha256()
        try:
            for i in range(0, len(data), chunk):
                P.feed(data[i:i + chunk])
                for ev in P.events():
                    h.update(bytes([ev.kind]) + ev.poc.to_bytes(4, "little", signed=True) + bytes([ev.slot]))
                    if ev.kind == hevcr.EVENT_DECODE:
                        h.update(syntax_bytes(ev.syntax.contents))
            P.flush()
            for ev in P.events():
                h.update(bytes([ev.kind]) + ev.poc.to_bytes(4, "little", signed=True) + bytes([ev.slot]))
        finally:
            P.close()
        return h.hexdigest()

    whole = digest(len(data))
    assert digest(997) == whole
    assert digest(64) == whole


def syntax_bytes(s: hevcr.FrameSyntax) -> bytes:
    """Every record of a picture's syntax, as bytes."""
    p = s.p
    wc = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2
    hc = (p.height + (1 << p.ctb_log2) - 1) >> p.ctb_log2
    out = [bytes(p)]
    for ptr, n, t in ((s.cus, s.num_cus, hevcr.CU), (s.pus, s.num_pus, hevcr.PU), (s.tus, s.num_tus, hevcr.TU),
                      (s.coeffs, s.num_coeffs, C.c_int16), (s.sao, wc * hc, hevcr.SAO),
                      (s.slices, s.num_slices, hevcr.Slice)):
        out.append(C.string_at(ptr, n * C.sizeof(t)) if n else b"")
    out.append(C.string_at(s.ctb_slice, wc * hc) if s.num_slices else b"")
    out.append(bytes(s.tiles.contents) if s.tiles else b"")
    out.append(bytes([s.pcm_loop_filter_disabled]))
    return b"".join(out)


def nal_units(data: bytes) -> list:
    """(start, end) of each NAL unit payload after its start code."""
    starts = [i + 3 for i in range(len(data) - 2) if data[i:i + 3] == b"\x00\x00\x01"]
    out = []
    for k, s in enumerate(starts):
        e = starts[k + 1] - 3 if k + 1 < len(starts) else len(data)
        while e > s and data[e - 1] == 0:
            e -= 1
        out.append((s, e))
    return out


def test_missing_slice_segment_is_an_error():
    data = cs.read("slices_dep_p")
    units = nal_units(data)
    vcl = [(s, e) for s, e in units if (data[s] >> 1) & 0x3f < 32]
    first_of_pic = [(s, e) for s, e in vcl if data[s + 2] & 0x80]
    # drop a segment that is not the first of its picture
    s, e = next((s, e) for s, e in vcl if (s, e) not in first_of_pic)
    cut = data[:s - 3] + data[e:]
    P = hevcr.Parser()
    try:
        with pytest.raises(hevcr.ParseFailed) as ei:
            P.feed(cut)
            P.flush()
            list(P.events())
        assert ei.value.rc < 0
    finally:
        P.close()
    # and a picture whose last segment never comes: the next picture's
    # first segment finds it open
    last_of_first_pic = [u for u in vcl if u[0] < first_of_pic[1][0]][-1]
    cut = data[:last_of_first_pic[0] - 3] + data[last_of_first_pic[1]:]
    P = hevcr.Parser()
    try:
        with pytest.raises(hevcr.ParseFailed):
            P.feed(cut)
            P.flush()
            list(P.events())
    finally:
        P.close()
