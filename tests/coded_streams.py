"""Coded-stream golden cases: streams written by tests/golden/stream_writer.cpp
(random content over slice segments, dependent segments, tiles, wavefront
entry points, PCM, transquant bypass, weighted prediction per segment and
the other picture and slice options; tests/golden/make_coded_goldens.py
lists the cases) with, for each, the MD5 of the YUV the reference decoder
writes for it (its CLI, src/main.cpp, run on the stream). Shared by the
CPU tests (front end + oracle) and the GPU tests (the hevc_api.h decoder)."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os

import numpy as np

import hevcr

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CODED_DIR = os.path.join(GOLDEN_DIR, "coded")


def load() -> list[dict]:
    with open(os.path.join(GOLDEN_DIR, "coded.json")) as f:
        return json.load(f)["streams"]


def read(name: str) -> bytes:
    with open(os.path.join(CODED_DIR, name + ".265"), "rb") as f:
        return f.read()


def copy_syntax(s: hevcr.FrameSyntax) -> hevcr.SyntaxArrays:
    """A copy of a parser's picture syntax that outlives the parser's buffers."""
    p = s.p
    wc = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2
    hc = (p.height + (1 << p.ctb_log2) - 1) >> p.ctb_log2

    def take(ptr, n, sz):
        return C.string_at(ptr, n * sz) if n else b""

    coeffs = np.frombuffer(take(s.coeffs, s.num_coeffs, 2), np.int16).copy()
    slices = ctb_slice = None
    if s.num_slices:
        slices = take(s.slices, s.num_slices, C.sizeof(hevcr.Slice))
        ctb_slice = take(s.ctb_slice, wc * hc, 1)
    return hevcr.SyntaxArrays(bytes(p), take(s.cus, s.num_cus, 16), take(s.pus, s.num_pus, 20),
                              take(s.tus, s.num_tus, 28), coeffs, take(s.sao, wc * hc, 24),
                              bytes(s.scaling.contents) if s.scaling else None, slices, ctb_slice,
                              bytes(s.tiles.contents) if s.tiles else None, s.pcm_loop_filter_disabled)


def parse(data: bytes) -> list:
    """[(event fields, SyntaxArrays copy or None)] in event order."""
    P = hevcr.Parser()
    out = []
    try:
        P.feed(data)
        P.flush()
        for ev in P.events():
            e = {f: getattr(ev, f) for f in ("kind", "sequence", "slot", "poc", "width", "height", "bit_depth",
                                              "dpb_slots", "decode_index")}
            e["crop"] = list(ev.crop)
            out.append((e, copy_syntax(ev.syntax.contents) if ev.kind == hevcr.EVENT_DECODE else None))
    finally:
        P.close()
    return out


def md5_of_outputs(pictures: list, crop, bit_depth: int) -> str:
    """MD5 of pictures written in output order as the reference's write_yuv
    does (picture.cpp:42-99): cropped to the conformance window, 8-bit
    samples as bytes, deeper ones as 16-bit little-endian."""
    h = hashlib.md5()
    cl, cr, ct, cb = crop
    for planes in pictures:
        for c in range(3):
            s = 1 if c == 0 else 2
            p = planes[c]
            p = p[ct // s:p.shape[0] - cb // s, cl // s:p.shape[1] - cr // s]
            h.update(p.astype(np.uint8).tobytes() if bit_depth <= 8 else p.astype("<u2").tobytes())
    return h.hexdigest()
