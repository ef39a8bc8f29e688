"""The host front end (libhevcr_front.so, include/hevcr_parse.h) against the
reference decoder's own parser: on every conformance bitstream of the
reference (tests/golden/bitstreams, x265 encodes incl. 50 pictures of 1080p
Big Buck Bunny with wavefront entry points, scaling lists and temporal MV
prediction), each picture's records equal the ones the reference's parser
handed to its reconstruction operators (tests/golden/streams, written by the
tap) field for field; pictures come out in POC order; a DPB slot is not
reused before its picture is output; feeding in small chunks changes
nothing; and the CPU oracle fed from the parser reproduces ffmpeg's YUV MD5."""
import ctypes as C
import os

import numpy as np
import pytest

import hevcr
import oracle_lib
import stream_goldens

BITSTREAMS = os.path.join(stream_goldens.GOLDEN_DIR, "bitstreams")
STREAMS = stream_goldens.supported()
PARAMS = ("width", "height", "bit_depth", "ctb_log2", "min_cb_log2", "slice_type", "poc", "strong_intra_smoothing",
          "sao_enabled", "pps_cb_qp_offset", "pps_cr_qp_offset", "slice_cb_qp_offset", "slice_cr_qp_offset",
          "deblocking_disabled", "beta_offset_div2", "tc_offset_div2", "weighted_pred")


def bitstream(name: str) -> bytes:
    with open(os.path.join(BITSTREAMS, name + ".265"), "rb") as f:
        return f.read()


def records(s: hevcr.FrameSyntax) -> dict:
    """Everything of a picture's syntax that reconstruction reads, slot numbers
    aside (they are the caller's), unused PU lists and weights cleared."""
    p = s.p
    out = {f: getattr(p, f) for f in PARAMS}
    out["refs"] = [[p.ref_poc[l][i] for i in range(p.num_ref[l])] for l in range(2)]
    if p.weighted_pred:
        out["wd"] = (p.luma_log2_wd, p.chroma_log2_wd)
        out["weights"] = [[bytes(p.weights[l][i]) for i in range(p.num_ref[l])] for l in range(2)]
    out["cus"] = C.string_at(s.cus, s.num_cus * C.sizeof(hevcr.CU)) if s.num_cus else b""
    out["tus"] = C.string_at(s.tus, s.num_tus * C.sizeof(hevcr.TU)) if s.num_tus else b""
    pus = []
    for i in range(s.num_pus):
        u = s.pus[i]
        pus.append((u.x, u.y, u.w, u.h, u.pred_flags) +
                   tuple((tuple(u.mv[l]), u.ref_idx[l]) if u.pred_flags >> l & 1 else None for l in range(2)))
    out["pus"] = pus
    out["coeffs"] = C.string_at(s.coeffs, s.num_coeffs * 2) if s.num_coeffs else b""
    if p.sao_enabled:
        wc = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2
        hc = (p.height + (1 << p.ctb_log2) - 1) >> p.ctb_log2
        out["sao"] = C.string_at(s.sao, wc * hc * C.sizeof(hevcr.SAO))
    out["scaling"] = bytes(s.scaling.contents) if s.scaling else None
    return out


def parse(data: bytes, chunk: int = 0):
    """[(event fields, records or None, syntax copy or None)] in event order."""
    P = hevcr.Parser()
    out = []

    def pull():
        for ev in P.events():
            e = {f: getattr(ev, f) for f in ("kind", "sequence", "slot", "poc", "width", "height", "bit_depth",
                                              "dpb_slots", "decode_index")}
            e["crop"] = list(ev.crop)
            if ev.kind == hevcr.EVENT_DECODE:
                s = ev.syntax.contents
                out.append((e, records(s), copy_syntax(s)))
            else:
                out.append((e, None, None))

    if chunk:
        for i in range(0, len(data), chunk):
            P.feed(data[i:i + chunk])
            pull()
    else:
        P.feed(data)
    P.flush()
    pull()
    P.close()
    return out


def copy_syntax(s: hevcr.FrameSyntax) -> hevcr.SyntaxArrays:
    p = s.p
    wc = (p.width + (1 << p.ctb_log2) - 1) >> p.ctb_log2
    hc = (p.height + (1 << p.ctb_log2) - 1) >> p.ctb_log2
    take = (lambda ptr, n, sz: C.string_at(ptr, n * sz) if n else b"")
    coeffs = np.frombuffer(take(s.coeffs, s.num_coeffs, 2), np.int16).copy()
    return hevcr.SyntaxArrays(bytes(p), take(s.cus, s.num_cus, 16), take(s.pus, s.num_pus, 20),
                              take(s.tus, s.num_tus, 28), coeffs, take(s.sao, wc * hc, 24),
                              bytes(s.scaling.contents) if s.scaling else None)


_cache = {}


def parsed(name: str):
    if name not in _cache:
        _cache[name] = parse(bitstream(name))
    return _cache[name]


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_parser_matches_reference_records(stream):
    got = [r for e, r, _ in parsed(stream["name"]) if e["kind"] == hevcr.EVENT_DECODE]
    want = [records(a.syntax) for _, a in stream_goldens.frames(stream)]
    assert len(got) == len(want)
    for k, (g, w) in enumerate(zip(got, want)):
        for key in w:
            assert g.get(key) == w[key], f"picture {k} (POC {w['poc']}): {key} differs"


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_parser_output_order_and_slots(stream):
    evs = parsed(stream["name"])
    frames = stream["frames"]
    decoded, outs = [], []
    held = {}  # slot -> decode index still waiting for output
    for e, _, _ in evs:
        if e["kind"] == hevcr.EVENT_DECODE:
            assert e["slot"] not in held, "slot reused before its picture was output"
            assert 0 <= e["slot"] < e["dpb_slots"]
            held[e["slot"]] = e["decode_index"]
            decoded.append(e["poc"])
        else:
            assert held.pop(e["slot"]) == e["decode_index"]
            outs.append((frames[e["decode_index"]]["cvs"], e["poc"]))
            assert e["crop"] == stream["crop"]
    assert not held
    assert [f["poc"] for f in frames] == decoded
    assert outs == sorted((f["cvs"], f["poc"]) for f in frames)  # POC order within each coded video sequence


@pytest.mark.parametrize("name", ["b_qcif_10f", "full_qcif_10f_10bit", "i_64x64_full"])
def test_parser_chunked_feed_is_identical(name):
    whole = [(e, r) for e, r, _ in parsed(name)]
    for chunk in (1, 1000):
        assert [(e, r) for e, r, _ in parse(bitstream(name), chunk)] == whole


@pytest.mark.parametrize("stream", [s for s in STREAMS if s["name"] != "bbb1080_50f"],
                         ids=[s["name"] for s in STREAMS if s["name"] != "bbb1080_50f"])
def test_parser_and_oracle_reproduce_ffmpeg_md5(stream):
    """Bitstream -> host front end -> CPU oracle -> write_yuv order: the
    stream's MD5 from the fixtures' README (ffmpeg's decode)."""
    evs = parsed(stream["name"])
    dpb = [None] * max(e["dpb_slots"] for e, _, _ in evs)
    outputs = []
    for e, _, syn in evs:
        if e["kind"] == hevcr.EVENT_DECODE:
            dpb[e["slot"]] = oracle_lib.recon_syntax(syn.syntax, dpb)[2]
        else:
            outputs.append(((stream["frames"][e["decode_index"]]["cvs"], e["poc"]), dpb[e["slot"]]))
    order = {(f["cvs"], f["poc"]): k for k, f in enumerate(stream["frames"])}
    pics = [None] * len(outputs)
    for key, pic in outputs:
        pics[order[key]] = pic
    md5 = stream_goldens.yuv_md5(stream, pics)
    assert md5 == stream["md5_yuv_reference"]
    if stream["md5_yuv_readme"]:
        assert md5 == stream["md5_yuv_readme"]


def test_parser_rejects_garbage_and_reports_the_error():
    P = hevcr.Parser()
    data = bitstream("i_64x64_full")
    # a picture whose slice data was cut short: the slice parser fails loudly
    cut = data[: len(data) - 200]
    P.feed(cut)
    with pytest.raises(hevcr.ParseFailed) as ei:
        P.flush()
        list(P.events())
    assert ei.value.rc < 0
    P.close()
    # a slice with no parameter sets before it
    P = hevcr.Parser()
    start = next(i for i in range(len(data) - 3)
                 if data[i:i + 3] == b"\x00\x00\x01" and 16 <= (data[i + 3] >> 1) & 0x3f <= 21)  # first IRAP slice
    with pytest.raises(hevcr.ParseFailed):
        P.feed(data[start:])
        P.flush()
    P.close()


def test_parser_survives_corrupted_bitstreams():
    """Bit flips, random bytes and truncation in the conformance streams end in
    a parse error or in pictures, never in a crash (the same mutations ran
    under AddressSanitizer / UBSan against a host build of the front end)."""
    import random
    rng = random.Random(20261204)
    names = ["i_64x64_full", "toy_qp30", "p_qcif_10f", "b_qcif_10f", "full_qcif_10f_10bit"]
    outcomes = {"pictures": 0, "error": 0}
    for _ in range(120):
        data = bytearray(bitstream(rng.choice(names)))
        for _ in range(rng.randint(1, 20)):
            k = rng.randrange(len(data))
            data[k] = rng.randrange(256) if rng.random() < 0.7 else data[k] ^ (1 << rng.randrange(8))
        if rng.random() < 0.2:
            data = data[:rng.randrange(len(data))]
        P = hevcr.Parser()
        try:
            P.feed(bytes(data))
            P.flush()
            for _ev in P.events():
                pass
            outcomes["pictures"] += 1
        except hevcr.ParseFailed as e:
            assert e.rc < 0
            outcomes["error"] += 1
        P.close()
    assert outcomes["error"] > 0 and sum(outcomes.values()) == 120


class _BitWriter:
    def __init__(self):
        self.bits = []

    def u(self, v, n):
        self.bits += [(v >> (n - 1 - i)) & 1 for i in range(n)]

    def ue(self, v):
        v += 1
        n = v.bit_length()
        self.u(0, n - 1)
        self.u(v, n)

    def nal(self, header: bytes) -> bytes:
        bits = self.bits + [1]
        bits += [0] * (-len(bits) % 8)
        raw = bytes(int("".join(map(str, bits[i:i + 8])), 2) for i in range(0, len(bits), 8))
        out, zeros = bytearray(), 0
        for byte in raw:  # emulation prevention (7.4.2)
            if zeros >= 2 and byte <= 3:
                out.append(3)
                zeros = 0
            out.append(byte)
            zeros = zeros + 1 if byte == 0 else 0
        return b"\x00\x00\x00\x01" + header + bytes(out)


def _sps_with_predicted_rps(num_positive: int) -> bytes:
    """An SPS whose set 0 has num_positive positive pictures and whose set 1
    is predicted from it with delta_rps = +1 and every entry used: the
    predicted set holds num_positive + 1 positive pictures (7-62)."""
    w = _BitWriter()
    w.u(0, 4); w.u(0, 3); w.u(1, 1)             # vps id, max_sub_layers_minus1, temporal nesting
    w.u(0, 2); w.u(0, 1); w.u(1, 5)             # profile space, tier, profile_idc (Main)
    w.u(0, 32 + 4 + 43 + 1); w.u(93, 8)         # flags, level_idc
    w.ue(0); w.ue(1); w.ue(64); w.ue(64); w.u(0, 1)  # sps id, 4:2:0, 64x64, no conformance window
    w.ue(0); w.ue(0); w.ue(4); w.u(1, 1)        # 8-bit, log2_max_poc_lsb 8, ordering info for all layers
    w.ue(15); w.ue(0); w.ue(0)                  # max_dec_pic_buffering 16
    w.ue(0); w.ue(3); w.ue(0); w.ue(3); w.ue(0); w.ue(0)  # CB 8..64, TB 4..32, depths 0
    w.u(0, 1); w.u(0, 1); w.u(0, 1); w.u(0, 1)  # scaling lists, AMP, SAO, PCM off
    w.ue(2)                                     # two short-term sets
    w.ue(0); w.ue(num_positive)                 # set 0: explicitly coded
    for _ in range(num_positive):
        w.ue(0); w.u(1, 1)
    w.u(1, 1); w.u(0, 1); w.ue(0)               # set 1: predicted from set 0, delta_rps = +1
    for _ in range(num_positive + 1):
        w.u(1, 1)                               # used_by_curr_pic_flag
    w.u(0, 1); w.u(0, 1); w.u(0, 1); w.u(0, 1)  # long-term, TMVP, strong smoothing, VUI
    w.u(0, 1)                                   # sps_extension_present_flag
    return w.nal(b"\x42\x01")


def test_parser_rejects_predicted_rps_past_16_entries():
    """A set predicted from one with 16 positive pictures would hold 17: the
    parser rejects it instead of writing past its 16-entry arrays, and it
    rejects sets larger than the SPS's DPB."""
    for num_positive in (16, 15):
        P = hevcr.Parser()
        with pytest.raises(hevcr.ParseFailed) as ei:
            P.feed(_sps_with_predicted_rps(num_positive))
            P.flush()
            list(P.events())
        assert "RPS" in str(ei.value)
        P.close()
