"""The oracle against the reference decoder on real bitstreams: every picture
of the reference's conformance streams (x265 encodes, tests/golden/streams),
reconstructed by the oracle from the syntax the reference's parser produced,
with its reference pictures taken from the oracle's own earlier outputs,
matches the reference's SHA-256 per plane after reconstruction, deblocking
and SAO; the whole decoded stream matches the MD5 the fixtures' README gives
for ffmpeg's decode."""
import pytest

import oracle_lib
import stream_goldens

STREAMS = stream_goldens.supported()


@pytest.mark.parametrize("stream", STREAMS, ids=[s["name"] for s in STREAMS])
def test_oracle_decodes_reference_streams(stream):
    dpb = [None] * stream["dpb_slots"]
    outputs = []
    for k, (fr, arrays) in enumerate(stream_goldens.frames(stream)):
        pics = oracle_lib.recon_syntax(arrays.syntax, dpb)
        for st, pic in zip(stream_goldens.STAGES, pics):
            assert pic.sha256() == fr["sha256"][st], f"picture {k} (POC {fr['poc']}) differs after {st}"
        dpb[fr["slot"]] = pics[2]
        outputs.append(pics[2])
    md5 = stream_goldens.yuv_md5(stream, outputs)
    assert md5 == stream["md5_yuv_reference"]
    if stream["md5_yuv_readme"]:
        assert md5 == stream["md5_yuv_readme"]


def test_stream_fixture_inventory():
    """Every conformance bitstream is either decoded above or listed with the
    reason the engine's syntax cannot describe it."""
    allv = stream_goldens.load()
    assert len(allv) >= 14
    for s in allv:
        assert ("unsupported" in s) != ("frames" in s)
