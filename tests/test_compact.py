"""The compact coefficient form (hevcr_frame_syntax.sig, made by
hevcr_syntax_compact): every coded block's list, expanded back, is the dense
block the reference's residual coding wrote (residual_coding.cpp:438), with
no position twice and nothing else changed in the syntax. Host only."""
import ctypes as C

import numpy as np
import pytest

import hevcr
import oracle_lib
import stream_goldens

CASES = [c for c in oracle_lib.load_goldens() if c["name"] in
         ("intra416_s1", "intra416_10bit", "intra200x136_ctb16", "b640x360_10bit", "intra416_scaling", "b1080_s1",
          "intra416_slices_pcm", "b200x136_ctb32_10bit_slices")]


def _blocks(syn):
    tus = np.frombuffer(C.string_at(syn.tus, syn.num_tus * C.sizeof(hevcr.TU)), np.uint8).reshape(-1, 28)
    for ti, t in enumerate(tus):
        log2, blk, cbf = int(t[8]), int(t[9]), int(t[10])
        offs = t[16:28].view(np.uint32)
        for c in range(3):
            if not (cbf >> c) & 1 or (c and not (log2 > 2 or blk == 3)):
                continue
            n = (1 << (log2 - 1) if log2 > 2 else 4) if c else 1 << log2
            yield ti, c, n, int(offs[c])


def _check(frame_syntax_owner):
    dense = frame_syntax_owner.syntax
    comp = hevcr.CompactSyntax(frame_syntax_owner)
    s = comp.syntax
    assert s.num_coeffs == 0 and s.num_sig > 0
    co = np.ctypeslib.as_array(dense.coeffs, (dense.num_coeffs,))
    sig = comp.sig[:s.num_sig]
    nz_dense = 0
    for (ti, c, n, off_d), (ti2, c2, n2, off_s) in zip(_blocks(dense), _blocks(s)):
        assert (ti, c, n) == (ti2, c2, n2)
        want = co[off_d:off_d + n * n]
        k = int(sig[off_s])
        words = sig[off_s + 1:off_s + 1 + k]
        pos = words & 0xFFFF
        assert len(set(pos.tolist())) == k and (pos < n * n).all()
        got = np.zeros(n * n, np.int16)
        got[pos] = (words >> 16).astype(np.uint16).view(np.int16)
        np.testing.assert_array_equal(got, want)
        nz_dense += int((want != 0).sum())
    assert s.num_sig == nz_dense + sum(1 for _ in _blocks(dense))
    for f in ("num_cus", "num_pus", "num_tus"):
        assert getattr(s, f) == getattr(dense, f)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_compact_form_expands_to_the_dense_blocks(case):
    _check(oracle_lib.make_frame(case))


def test_compact_form_of_a_real_stream():
    stream = next(s for s in stream_goldens.supported() if s["name"] == "bbb1080_50f")
    for k, (_, arrays) in enumerate(stream_goldens.frames(stream)):
        if k % 10 == 0:
            _check(arrays)


def test_compacting_twice_is_rejected():
    f = oracle_lib.make_frame(CASES[0])
    comp = hevcr.CompactSyntax(f)
    with pytest.raises(hevcr.EngineError):
        hevcr.CompactSyntax(comp)
