"""Invalid syntax and arguments are rejected on the host, before any kernel
runs, with the documented error codes (mirroring the reference C API's
null/range checks, src/wasm/hevc_api.cpp:28, 50-51, 108-109)."""
import ctypes as C

import pytest

import hevcr

pytestmark = pytest.mark.gpu
EINVAL, EUNSUPPORTED = -3, -6


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _rc(fn):
    try:
        fn()
    except hevcr.EngineError as e:
        return e.rc
    return 0


def test_rejects_mismatched_and_broken_syntax(engine):
    f = hevcr.Frame(hevcr.preset(0), 21)
    sid = engine.stream(f.width, f.height, 8, 2)
    try:
        s = f.syntax
        assert _rc(lambda: engine.prepare(sid, s).release()) == 0
        s.p.width = 408
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.p.width = 416
        x0 = s.cus[3].x
        s.cus[3].x = 4000
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.cus[3].x = x0
        off = s.tus[5].coeff_off[0]
        cbf = s.tus[5].cbf
        s.tus[5].cbf |= 1
        s.tus[5].coeff_off[0] = s.num_coeffs
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.tus[5].coeff_off[0] = off
        s.tus[5].cbf = cbf
        mode = s.cus[0].intra_luma[0]
        s.cus[0].intra_luma[0] = 35
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.cus[0].intra_luma[0] = mode
        s.p.min_cb_log2 = 4  # the frame has 8x8 CUs
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.p.min_cb_log2 = 2
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.p.min_cb_log2 = 3
        k = next(i for i in range(s.num_tus) if s.tus[i].log2_size >= 3 and s.tus[i].cbf & 1)
        s.tus[k].tskip = 1  # transform skip above 4x4
        assert _rc(lambda: engine.prepare(sid, s)) == EUNSUPPORTED
        s.tus[k].tskip = 0
        assert _rc(lambda: engine.prepare(sid, s).release()) == 0
    finally:
        engine.destroy_stream(sid)


def test_rejects_overlapping_and_uncovered_blocks(engine):
    """Coverage is checked from the records' z-order runs: a CU on top of the
    one before it, a picture with a CU missing, transform units on top of
    each other and an intra CU its transform units do not cover are all
    rejected; the untouched frame is accepted after each."""
    f = hevcr.Frame(hevcr.preset(0), 21)
    sid = engine.stream(f.width, f.height, 8, 2)
    try:
        s = f.syntax
        i = next(i for i in range(1, s.num_cus) if s.cus[i].log2_size == s.cus[i - 1].log2_size)
        x, y = s.cus[i].x, s.cus[i].y
        s.cus[i].x, s.cus[i].y = s.cus[i - 1].x, s.cus[i - 1].y
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.cus[i].x, s.cus[i].y = x, y
        n = s.num_cus
        last_tus = sum(1 for k in range(s.num_tus) if s.tus[k].cu_idx == n - 1)
        s.num_cus, s.num_tus = n - 1, s.num_tus - last_tus  # the last CU and its units gone
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.num_cus, s.num_tus = n, s.num_tus + last_tus
        assert _rc(lambda: engine.prepare(sid, s).release()) == 0
        k = next(k for k in range(1, s.num_tus) if s.tus[k].cu_idx == s.tus[k - 1].cu_idx
                 and s.tus[k].log2_size == s.tus[k - 1].log2_size and s.tus[k].log2_size > 2)
        x, y, xb, yb = s.tus[k].x, s.tus[k].y, s.tus[k].xbase, s.tus[k].ybase
        s.tus[k].x, s.tus[k].y = s.tus[k - 1].x, s.tus[k - 1].y  # two units on one square
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.tus[k].x, s.tus[k].y = x, y
        s.tus[k].log2_size -= 1  # a smaller unit leaves part of the intra CU uncovered
        assert _rc(lambda: engine.prepare(sid, s)) == EINVAL
        s.tus[k].log2_size += 1
        assert (s.tus[k].xbase, s.tus[k].ybase) == (xb, yb)
        assert _rc(lambda: engine.prepare(sid, s).release()) == 0
    finally:
        engine.destroy_stream(sid)


def test_rejects_bad_reference_use(engine):
    f = hevcr.Frame(hevcr.preset(1, width=256, height=128), 3)
    sid = engine.stream(f.width, f.height, 8, 5)
    try:
        pf = engine.prepare(sid, f.syntax)
        assert _rc(lambda: pf.run(0)) == EINVAL  # slot 0 is a reference of this frame
        assert _rc(lambda: pf.run(7)) == EINVAL  # no such slot
        pf.release()
        f.syntax.pus[0].ref_idx[0] = 15
        f.syntax.pus[0].pred_flags = 1
        assert _rc(lambda: engine.prepare(sid, f.syntax)) == EINVAL
    finally:
        engine.destroy_stream(sid)


def test_rejects_bad_arguments(engine):
    lib = hevcr.engine_lib()
    sid = C.c_int()
    assert lib.hevcr_stream_create(engine.h, 100, 64, 8, 2, C.byref(sid)) == EINVAL
    assert lib.hevcr_stream_create(engine.h, 64, 64, 12, 2, C.byref(sid)) == EUNSUPPORTED
    assert lib.hevcr_stream_create(engine.h, 64, 64, 8, 0, C.byref(sid)) == EINVAL
    assert lib.hevcr_stream_destroy(engine.h, 12345) == EINVAL
    assert lib.hevcr_frame_run(engine.h, None, 0, 0) == EINVAL
