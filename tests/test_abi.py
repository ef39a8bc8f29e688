"""The engine's C-ABI libraries load without a GPU and export every function
their headers declare; the generator is deterministic."""
import ctypes as C
import os
import re

import hevcr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hevcr_[a-z0-9_]+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    lib = hevcr.engine_lib()
    names = declared("hevcr.h")
    assert "hevcr_decode_frame" in names and len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_kernel_names_match_the_header():
    """Every timed kernel the header numbers has its name, in the header's order."""
    text = open(os.path.join(ROOT, "include", "hevcr.h")).read()
    n = int(re.search(r"#define HEVCR_NUM_KERNELS\s+(\d+)", text).group(1))
    lib = hevcr.engine_lib()
    names = [lib.hevcr_kernel_name(k) for k in range(n)]
    assert all(names) and len(set(names)) == n
    assert tuple(x.decode() for x in names) == hevcr.KERNELS
    assert lib.hevcr_kernel_name(n) == b""


def test_front_end_exports_every_declared_symbol():
    lib = hevcr.front_lib()
    names = declared("hevcr_parse.h")
    assert "hevcr_parser_next" in names
    for n in names:
        assert hasattr(lib, n), n


def test_decoder_api_exports_hevc_api_names():
    """The drop-in decoder exports hevc_api.h's functions under their own names."""
    text = open(os.path.join(ROOT, "include", "hevcr_api.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b((?:hevc|hevcr)_decoder_[a-z_]+)\s*\(", text)))
    assert len(names) == 11
    lib = hevcr.decoder_lib()
    for n in names:
        assert hasattr(lib, n), n
    # no device is touched before the first decode call
    d = lib.hevc_decoder_create()
    assert d
    assert lib.hevc_decoder_get_frame_count(d) == 0
    assert lib.hevc_decoder_get_frame(d, 0, C.byref(hevcr.HEVCFrame())) == -1
    assert lib.hevc_decoder_get_info(d, C.byref(hevcr.HEVCStreamInfo())) == -1
    cnt = C.c_int(7)
    assert lib.hevc_decoder_drain(d, C.byref(cnt)) == 0 and cnt.value == 0
    assert lib.hevc_decoder_decode(d, None, 0) == -1
    lib.hevc_decoder_destroy(d)


def test_synth_exports_every_declared_symbol():
    lib = hevcr.synth_lib()
    for n in declared("hevcr_synth.h"):
        assert hasattr(lib, n), n


def test_engine_create_reports_missing_device_without_computing():
    import torch  # noqa: F401  (same HIP runtime as on the GPU host)
    if torch.cuda.is_available():
        return
    h = C.c_void_p()
    rc = hevcr.engine_lib().hevcr_engine_create(0, C.byref(h))
    assert rc == -5 and not h.value
    assert b"GPU" in hevcr.engine_lib().hevcr_last_error()


def test_generator_is_deterministic_and_seeded():
    a = hevcr.Frame(hevcr.preset(0), 7)
    b = hevcr.Frame(hevcr.preset(0), 7)
    c = hevcr.Frame(hevcr.preset(0), 8)
    import oracle_lib
    assert oracle_lib.syntax_digest(a) == oracle_lib.syntax_digest(b)
    assert oracle_lib.syntax_digest(a) != oracle_lib.syntax_digest(c)


def test_generator_rejects_bad_config():
    cfg = hevcr.preset(0, width=100)  # not a multiple of 8
    try:
        hevcr.Frame(cfg, 1)
    except ValueError:
        return
    raise AssertionError("accepted a width that is not a multiple of 8")


def test_egress_sizes():
    """hevcr_egress_size (host only): write_yuv bytes of the conformance
    window (picture.cpp:42-99: chroma offsets are the luma ones / 2; 8-bit
    samples as bytes, deeper ones as 2 bytes), or the whole uint16 planes."""
    for w, h, bd in ((1920, 1088, 8), (176, 144, 10), (416, 240, 8), (64, 64, 12)):
        for crop in (None, (0, 0, 0, 8), (2, 4, 6, 8)):
            l, r, t, b = crop or (0, 0, 0, 0)
            bps = 2 if bd > 8 else 1
            want = ((w - l - r) * (h - t - b) + 2 * (w // 2 - l // 2 - r // 2) * (h // 2 - t // 2 - b // 2)) * bps
            assert hevcr.egress_size(w, h, bd, hevcr.EGRESS_YUV, crop) == want
            assert hevcr.egress_size(w, h, bd, hevcr.EGRESS_U16, crop) == 2 * (w * h + 2 * (w // 2) * (h // 2))
    for bad in ((1, 0, 0, 0), (0, 0, -2, 0), (100, 100, 0, 0), (0, 0, 30, 34)):
        try:
            hevcr.egress_size(176, 64, 8, hevcr.EGRESS_YUV, bad)
        except hevcr.EngineError as e:
            assert e.rc == -3  # HEVCR_EINVAL
        else:
            raise AssertionError(f"crop {bad} accepted")
    try:
        hevcr.egress_size(176, 144, 8, 7, None)
    except hevcr.EngineError as e:
        assert e.rc == -3  # HEVCR_EINVAL
    else:
        raise AssertionError("format 7 accepted")
