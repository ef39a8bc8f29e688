"""ctypes binding of the reconstruction engine's C-ABI (include/hevcr.h) and of
the seeded syntax generator (include/hevcr_synth.h).

The engine itself is C++/HIP (libhevcr.so); this module only marshals
numpy buffers across the C-ABI for tests and the benchmark. Loading fails
loudly when a library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")

SLICE_B, SLICE_P, SLICE_I = 0, 1, 2
MODE_INTER, MODE_INTRA, MODE_SKIP = 0, 1, 2
MAX_REFS = 16


class CU(C.Structure):
    _fields_ = [("x", C.c_uint16), ("y", C.c_uint16), ("log2_size", C.c_uint8),
                ("pred_mode", C.c_uint8), ("part_mode", C.c_uint8), ("qp_y", C.c_int8),
                ("intra_luma", C.c_uint8 * 4), ("intra_chroma", C.c_uint8),
                ("num_pu", C.c_uint8), ("flags", C.c_uint16)]


class PU(C.Structure):
    _fields_ = [("x", C.c_uint16), ("y", C.c_uint16), ("w", C.c_uint16), ("h", C.c_uint16),
                ("mv", (C.c_int16 * 2) * 2), ("ref_idx", C.c_int8 * 2),
                ("pred_flags", C.c_uint8), ("reserved", C.c_uint8)]


class TU(C.Structure):
    _fields_ = [("x", C.c_uint16), ("y", C.c_uint16), ("xbase", C.c_uint16), ("ybase", C.c_uint16),
                ("log2_size", C.c_uint8), ("blk_idx", C.c_uint8), ("cbf", C.c_uint8),
                ("tskip", C.c_uint8), ("cu_idx", C.c_uint32), ("coeff_off", C.c_uint32 * 3)]


class SAO(C.Structure):
    _fields_ = [("type", C.c_uint8 * 3), ("band_pos", C.c_uint8 * 3), ("eo_class", C.c_uint8 * 3),
                ("offset", (C.c_int8 * 5) * 3)]


class Weight(C.Structure):
    _fields_ = [("w", C.c_int16 * 3), ("o", C.c_int16 * 3)]


class Scaling(C.Structure):
    _fields_ = [("list", ((C.c_uint8 * 64) * 6) * 4), ("dc", (C.c_uint8 * 6) * 2)]


class FrameParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32),
                ("ctb_log2", C.c_int32), ("min_cb_log2", C.c_int32), ("slice_type", C.c_int32),
                ("poc", C.c_int32), ("strong_intra_smoothing", C.c_int32),
                ("intra_smoothing_disabled", C.c_int32), ("sao_enabled", C.c_int32),
                ("pps_cb_qp_offset", C.c_int32), ("pps_cr_qp_offset", C.c_int32),
                ("slice_cb_qp_offset", C.c_int32), ("slice_cr_qp_offset", C.c_int32),
                ("deblocking_disabled", C.c_int32), ("beta_offset_div2", C.c_int32),
                ("tc_offset_div2", C.c_int32), ("weighted_pred", C.c_int32),
                ("luma_log2_wd", C.c_int32), ("chroma_log2_wd", C.c_int32),
                ("num_ref", C.c_int32 * 2), ("ref_slot", (C.c_int32 * MAX_REFS) * 2),
                ("ref_poc", (C.c_int32 * MAX_REFS) * 2), ("weights", (Weight * MAX_REFS) * 2)]


MAX_SLICES, MAX_TILE_COLS, MAX_TILE_ROWS = 256, 20, 22
CU_PCM, CU_BYPASS = 1, 2


class Slice(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("slice_type", "slice_cb_qp_offset", "slice_cr_qp_offset",
                                          "deblocking_disabled", "beta_offset_div2", "tc_offset_div2",
                                          "loop_filter_across_slices", "weighted_pred", "luma_log2_wd",
                                          "chroma_log2_wd")] + \
               [("weights", (Weight * MAX_REFS) * 2)]


class Tiles(C.Structure):
    _fields_ = [("cols", C.c_int32), ("rows", C.c_int32), ("col_width", C.c_uint16 * MAX_TILE_COLS),
                ("row_height", C.c_uint16 * MAX_TILE_ROWS), ("loop_filter_across_tiles", C.c_int32)]


class FrameSyntax(C.Structure):
    _fields_ = [("p", FrameParams),
                ("cus", C.POINTER(CU)), ("num_cus", C.c_uint32),
                ("pus", C.POINTER(PU)), ("num_pus", C.c_uint32),
                ("tus", C.POINTER(TU)), ("num_tus", C.c_uint32),
                ("coeffs", C.POINTER(C.c_int16)), ("num_coeffs", C.c_uint64),
                ("sao", C.POINTER(SAO)), ("scaling", C.POINTER(Scaling)),
                ("sig", C.POINTER(C.c_uint32)), ("num_sig", C.c_uint64),
                ("slices", C.POINTER(Slice)), ("num_slices", C.c_uint32),
                ("ctb_slice", C.POINTER(C.c_uint8)), ("tiles", C.POINTER(Tiles)),
                ("pcm_loop_filter_disabled", C.c_int32)]


class SynthCfg(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("width", "height", "bit_depth", "ctb_log2", "slice_type",
                                          "poc", "qp_min", "qp_max", "level_mul")] + \
               [(n, C.c_uint32) for n in ("p_cu_split", "p_tu_split", "p_intra_nxn", "p_cbf_luma",
                                           "p_cbf_chroma", "p_intra_cu", "p_skip", "p_root_cbf_zero",
                                           "p_bi", "p_wp")] + \
               [(n, C.c_int32) for n in ("mv_sigma_qpel", "mv_clamp_qpel", "num_refs",
                                          "num_refs_before", "weighted", "amp", "sao_enabled")] + \
               [("p_sao_off", C.c_uint32), ("deblocking_disabled", C.c_int32),
                ("strong_intra_smoothing", C.c_int32), ("scaling_lists", C.c_int32),
                ("num_slices", C.c_int32), ("p_dependent", C.c_uint32), ("p_slice_type_change", C.c_uint32),
                ("tile_cols", C.c_int32), ("tile_rows", C.c_int32), ("loop_filter_across_tiles", C.c_int32),
                ("p_pcm", C.c_uint32), ("pcm_bits", C.c_int32), ("pcm_loop_filter_disabled", C.c_int32),
                ("p_bypass", C.c_uint32)]


def prob(p: float) -> int:
    """Fixed-point probability with 16 fractional bits."""
    return int(p * 65536.0 + 0.5)


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run __graft_entry__.build() first")
    return C.CDLL(path)


_synth = None


def synth_lib() -> C.CDLL:
    global _synth
    if _synth is None:
        lib = _load("libhevcr_synth.so")
        lib.hevcr_synth_preset.argtypes = [C.c_int, C.POINTER(SynthCfg)]
        lib.hevcr_synth_generate.argtypes = [C.POINTER(SynthCfg), C.c_uint64]
        lib.hevcr_synth_generate.restype = C.c_void_p
        lib.hevcr_synth_syntax.argtypes = [C.c_void_p]
        lib.hevcr_synth_syntax.restype = C.POINTER(FrameSyntax)
        lib.hevcr_synth_free.argtypes = [C.c_void_p]
        lib.hevcr_synth_num_ref_pictures.argtypes = [C.c_void_p]
        lib.hevcr_synth_ref_picture.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64] + [C.c_void_p] * 3
        _synth = lib
    return _synth


def preset(n: int, **overrides) -> SynthCfg:
    cfg = SynthCfg()
    if synth_lib().hevcr_synth_preset(n, C.byref(cfg)) != 0:
        raise ValueError(f"unknown preset {n}")
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise AttributeError(k)
        setattr(cfg, k, prob(v) if k.startswith("p_") and isinstance(v, float) else v)
    return cfg


class Picture:
    """Three uint16 planes, stride == width, 4:2:0."""

    def __init__(self, width: int, height: int, bit_depth: int = 8, planes=None):
        self.width, self.height, self.bit_depth = width, height, bit_depth
        if planes is None:
            planes = [np.zeros((height, width), np.uint16),
                      np.zeros((height // 2, width // 2), np.uint16),
                      np.zeros((height // 2, width // 2), np.uint16)]
        self.planes = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]

    def ptrs(self):
        return [p.ctypes.data for p in self.planes]

    def plane_bytes(self, c: int) -> bytes:
        """Serialisation of write_yuv: 1 byte/sample up to 8 bits, uint16 LE above."""
        p = self.planes[c]
        return p.astype(np.uint8).tobytes() if self.bit_depth <= 8 else p.astype("<u2").tobytes()

    def sha256(self) -> list[str]:
        return [hashlib.sha256(self.plane_bytes(c)).hexdigest() for c in range(3)]

    def copy(self) -> "Picture":
        return Picture(self.width, self.height, self.bit_depth, [p.copy() for p in self.planes])


def synth_ref_picture(width: int, height: int, bit_depth: int, seed: int) -> Picture:
    pic = Picture(width, height, bit_depth)
    rc = synth_lib().hevcr_synth_ref_picture(width, height, bit_depth, seed & (2**64 - 1), *pic.ptrs())
    if rc != 0:
        raise ValueError("hevcr_synth_ref_picture failed")
    return pic


class Frame:
    """One generated frame of syntax plus the reference pictures its lists name."""

    def __init__(self, cfg: SynthCfg, seed: int, refs: list[Picture] | None = None, with_refs: bool = True):
        lib = synth_lib()
        self.cfg, self.seed = cfg, seed
        self._h = lib.hevcr_synth_generate(C.byref(cfg), seed & (2**64 - 1))
        if not self._h:
            raise ValueError("invalid synth configuration")
        self.syntax = lib.hevcr_synth_syntax(self._h).contents
        self.num_ref_pics = lib.hevcr_synth_num_ref_pictures(self._h)
        if not with_refs:  # syntax only: the caller maps list entries to its own pictures
            self.refs = []
            return
        if refs is None:
            refs = [synth_ref_picture(cfg.width, cfg.height, cfg.bit_depth, seed * 1000003 + 17 + k)
                    for k in range(self.num_ref_pics)]
        if len(refs) < self.num_ref_pics:
            raise ValueError("not enough reference pictures")
        self.refs = refs

    def ref_plane_array(self):
        arr = (C.c_void_p * max(1, 3 * len(self.refs)))()
        for k, r in enumerate(self.refs):
            for c in range(3):
                arr[3 * k + c] = r.planes[c].ctypes.data
        return arr

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            synth_lib().hevcr_synth_free(h)
            self._h = None

    @property
    def width(self):
        return self.syntax.p.width

    @property
    def height(self):
        return self.syntax.p.height

    @property
    def bit_depth(self):
        return self.syntax.p.bit_depth

    def new_picture(self) -> Picture:
        return Picture(self.width, self.height, self.bit_depth)


EVENT_DECODE, EVENT_OUTPUT = 1, 2


class ParserEvent(C.Structure):
    _fields_ = [("kind", C.c_int32), ("sequence", C.c_int32), ("slot", C.c_int32), ("poc", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32), ("dpb_slots", C.c_int32),
                ("crop", C.c_int32 * 4), ("decode_index", C.c_int32), ("profile_idc", C.c_int32),
                ("level_idc", C.c_int32), ("syntax", C.POINTER(FrameSyntax))]


def front_lib() -> C.CDLL:
    """The host front end (include/hevcr_parse.h): bitstream -> syntax."""
    global _front
    if _front is None:
        L = _load("libhevcr_front.so")
        L.hevcr_parser_create.argtypes = [C.POINTER(C.c_void_p)]
        L.hevcr_parser_destroy.argtypes = [C.c_void_p]
        L.hevcr_parser_destroy.restype = None
        L.hevcr_parser_feed.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.hevcr_parser_flush.argtypes = [C.c_void_p]
        L.hevcr_parser_next.argtypes = [C.c_void_p, C.POINTER(ParserEvent)]
        L.hevcr_parser_error.argtypes = [C.c_void_p]
        L.hevcr_parser_error.restype = C.c_char_p
        _front = L
    return _front


_front = None


class ParseFailed(RuntimeError):
    def __init__(self, rc: int, what: str):
        super().__init__(f"{what} ({rc})")
        self.rc = rc


class Parser:
    """Annex-B bytes in; DECODE / OUTPUT events out (include/hevcr_parse.h).
    A DECODE event's syntax is valid until the next call to events()' next item."""

    def __init__(self):
        h = C.c_void_p()
        _check(front_lib().hevcr_parser_create(C.byref(h)), "hevcr_parser_create")
        self.h = h

    def close(self):
        if self.h:
            front_lib().hevcr_parser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rc(self, rc: int, what: str):
        if rc < 0:
            raise ParseFailed(rc, f"{what}: {front_lib().hevcr_parser_error(self.h).decode()}")

    def feed(self, data: bytes):
        self._rc(front_lib().hevcr_parser_feed(self.h, data, len(data)), "feed")

    def flush(self):
        self._rc(front_lib().hevcr_parser_flush(self.h), "flush")

    def events(self):
        ev = ParserEvent()
        while True:
            rc = front_lib().hevcr_parser_next(self.h, C.byref(ev))
            self._rc(rc, "next")
            if rc == 0:
                return
            yield ev


class HEVCFrame(C.Structure):
    _fields_ = [("y", C.POINTER(C.c_uint16)), ("cb", C.POINTER(C.c_uint16)), ("cr", C.POINTER(C.c_uint16)),
                ("width", C.c_int), ("height", C.c_int), ("stride_y", C.c_int), ("stride_c", C.c_int),
                ("chroma_width", C.c_int), ("chroma_height", C.c_int), ("bit_depth", C.c_int), ("poc", C.c_int)]


class HEVCStreamInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("width", "height", "bit_depth", "chroma_format", "profile", "level")]


_decoder = None


def decoder_lib() -> C.CDLL:
    """hevc_api.h's decoder over the front end and the engine (include/hevcr_api.h)."""
    global _decoder
    if _decoder is None:
        L = _load("libhevcr_decoder.so")
        L.hevc_decoder_create.restype = C.c_void_p
        L.hevc_decoder_destroy.argtypes = [C.c_void_p]
        L.hevc_decoder_destroy.restype = None
        L.hevc_decoder_decode.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.hevc_decoder_get_frame_count.argtypes = [C.c_void_p]
        L.hevc_decoder_get_frame.argtypes = [C.c_void_p, C.c_int, C.POINTER(HEVCFrame)]
        L.hevc_decoder_get_info.argtypes = [C.c_void_p, C.POINTER(HEVCStreamInfo)]
        L.hevc_decoder_feed.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.hevc_decoder_drain.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.hevc_decoder_get_drained_frame.argtypes = [C.c_void_p, C.c_int, C.POINTER(HEVCFrame)]
        L.hevc_decoder_flush.argtypes = [C.c_void_p]
        L.hevcr_decoder_error.argtypes = [C.c_void_p]
        L.hevcr_decoder_error.restype = C.c_char_p
        _decoder = L
    return _decoder


def frame_planes(f: HEVCFrame) -> list:
    """The cropped Y, Cb, Cr planes of an HEVCFrame as uint16 arrays (copies)."""
    out = []
    for ptr, w, h, stride in ((f.y, f.width, f.height, f.stride_y), (f.cb, f.chroma_width, f.chroma_height, f.stride_c),
                              (f.cr, f.chroma_width, f.chroma_height, f.stride_c)):
        a = np.ctypeslib.as_array(ptr, shape=((h - 1) * stride + w,)) if h else np.zeros(0, np.uint16)
        out.append(np.lib.stride_tricks.as_strided(a, shape=(h, w), strides=(stride * 2, 2)).copy())
    return out


class SyntaxArrays:
    """A FrameSyntax over record arrays held in numpy buffers: the form a host
    parser (or a stored stream) hands over. The buffers are kept alive here
    for as long as the syntax is in use.

    params: the raw bytes of one hevcr_frame_params; cus, pus, tus, sao: raw
    bytes of the record arrays (16, 20, 28 and 24 bytes per record); coeffs:
    the int16 coefficient pool the TU records index; scaling: the raw bytes of
    one hevcr_scaling, or None for flat factors. Pictures of several slice
    segments add slices (raw hevcr_slice records) and ctb_slice (one byte per
    CTB); tiles is the raw bytes of one hevcr_tiles or None;
    pcm_loop_filter_disabled is the SPS flag."""

    def __init__(self, params, cus, pus, tus, coeffs, sao, scaling=None, slices=None, ctb_slice=None,
                 tiles=None, pcm_loop_filter_disabled=0):
        self._bufs = [np.ascontiguousarray(np.frombuffer(b, np.uint8)) for b in (cus, pus, tus, sao)]
        self._coeffs = np.ascontiguousarray(coeffs, dtype=np.int16)
        pb = bytes(params)
        if len(pb) != C.sizeof(FrameParams):
            raise ValueError("params is not one hevcr_frame_params")
        s = FrameSyntax()
        C.memmove(C.addressof(s.p), pb, len(pb))
        sizes = (C.sizeof(CU), C.sizeof(PU), C.sizeof(TU), C.sizeof(SAO))
        for b, sz in zip(self._bufs, sizes):
            if len(b) % sz:
                raise ValueError("record array length is not a whole number of records")
        cu, pu, tu, sa = self._bufs
        s.cus, s.num_cus = C.cast(cu.ctypes.data, C.POINTER(CU)), len(cu) // sizes[0]
        s.pus, s.num_pus = C.cast(pu.ctypes.data, C.POINTER(PU)), len(pu) // sizes[1]
        s.tus, s.num_tus = C.cast(tu.ctypes.data, C.POINTER(TU)), len(tu) // sizes[2]
        s.coeffs, s.num_coeffs = C.cast(self._coeffs.ctypes.data, C.POINTER(C.c_int16)), len(self._coeffs)
        s.sao = C.cast(sa.ctypes.data, C.POINTER(SAO)) if len(sa) else None
        self._scaling = None
        if scaling is not None and len(scaling):
            self._scaling = Scaling.from_buffer_copy(bytes(scaling))
            s.scaling = C.pointer(self._scaling)
        if slices is not None and len(slices):
            if len(slices) % C.sizeof(Slice) or ctb_slice is None:
                raise ValueError("slices is not a whole number of hevcr_slice records, or ctb_slice is missing")
            self._slices = np.ascontiguousarray(np.frombuffer(bytes(slices), np.uint8))
            self._ctb_slice = np.ascontiguousarray(np.frombuffer(bytes(ctb_slice), np.uint8))
            s.slices = C.cast(self._slices.ctypes.data, C.POINTER(Slice))
            s.num_slices = len(self._slices) // C.sizeof(Slice)
            s.ctb_slice = C.cast(self._ctb_slice.ctypes.data, C.POINTER(C.c_uint8))
        self._tiles = None
        if tiles is not None:
            self._tiles = Tiles.from_buffer_copy(bytes(tiles))
            s.tiles = C.pointer(self._tiles)
        s.pcm_loop_filter_disabled = int(pcm_loop_filter_disabled)
        self.syntax = s

    @property
    def params(self) -> FrameParams:
        return self.syntax.p


class CompactSyntax:
    """The compact coefficient form of a dense frame's syntax
    (hevcr_syntax_compact): a copy of its FrameSyntax whose TU records point
    into a list of the nonzero coefficients; the other record arrays are
    shared with `source`, which is kept alive."""

    def __init__(self, source):
        src = source.syntax
        n = C.c_uint64()
        L = engine_lib()
        _check(L.hevcr_syntax_compact(C.byref(src), None, 0, C.byref(n), None), "hevcr_syntax_compact")
        self.sig = np.zeros(max(1, n.value), np.uint32)
        offs = np.zeros(3 * max(1, src.num_tus), np.uint32)
        _check(L.hevcr_syntax_compact(C.byref(src), self.sig.ctypes.data, self.sig.size, C.byref(n),
                                      offs.ctypes.data), "hevcr_syntax_compact")
        self.tus = np.frombuffer(C.string_at(src.tus, src.num_tus * C.sizeof(TU)) if src.num_tus else b"",
                                 np.uint8).copy().reshape(-1, C.sizeof(TU))
        if src.num_tus:  # coeff_off[3] are the last 12 bytes of a TU record
            self.tus[:, 16:28] = offs[:3 * src.num_tus].view(np.uint8).reshape(-1, 12)
        s = FrameSyntax.from_buffer_copy(src)
        s.tus = C.cast(self.tus.ctypes.data, C.POINTER(TU))
        s.coeffs, s.num_coeffs = None, 0
        s.sig, s.num_sig = C.cast(self.sig.ctypes.data, C.POINTER(C.c_uint32)), n.value
        s._owner = self  # the structure points into this object's buffers: keep them alive with it
        self.source, self.syntax = source, s


def plane_out_array(pic: Picture | None):
    if pic is None:
        return None
    arr = (C.c_void_p * 3)()
    for c in range(3):
        arr[c] = pic.planes[c].ctypes.data
    return arr


# ---------------------------------------------------------------------------
# Engine (libhevcr.so)
# ---------------------------------------------------------------------------

RUN_FULL, RUN_STOP_RECON, RUN_STOP_DEBLOCK = 0, 1, 2
STAGES = ("expand", "mc", "residual", "intra", "deblock", "sao")
KERNELS = ("k_expand", "k_mc", "k_mc8", "k_residual<4x4>", "k_residual<8x8>", "k_residual<16x16>",
           "k_residual<32x32>", "k_intra", "k_deblock", "k_sao", "k_loopfilter")

_engine_lib = None


def engine_lib() -> C.CDLL:
    global _engine_lib
    if _engine_lib is None:
        lib = _load("libhevcr.so")
        vp, i, u32, u64 = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64
        lib.hevcr_last_error.restype = C.c_char_p
        lib.hevcr_engine_create.argtypes = [i, C.POINTER(vp)]
        lib.hevcr_engine_destroy.argtypes = [vp]
        lib.hevcr_engine_destroy.restype = None
        lib.hevcr_engine_sync.argtypes = [vp]
        lib.hevcr_stream_create.argtypes = [vp, i, i, i, i, C.POINTER(i)]
        lib.hevcr_stream_destroy.argtypes = [vp, i]
        lib.hevcr_picture_upload.argtypes = [vp, i, i, vp, vp, vp]
        lib.hevcr_picture_download.argtypes = [vp, i, i, vp, vp, vp]
        lib.hevcr_decode_frame.argtypes = [vp, i, C.POINTER(FrameSyntax), i]
        lib.hevcr_frame_prepare.argtypes = [vp, i, C.POINTER(FrameSyntax), C.POINTER(vp)]
        lib.hevcr_frame_run.argtypes = [vp, vp, i, i]
        lib.hevcr_frame_release.argtypes = [vp, vp]
        lib.hevcr_frame_release.restype = None
        lib.hevcr_frame_bytes.argtypes = [vp, C.POINTER(C.c_double)]
        lib.hevcr_frame_kernel_bytes.argtypes = [vp, C.POINTER(C.c_double)]
        lib.hevcr_batch_kernel_bytes.argtypes = [vp, C.POINTER(C.c_double)]
        lib.hevcr_set_profiling.argtypes = [vp, i]
        lib.hevcr_stage_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(u64)]
        lib.hevcr_reset_stage_times.argtypes = [vp]
        lib.hevcr_kernel_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(u64)]
        lib.hevcr_kernel_name.argtypes = [i]
        lib.hevcr_kernel_name.restype = C.c_char_p
        lib.hevcr_residual_blocks.argtypes = [vp, i, u32, vp, vp, vp, vp, vp, vp, u64, vp]
        lib.hevcr_batch_create.argtypes = [vp, vp, vp, i, i, C.POINTER(vp)]
        lib.hevcr_batch_run.argtypes = [vp, vp, i]
        lib.hevcr_batch_release.argtypes = [vp, vp]
        lib.hevcr_batch_release.restype = None
        lib.hevcr_batch_bytes.argtypes = [vp, C.POINTER(C.c_double)]
        lib.hevcr_bench_copy.argtypes = [vp, u64, i]
        lib.hevcr_picture_digest.argtypes = [vp, i, vp, vp, vp]
        lib.hevcr_egress_size.argtypes = [i, i, i, i, vp, C.POINTER(u64)]
        lib.hevcr_picture_sha256.argtypes = [vp, i, vp, vp, vp, vp]
        lib.hevcr_residual_blocks_noalloc.argtypes = [vp, i, u32, vp, vp, vp, vp, vp, vp, u64, vp]
        lib.hevcr_tuning_set.argtypes = [vp, i, i, i, C.POINTER(Tuning)]
        lib.hevcr_tuning_get.argtypes = [vp, i, i, i, C.POINTER(Tuning)]
        lib.hevcr_autotune.argtypes = [vp, i, vp, vp, vp, i, C.POINTER(Tuning), C.POINTER(C.c_double)]
        lib.hevcr_picture_egress.argtypes = [vp, i, vp, vp, vp, i, vp, u64]
        lib.hevcr_host_alloc.argtypes = [u64, C.POINTER(vp)]
        lib.hevcr_host_free.argtypes = [vp]
        lib.hevcr_host_free.restype = None
        lib.hevcr_server_create.argtypes = [vp, C.POINTER(ServerCfg), C.POINTER(vp)]
        lib.hevcr_server_destroy.argtypes = [vp]
        lib.hevcr_server_destroy.restype = None
        lib.hevcr_server_clock.restype = C.c_double
        lib.hevcr_server_open.argtypes = [vp, i, i, i, i, C.POINTER(i)]
        lib.hevcr_server_close.argtypes = [vp, i]
        lib.hevcr_server_admitted.argtypes = [vp, i, C.POINTER(i)]
        lib.hevcr_server_submit.argtypes = [vp, vp, i, u64, C.c_double]
        lib.hevcr_server_submit_syntax.argtypes = [vp, i, C.POINTER(FrameSyntax), i, u64, C.c_double]
        lib.hevcr_syntax_compact.argtypes = [C.POINTER(FrameSyntax), vp, u64, C.POINTER(u64), vp]
        lib.hevcr_server_step.argtypes = [vp, C.POINTER(i)]
        lib.hevcr_server_poll.argtypes = [vp, i, C.POINTER(Done), i, C.POINTER(i)]
        lib.hevcr_server_info.argtypes = [vp, C.POINTER(ServerState)]
        _engine_lib = lib
    return _engine_lib


EGRESS_YUV, EGRESS_U16 = 0, 1
TUNE_CANDIDATES = 32


class Tuning(C.Structure):
    """hevcr_tuning: how frames of one size and bit depth are cut into work."""
    _fields_ = [("mc8", C.c_int32), ("lf_fused", C.c_int32), ("intra_group4", C.c_int32),
                ("intra_persistent", C.c_int32), ("mc_strips", C.c_int32)]

    def as_dict(self) -> dict:
        return {"mc8": self.mc8, "lf_fused": self.lf_fused, "intra_group4": self.intra_group4,
                "intra_persistent": self.intra_persistent, "mc_strips": self.mc_strips}

    @staticmethod
    def candidate(k: int) -> "Tuning":
        return Tuning(k & 1, (k >> 1) & 1, (k >> 2) & 1, (k >> 3) & 1, (k >> 4) & 1)



def egress_size(width: int, height: int, bit_depth: int, fmt: int, crop=None) -> int:
    """Bytes of one picture in an egress format (hevcr_egress_size)."""
    out = C.c_uint64()
    c = (C.c_int * 4)(*crop) if crop else None
    _check(engine_lib().hevcr_egress_size(width, height, bit_depth, fmt, c, C.byref(out)), "hevcr_egress_size")
    return int(out.value)


class PinnedBuffer:
    """Pinned host memory from hevcr_host_alloc: egress into it is written by
    the GPU directly. .array is a uint8 numpy view of it."""

    def __init__(self, nbytes: int):
        h = C.c_void_p()
        _check(engine_lib().hevcr_host_alloc(nbytes, C.byref(h)), "hevcr_host_alloc")
        self.addr, self.nbytes = h.value, nbytes
        self.array = np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(h.value))

    def close(self):
        if self.addr:
            self.array = None
            engine_lib().hevcr_host_free(self.addr)
            self.addr = None


class ServerCfg(C.Structure):
    _fields_ = [("dpb_budget_bytes", C.c_uint64), ("max_batch", C.c_int32), ("max_inflight", C.c_int32),
                ("lanes", C.c_int32), ("ingest_threads", C.c_int32)]


class Done(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("stream_id", C.c_int32), ("out_slot", C.c_int32),
                ("status", C.c_int32), ("step_frames", C.c_int32),
                ("t_submit", C.c_double), ("t_launch", C.c_double), ("t_done", C.c_double)]


class ServerState(C.Structure):
    _fields_ = [("streams", C.c_int32), ("admitted", C.c_int32), ("queued_frames", C.c_int32),
                ("inflight_steps", C.c_int32), ("budget_used", C.c_uint64), ("frames_done", C.c_uint64),
                ("steps_done", C.c_uint64), ("ingested_frames", C.c_uint64), ("ingest_bytes", C.c_uint64),
                ("ingest_busy_s", C.c_double), ("ingest_validate_s", C.c_double), ("ingest_lists_s", C.c_double),
                ("ingest_stage_s", C.c_double), ("ingest_upload_s", C.c_double)]


class EngineError(RuntimeError):
    def __init__(self, rc: int, what: str):
        msg = engine_lib().hevcr_last_error().decode()
        super().__init__(f"{what} failed ({rc}): {msg}")
        self.rc = rc


def _check(rc: int, what: str):
    if rc != 0:
        raise EngineError(rc, what)


class PreparedFrame:
    def __init__(self, engine: "Engine", handle):
        self.engine, self.handle = engine, handle

    def run(self, out_slot: int, stop: int = RUN_FULL):
        _check(engine_lib().hevcr_frame_run(self.engine.h, self.handle, out_slot, stop), "hevcr_frame_run")

    def bytes(self) -> dict:
        arr = (C.c_double * len(STAGES))()
        _check(engine_lib().hevcr_frame_bytes(self.handle, arr), "hevcr_frame_bytes")
        return dict(zip(STAGES, arr))

    def kernel_bytes(self) -> dict:
        arr = (C.c_double * len(KERNELS))()
        _check(engine_lib().hevcr_frame_kernel_bytes(self.handle, arr), "hevcr_frame_kernel_bytes")
        return dict(zip(KERNELS, arr))

    def release(self):
        if self.handle:
            engine_lib().hevcr_frame_release(self.engine.h, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Batch:
    """Prepared frames reconstructed together (one launch per stage)."""

    def __init__(self, engine: "Engine", frames: list, slots: list, stop: int = RUN_FULL):
        arr = (C.c_void_p * len(frames))(*[f.handle for f in frames])
        sl = (C.c_int * len(slots))(*slots)
        h = C.c_void_p()
        _check(engine_lib().hevcr_batch_create(engine.h, arr, sl, len(frames), stop, C.byref(h)),
               "hevcr_batch_create")
        self.engine, self.handle, self.frames = engine, h, frames

    def run(self, graph: bool = True):
        _check(engine_lib().hevcr_batch_run(self.engine.h, self.handle, int(graph)), "hevcr_batch_run")

    def bytes(self) -> dict:
        arr = (C.c_double * len(STAGES))()
        _check(engine_lib().hevcr_batch_bytes(self.handle, arr), "hevcr_batch_bytes")
        return dict(zip(STAGES, arr))

    def kernel_bytes(self) -> dict:
        arr = (C.c_double * len(KERNELS))()
        _check(engine_lib().hevcr_batch_kernel_bytes(self.handle, arr), "hevcr_batch_kernel_bytes")
        return dict(zip(KERNELS, arr))

    def release(self):
        if self.handle:
            engine_lib().hevcr_batch_release(self.engine.h, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Engine:
    """One reconstruction engine on one GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(engine_lib().hevcr_engine_create(device, C.byref(h)), "hevcr_engine_create")
        self.h = h

    def close(self):
        if self.h:
            engine_lib().hevcr_engine_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _check(engine_lib().hevcr_engine_sync(self.h), "hevcr_engine_sync")

    def stream(self, width: int, height: int, bit_depth: int, slots: int) -> int:
        sid = C.c_int()
        _check(engine_lib().hevcr_stream_create(self.h, width, height, bit_depth, slots, C.byref(sid)),
               "hevcr_stream_create")
        return sid.value

    def destroy_stream(self, sid: int):
        _check(engine_lib().hevcr_stream_destroy(self.h, sid), "hevcr_stream_destroy")

    def upload(self, sid: int, slot: int, pic: Picture):
        _check(engine_lib().hevcr_picture_upload(self.h, sid, slot, *pic.ptrs()), "hevcr_picture_upload")

    def download(self, sid: int, slot: int, width: int, height: int, bit_depth: int) -> Picture:
        pic = Picture(width, height, bit_depth)
        _check(engine_lib().hevcr_picture_download(self.h, sid, slot, *pic.ptrs()), "hevcr_picture_download")
        return pic

    def digests(self, pictures: list) -> list:
        """Device digests (Y, Cb, Cr) of [(stream id, slot), ...]."""
        n = len(pictures)
        sids = np.array([p[0] for p in pictures], np.int32)
        slots = np.array([p[1] for p in pictures], np.int32)
        out = np.zeros(3 * max(n, 1), np.uint64)
        _check(engine_lib().hevcr_picture_digest(self.h, n, sids.ctypes.data, slots.ctypes.data, out.ctypes.data),
               "hevcr_picture_digest")
        return [tuple(int(v) for v in out[3 * k:3 * k + 3]) for k in range(n)]

    def sha256(self, pictures: list, crops: list = None) -> list:
        """SHA-256 on the device of the write_yuv bytes of each plane of
        [(stream id, slot), ...] (crops: [left, right, top, bottom] per
        picture, or None): [[Y, Cb, Cr] hex digests, ...]."""
        n = len(pictures)
        sids = np.array([p[0] for p in pictures], np.int32)
        slots = np.array([p[1] for p in pictures], np.int32)
        cr = np.array(crops, np.int32).reshape(-1) if crops else None
        out = np.zeros(96 * max(n, 1), np.uint8)
        _check(engine_lib().hevcr_picture_sha256(self.h, n, sids.ctypes.data, slots.ctypes.data,
                                                 cr.ctypes.data if cr is not None else None, out.ctypes.data),
               "hevcr_picture_sha256")
        return [[out[96 * k + 32 * c:96 * k + 32 * c + 32].tobytes().hex() for c in range(3)] for k in range(n)]

    def egress(self, pictures: list, fmt: int = None, crops: list = None, out=None):
        """Pictures [(stream id, slot, width, height, bit depth), ...] serialised
        by hevcr_picture_egress (EGRESS_YUV: write_yuv bytes within each crop;
        EGRESS_U16: whole uint16 planes) one after another. out: None (a new
        numpy buffer), a numpy uint8 array, a PinnedBuffer, or (device
        address, bytes). Returns the bytes written as a uint8 numpy array (the
        view of out, or None for a device address)."""
        fmt = EGRESS_YUV if fmt is None else fmt
        n = len(pictures)
        total = sum(egress_size(w, h, bd, fmt, crops[k] if crops else None)
                    for k, (_, _, w, h, bd) in enumerate(pictures))
        sids = np.array([p[0] for p in pictures], np.int32)
        slots = np.array([p[1] for p in pictures], np.int32)
        cr = np.array(crops, np.int32).reshape(-1) if crops else None
        if out is None:
            out = np.empty(max(total, 1), np.uint8)
        if isinstance(out, PinnedBuffer):
            addr, cap, view = out.addr, out.nbytes, out.array
        elif isinstance(out, tuple):
            addr, cap, view = out[0], out[1], None
        else:
            addr, cap, view = out.ctypes.data, out.nbytes, out
        _check(engine_lib().hevcr_picture_egress(self.h, n, sids.ctypes.data, slots.ctypes.data,
                                                 cr.ctypes.data if cr is not None else None, fmt, addr, cap),
               "hevcr_picture_egress")
        return view[:total] if view is not None else None

    def set_tuning(self, width: int, height: int, bit_depth: int, mc8=1, lf_fused=1, intra_group4=1,
                   intra_persistent=1, mc_strips=1):
        t = Tuning(mc8, lf_fused, intra_group4, intra_persistent, mc_strips)
        _check(engine_lib().hevcr_tuning_set(self.h, width, height, bit_depth, C.byref(t)), "hevcr_tuning_set")

    def tuning(self, width: int, height: int, bit_depth: int) -> dict:
        t = Tuning()
        _check(engine_lib().hevcr_tuning_get(self.h, width, height, bit_depth, C.byref(t)), "hevcr_tuning_get")
        return t.as_dict()

    def autotune(self, sids: list, syntaxes: list, slots: list, reps: int = 10):
        """hevcr_autotune over these frames (one size and bit depth): (the
        chosen tuning, ms per replay of each candidate)."""
        n = len(syntaxes)
        sid_a = np.array(sids, np.int32)
        slot_a = np.array(slots, np.int32)
        ptrs = (C.c_void_p * n)(*[C.addressof(s) for s in syntaxes])
        best = Tuning()
        ms = (C.c_double * TUNE_CANDIDATES)()
        _check(engine_lib().hevcr_autotune(self.h, n, sid_a.ctypes.data, ptrs, slot_a.ctypes.data, reps,
                                           C.byref(best), ms), "hevcr_autotune")
        return best.as_dict(), list(ms)

    def decode(self, sid: int, syntax: FrameSyntax, out_slot: int):
        _check(engine_lib().hevcr_decode_frame(self.h, sid, C.byref(syntax), out_slot), "hevcr_decode_frame")

    def prepare(self, sid: int, syntax: FrameSyntax) -> PreparedFrame:
        h = C.c_void_p()
        _check(engine_lib().hevcr_frame_prepare(self.h, sid, C.byref(syntax), C.byref(h)), "hevcr_frame_prepare")
        return PreparedFrame(self, h)

    def batch(self, frames: list, slots: list, stop: int = RUN_FULL) -> Batch:
        return Batch(self, frames, slots, stop)

    def profiling(self, on: bool):
        _check(engine_lib().hevcr_set_profiling(self.h, int(on)), "hevcr_set_profiling")

    def stage_times(self) -> dict:
        ms = (C.c_double * len(STAGES))()
        n = (C.c_uint64 * len(STAGES))()
        _check(engine_lib().hevcr_stage_times(self.h, ms, n), "hevcr_stage_times")
        return {s: (ms[k], n[k]) for k, s in enumerate(STAGES)}

    def kernel_times(self) -> dict:
        """{kernel name: (ms, launches)} accumulated since the last reset (profiling on)."""
        n = len(KERNELS)
        ms = (C.c_double * n)()
        cnt = (C.c_uint64 * n)()
        _check(engine_lib().hevcr_kernel_times(self.h, ms, cnt), "hevcr_kernel_times")
        return {engine_lib().hevcr_kernel_name(k).decode(): (ms[k], cnt[k]) for k in range(n)}

    def reset_stage_times(self):
        _check(engine_lib().hevcr_reset_stage_times(self.h), "hevcr_reset_stage_times")

    def residual_blocks(self, bit_depth, log2, cidx, qp, flags, off, coeffs) -> np.ndarray:
        n = len(log2)
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in (log2, cidx, qp, flags)]
        off = np.ascontiguousarray(off, dtype=np.uint32)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.int16)
        out = np.zeros_like(coeffs)
        _check(engine_lib().hevcr_residual_blocks(self.h, bit_depth, n, *[a.ctypes.data for a in arrs],
                                                  off.ctypes.data, coeffs.ctypes.data, coeffs.size,
                                                  out.ctypes.data), "hevcr_residual_blocks")
        return out


class Server:
    """Continuous-batching request loop over one engine (hevcr_server_*)."""

    def __init__(self, engine: Engine, dpb_budget_bytes: int, max_batch: int = 64, max_inflight: int = 2,
                 lanes: int = 1, ingest_threads: int = 0):
        cfg = ServerCfg(dpb_budget_bytes, max_batch, max_inflight, lanes, ingest_threads)
        h = C.c_void_p()
        _check(engine_lib().hevcr_server_create(engine.h, C.byref(cfg), C.byref(h)), "hevcr_server_create")
        self.engine, self.h = engine, h
        self._buf = (Done * 1024)()
        self._held = {}  # tag -> prepared frame, kept alive until its completion is polled

    @staticmethod
    def clock() -> float:
        return engine_lib().hevcr_server_clock()

    def open(self, width: int, height: int, bit_depth: int, dpb_slots: int) -> int:
        sid = C.c_int()
        _check(engine_lib().hevcr_server_open(self.h, width, height, bit_depth, dpb_slots, C.byref(sid)),
               "hevcr_server_open")
        return sid.value

    def close_stream(self, sid: int):
        _check(engine_lib().hevcr_server_close(self.h, sid), "hevcr_server_close")

    def admitted(self, sid: int) -> bool:
        a = C.c_int()
        _check(engine_lib().hevcr_server_admitted(self.h, sid, C.byref(a)), "hevcr_server_admitted")
        return bool(a.value)

    def submit(self, frame: PreparedFrame, out_slot: int, tag: int, t_arrival: float = -1.0):
        _check(engine_lib().hevcr_server_submit(self.h, frame.handle, out_slot, tag, t_arrival), "hevcr_server_submit")
        self._held[tag] = frame

    def submit_syntax(self, sid: int, owner, out_slot: int, tag: int, t_arrival: float = -1.0):
        """A frame as host syntax (owner.syntax, a FrameSyntax, kept alive until
        its completion is polled), prepared by the server's ingest threads."""
        _check(engine_lib().hevcr_server_submit_syntax(self.h, sid, C.byref(owner.syntax), out_slot, tag, t_arrival),
               "hevcr_server_submit_syntax")
        self._held[tag] = owner

    def step(self) -> int:
        n = C.c_int()
        _check(engine_lib().hevcr_server_step(self.h, C.byref(n)), "hevcr_server_step")
        return n.value

    def poll(self, wait: bool = False) -> list:
        n = C.c_int()
        _check(engine_lib().hevcr_server_poll(self.h, int(wait), self._buf, len(self._buf), C.byref(n)),
               "hevcr_server_poll")
        out = []
        for k in range(n.value):
            d = self._buf[k]
            out.append(dict(tag=d.tag, stream_id=d.stream_id, out_slot=d.out_slot, status=d.status,
                            step_frames=d.step_frames, t_submit=d.t_submit, t_launch=d.t_launch, t_done=d.t_done))
            self._held.pop(d.tag, None)
        return out

    def info(self) -> dict:
        st = ServerState()
        _check(engine_lib().hevcr_server_info(self.h, C.byref(st)), "hevcr_server_info")
        return {k: getattr(st, k) for k, _ in ServerState._fields_}

    def close(self):
        if self.h:
            engine_lib().hevcr_server_destroy(self.h)
            self.h = None
            self._held.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def decode_on_engine(engine: Engine, frame: Frame, stop: int = RUN_FULL, syntax: FrameSyntax = None) -> Picture:
    """Reconstructs one generated frame on the GPU: its reference pictures go to
    DPB slots 0..k-1 and the output to slot k. syntax: another form of the
    frame's syntax (the compact one), else the frame's own."""
    k = frame.num_ref_pics
    sid = engine.stream(frame.width, frame.height, frame.bit_depth, k + 1)
    try:
        for slot in range(k):
            engine.upload(sid, slot, frame.refs[slot])
        pf = engine.prepare(sid, syntax if syntax is not None else frame.syntax)
        pf.run(k, stop)
        pf.release()
        engine.sync()
        return engine.download(sid, k, frame.width, frame.height, frame.bit_depth)
    finally:
        engine.destroy_stream(sid)
