"""Test-side access to the CPU oracle (oracle/liboracle.so): the checker the
HIP engine is compared against. Only tests and the benchmark's CPU-baseline
leg import this module."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os

import numpy as np

import hevcr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built: run __graft_entry__.build() first")
        L = C.CDLL(ORACLE_SO)
        L.oracle_recon_frame.argtypes = [C.POINTER(hevcr.FrameSyntax), C.c_void_p, C.c_int,
                                         C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_residual_block.argtypes = [C.c_int] * 6 + [C.c_void_p, C.c_void_p]
        L.oracle_filters.argtypes = [C.POINTER(hevcr.FrameSyntax), C.c_void_p, C.c_int, C.c_int]
        _lib = L
    return _lib


def recon(frame: hevcr.Frame, L=None, fn="oracle_recon_frame"):
    """[pre-deblock, post-deblock, post-SAO] pictures of one frame."""
    L = L or lib()
    outs = [frame.new_picture() for _ in range(3)]
    rc = getattr(L, fn)(C.byref(frame.syntax), frame.ref_plane_array(), frame.num_ref_pics,
                        *[hevcr.plane_out_array(o) for o in outs])
    if rc != 0:
        raise RuntimeError(f"{fn} returned {rc}")
    return outs


def recon_syntax(syntax: hevcr.FrameSyntax, dpb: list, L=None, fn="oracle_recon_frame"):
    """[pre-deblock, post-deblock, post-SAO] pictures of one frame whose
    reference lists name slots of dpb (a list of Pictures, None for empty)."""
    L = L or lib()
    p = syntax.p
    outs = [hevcr.Picture(p.width, p.height, p.bit_depth) for _ in range(3)]
    blank = hevcr.Picture(p.width, p.height, p.bit_depth)
    arr = (C.c_void_p * max(1, 3 * len(dpb)))()
    for k, r in enumerate(dpb):
        for c in range(3):
            arr[3 * k + c] = (r or blank).planes[c].ctypes.data
    rc = getattr(L, fn)(C.byref(syntax), arr, len(dpb), *[hevcr.plane_out_array(o) for o in outs])
    if rc != 0:
        raise RuntimeError(f"{fn} returned {rc}")
    return outs


def residual_block(log2, cidx, intra, qp, bd, tskip, coeffs) -> np.ndarray:
    c = np.ascontiguousarray(coeffs, dtype=np.int16)
    out = np.zeros(1 << (2 * log2), np.int16)
    rc = lib().oracle_residual_block(log2, cidx, intra, qp, bd, tskip, c.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_residual_block failed")
    return out


def syntax_digest(frame: hevcr.Frame) -> str:
    """SHA-256 over the generated syntax (parameters + every record array)."""
    s = frame.syntax
    h = hashlib.sha256()
    h.update(C.string_at(C.addressof(s.p), C.sizeof(s.p)))
    for ptr, n, t in ((s.cus, s.num_cus, hevcr.CU), (s.pus, s.num_pus, hevcr.PU), (s.tus, s.num_tus, hevcr.TU),
                      (s.coeffs, s.num_coeffs, C.c_int16)):
        if n:
            h.update(C.string_at(ptr, n * C.sizeof(t)))
    w = (s.p.width + (1 << s.p.ctb_log2) - 1) >> s.p.ctb_log2
    hh = (s.p.height + (1 << s.p.ctb_log2) - 1) >> s.p.ctb_log2
    h.update(C.string_at(s.sao, w * hh * C.sizeof(hevcr.SAO)))
    if s.scaling:
        h.update(C.string_at(s.scaling, C.sizeof(hevcr.Scaling)))
    if s.slices:
        h.update(C.string_at(s.slices, s.num_slices * C.sizeof(hevcr.Slice)))
        h.update(C.string_at(s.ctb_slice, w * hh))
    if s.tiles:
        h.update(C.string_at(s.tiles, C.sizeof(hevcr.Tiles)))
    if s.pcm_loop_filter_disabled:
        h.update(b"pcm_loop_filter_disabled")
    for r in frame.refs[:frame.num_ref_pics]:
        for p in r.planes:
            h.update(p.tobytes())
    return h.hexdigest()


def load_goldens() -> list[dict]:
    with open(os.path.join(GOLDEN_DIR, "frames.json")) as f:
        return json.load(f)["cases"]


def make_frame(case: dict) -> hevcr.Frame:
    f = hevcr.Frame(hevcr.preset(case["preset"], **case.get("overrides", {})), case["seed"])
    for k, v in case.get("params", {}).items():
        setattr(f.syntax.p, k, v)
    return f
