"""Real-stream golden vectors (tests/golden/streams.json + streams/*.npz,
written by tests/golden/make_stream_goldens.py from the reference decoder run
on the reference's conformance bitstreams): loading and output hashing shared
by the oracle and engine tests."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

import hevcr

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAGES = ("recon", "deblock", "sao")


def load() -> list[dict]:
    with open(os.path.join(GOLDEN_DIR, "streams.json")) as f:
        return json.load(f)["streams"]


def supported() -> list[dict]:
    return [s for s in load() if "unsupported" not in s]


def frames(stream: dict):
    """Yields (frame entry, SyntaxArrays) for each picture in decoding order."""
    z = np.load(os.path.join(GOLDEN_DIR, "streams", stream["name"] + ".npz"))
    params, cus, pus, tus, sao, coeffs = z["params"], z["cus"], z["pus"], z["tus"], z["sao"], z["coeffs"]
    co, po, to, so, ko = z["cu_off"], z["pu_off"], z["tu_off"], z["sao_off"], z["coeff_off"]
    sc, sco = z["scaling"], z["scaling_off"]
    for k, fr in enumerate(stream["frames"]):
        if int(z["slot"][k]) != fr["slot"]:
            raise ValueError("stream fixture and index disagree")
        yield fr, hevcr.SyntaxArrays(params[k].tobytes(), cus[co[k] * 16:co[k + 1] * 16],
                                     pus[po[k] * 20:po[k + 1] * 20], tus[to[k] * 28:to[k + 1] * 28],
                                     coeffs[ko[k]:ko[k + 1]], sao[so[k] * 24:so[k + 1] * 24],
                                     sc[sco[k]:sco[k + 1]])


def yuv_md5(stream: dict, outputs: list) -> str:
    """MD5 of the decoded pictures written as the reference's write_yuv does
    (picture.cpp:42-99): POC order within each coded video sequence, cropped
    to the conformance window. outputs[k] is the picture of frame k."""
    h = hashlib.md5()
    cl, cr, ct, cb = stream["crop"]
    order = sorted(range(len(outputs)), key=lambda k: (stream["frames"][k]["cvs"], stream["frames"][k]["poc"]))
    for k in order:
        pic = outputs[k]
        for c in range(3):
            s = 1 if c == 0 else 2
            p = pic.planes[c]
            p = p[ct // s:p.shape[0] - cb // s, cl // s:p.shape[1] - cr // s]
            h.update(p.astype(np.uint8).tobytes() if pic.bit_depth <= 8 else p.astype("<u2").tobytes())
    return h.hexdigest()
