"""Writes the golden vectors of tests/golden from the reference decoder's own
reconstruction code (oracle/_ref/libhevcref_harness.so, built by
`make -C oracle ref` from the reference sources). Run in the build container:

    python tests/golden/make_goldens.py

Outputs:
  frames.json       per case: generator settings, syntax digest, SHA-256 of the
                    Y/Cb/Cr planes after reconstruction, deblocking and SAO
  residual_kat.npz  random coefficient blocks with the reference's
                    perform_dequant + perform_transform_inverse outputs
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hevc.js_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, HERE)

import hevcr  # noqa: E402
import oracle_lib  # noqa: E402
from cases import CASES  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "libhevcref_harness.so")


def main():
    ref = C.CDLL(HARNESS)
    ref.ref_recon_frame.argtypes = [C.POINTER(hevcr.FrameSyntax), C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
    ref.ref_residual_block.argtypes = [C.c_int] * 6 + [C.c_void_p, C.c_void_p]
    out = []
    for case in CASES:
        name, preset, seed, over = case[:4]
        params = case[4] if len(case) > 4 else {}
        c = {"name": name, "preset": preset, "seed": seed, "overrides": over, "params": params}
        f = oracle_lib.make_frame(c)
        pics = oracle_lib.recon(f, ref, "ref_recon_frame")
        c["syntax_sha256"] = oracle_lib.syntax_digest(f)
        c["width"], c["height"], c["bit_depth"] = f.width, f.height, f.bit_depth
        c["stages"] = {st: p.sha256() for st, p in zip(("recon", "deblock", "sao"), pics)}
        out.append(c)
        print(name, c["stages"]["sao"][0][:16], flush=True)
    with open(os.path.join(HERE, "frames.json"), "w") as fh:
        json.dump({"generator": "tests/golden/make_goldens.py", "source": "reference reconstruction code",
                   "cases": out}, fh, indent=1)

    # Residual KATs: every size, component, intra/inter, transform skip, 8 and 10 bit.
    rng = np.random.default_rng(20261203)
    rows = []
    coeffs, outs = [], []
    off = 0
    for bd in (8, 10):
        for log2 in (2, 3, 4, 5):
            for cidx in (0, 1):
                for intra in (0, 1):
                    for tskip in ((0, 1) if log2 == 2 else (0,)):
                        for _ in range(6):
                            n = 1 << log2
                            qp = int(rng.integers(0, 52 + 6 * (bd - 8)))
                            kind = rng.integers(0, 3)
                            if kind == 0:  # sparse small levels
                                c = (rng.laplace(0, 2, n * n) * (rng.random(n * n) < 0.3)).astype(np.int16)
                            elif kind == 1:  # dense large levels: clipping paths
                                c = rng.integers(-32768, 32768, n * n).astype(np.int16)
                            else:
                                c = rng.integers(-300, 301, n * n).astype(np.int16)
                            r = np.zeros(n * n, np.int16)
                            ref.ref_residual_block(log2, cidx, intra, qp, bd, tskip, c.ctypes.data, r.ctypes.data)
                            rows.append((bd, log2, cidx, intra, qp, tskip, off))
                            coeffs.append(c)
                            outs.append(r)
                            off += n * n
    np.savez_compressed(os.path.join(HERE, "residual_kat.npz"), blocks=np.array(rows, np.int32),
                        coeffs=np.concatenate(coeffs), residual=np.concatenate(outs))
    print("residual KATs:", len(rows))


if __name__ == "__main__":
    main()
