"""Writes the coded-stream golden cases of tests/golden/coded/: HEVC streams of
random content written by tests/golden/stream_writer.cpp over slice segments
(dependent ones included), tiles, wavefront entry points, PCM, transquant
bypass, weighted prediction per segment and the other picture and slice
options, each with the MD5 of the YUV the reference decoder's CLI
(oracle/_ref/hevc-decode, built by `make -C oracle ref` from the reference
sources) writes for it. Run in the build container:

    python tests/golden/make_coded_goldens.py

Outputs coded/<name>.265 and coded.json (per stream: the writer's arguments,
size, bit depth, conformance window, picture count, the reference's YUV MD5).

The streams are what the reference decoder reads as intended; where its
reading differs from H.265 (the alignment after a terminating bin that ends
on a byte boundary, dependent segment headers, no context reset at tile
starts), the writer follows the reference, as the front end does. The
reference's wavefront path is not exercised with more than one slice segment
per picture or a picture one CTB wide: it decodes every row from a segment's
first to the picture's last (coding_tree.cpp:404-528) and waits for a second
CTB that a one-CTB-wide picture does not have.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT_DIR = os.path.join(HERE, "coded")
BUILD = os.path.join(HERE, "_build")
WRITER = os.path.join(BUILD, "stream-writer")
REF = os.path.join(ROOT, "oracle", "_ref", "hevc-decode")
FRONT = os.path.join(ROOT, "hevc.js_amd", "csrc", "front")

# name -> writer arguments (see the Config keys of stream_writer.cpp)
CASES = {
    "slices_dep_p": "seed=11 pics=4 segs=4 pdep=0.5 types=p refs=2",
    "slices_types_wp": "seed=12 pics=4 segs=3 pdep=0.3 types=pb refs=3 wp=1 wbp=1 pseg_i=0.3 nref0=2 nref1=2",
    "tiles3x2_lf": "seed=13 pics=3 segs=3 tcols=3 trows=2 lf_tiles=1 types=b",
    "tiles4x2_explicit_nolf": "seed=14 pics=3 segs=4 pdep=0.5 tcols=4 trows=2 tuniform=0 lf_tiles=0 lf_slices=0 types=pb",
    "pcm_slices": "seed=15 pics=3 segs=3 pdep=0.5 pcm=1 pcm_bits=7 pcm_bits_c=6 p_pcm=0.4 types=pb",
    "pcm_nolf_tiles": "seed=16 pics=3 segs=2 pcm=1 pcm_lf_disabled=1 pcm_min=4 tcols=2 trows=2 p_pcm=0.5 types=pb",
    "bypass_qpdelta_sdh": "seed=17 pics=3 segs=3 tqb=1 p_bypass=0.3 cuqpd=1 qpd_depth=2 sdh=1 tskip=1 types=pb",
    "ctb16_10bit_pcm_wp": "seed=18 pics=4 segs=4 pdep=0.5 bd=10 ctb=4 maxtb=4 w=200 h=136 pcm=1 pcm_bits=9 "
                          "pcm_max=4 wp=1 wbp=1 types=pb",
    "ra_listmod_pml": "seed=19 pics=5 segs=3 gop=ra refs=4 listmod=1 nref0=3 nref1=2 pml=3 types=b",
    "ctb64_crop_scaling": "seed=20 pics=3 segs=3 ctb=6 w=416 h=240 crop_r=4 crop_b=4 scaling=1 amp=1 pml=4 "
                          "dbover=1 types=pb",
    "deblock_sao_variants": "seed=21 pics=3 segs=5 pdep=0.4 dbctl=1 dbover=1 beta=3 tc=-2 lf_slices=1 "
                            "cb=2 cr=-2 qp=34 types=pb",
    "wpp_one_segment": "seed=22 pics=3 segs=1 wpp=1 types=pb",
    "dependent_aligned_header": "seed=1 pics=2 segs=5 pdep=1 w=64 h=136 types=p",
    "intra_slices": "seed=23 pics=2 segs=4 pdep=0.5 types=i p_intra=1",
    "tiles_slices_832x480": "seed=24 pics=2 segs=4 pdep=0.3 w=832 h=480 ctb=6 tcols=3 trows=2 tuniform=0 "
                            "pcm=1 tqb=1 types=pb",
}


def build_writer():
    os.makedirs(BUILD, exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    "-I" + FRONT, "-o", WRITER, os.path.join(HERE, "stream_writer.cpp"),
                    os.path.join(FRONT, "params.cpp")], check=True)


def stream_info(data: bytes) -> dict:
    """Size, bit depth, crop and picture count from the front end's parse."""
    sys.path.insert(0, os.path.join(ROOT, "hevc.js_amd"))
    import hevcr
    P = hevcr.Parser()
    info, n = None, 0
    try:
        P.feed(data)
        P.flush()
        for e in P.events():
            if e.kind == hevcr.EVENT_DECODE:
                n += 1
                if info is None:
                    info = {"width": e.width, "height": e.height, "bit_depth": e.bit_depth, "crop": list(e.crop)}
    finally:
        P.close()
    info["pictures"] = n
    return info


def main():
    build_writer()
    os.makedirs(OUT_DIR, exist_ok=True)
    out = []
    for name, args in CASES.items():
        path = os.path.join(OUT_DIR, name + ".265")
        subprocess.run([WRITER, path] + args.split(), check=True, capture_output=True)
        with tempfile.NamedTemporaryFile(suffix=".yuv") as tf:
            subprocess.run([REF, "-o", tf.name, path], check=True, capture_output=True, timeout=300)
            md5 = hashlib.md5(open(tf.name, "rb").read()).hexdigest()
        data = open(path, "rb").read()
        entry = {"name": name, "writer_args": args, "bytes": len(data), "md5_yuv_reference": md5}
        entry.update(stream_info(data))
        out.append(entry)
        print(f"{name}: {len(data)} bytes, {entry['pictures']} pictures, reference YUV md5 {md5}", flush=True)
    with open(os.path.join(HERE, "coded.json"), "w") as fh:
        json.dump({"generator": "tests/golden/make_coded_goldens.py", "writer": "tests/golden/stream_writer.cpp",
                   "reference": "oracle/_ref/hevc-decode (the reference decoder's CLI, src/main.cpp)",
                   "streams": out}, fh, indent=1)


if __name__ == "__main__":
    main()
