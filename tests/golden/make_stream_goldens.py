"""Writes the real-stream golden vectors of tests/golden/streams/ from the
reference decoder run on the reference's own conformance bitstreams
(tests/conformance/fixtures/*.265, x265 encodes; their README holds the MD5
of each stream's decoded YUV as computed by ffmpeg).

oracle/_ref/hevc-tap (built by `make -C oracle ref` from the reference
sources) decodes each bitstream with the reference decoder and records, per
picture, the reconstruction syntax its parser handed to the reconstruction
operators (as hevcr_frame_syntax records) together with the picture after
reconstruction, after deblocking and after SAO. Run in the build container:

    python tests/golden/make_stream_goldens.py

Outputs:
  streams/<name>.npz  per picture: hevcr_frame_params (reference lists mapped
                      to DPB slots), CU/PU/TU/SAO records, coefficient pool,
                      scaling factors (when the stream enables them),
                      the DPB slot the picture is written to
  streams.json        per stream: size, bit depth, DPB slots, conformance
                      window, the README's MD5 of the decoded YUV, and per
                      picture its POC and the SHA-256 of each plane after each
                      stage; streams the engine's syntax cannot describe are
                      listed with the reason
"""
import json
import os
import re
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hevc.js_amd"))

import hevcr  # noqa: E402

FIXTURES = "/root/reference/tests/conformance/fixtures"
TAP = os.path.join(ROOT, "oracle", "_ref", "hevc-tap")
OUT_DIR = os.path.join(HERE, "streams")
MAX_SLOTS = 16
# Header written by oracle/ref_tap.cpp before each picture (136 bytes).
HDR = struct.Struct("<4s i i i i i i 4i I I I I I Q 64s")
STAGES = ("recon", "deblock", "sao")


def readme_md5() -> dict:
    """MD5 of the decoded YUV per bitstream, from the fixtures' README table."""
    md5 = {}
    with open(os.path.join(FIXTURES, "README.md"), encoding="utf-8") as f:
        for line in f:
            m = re.match(r"\|\s*`([\w.]+\.265)`.*`([0-9a-f]{32})`\s*\|\s*$", line)
            if m:
                md5[m.group(1)] = m.group(2)
    return md5


def read_tap(path: str) -> list:
    data = open(path, "rb").read()
    o, pics = 0, []
    psz = hevcr.C.sizeof(hevcr.FrameParams)
    while o < len(data):
        (magic, pid, cvs, status, w, h, bd, cl, cr, ct, cb, ncu, npu, ntu, nsao, has_scaling, ncoef,
         reason) = HDR.unpack_from(data, o)
        if magic != b"HTAP":
            raise ValueError(f"{path}: bad record at {o}")
        o += HDR.size
        pic = {"pic_id": pid, "cvs": cvs, "status": status, "reason": reason.split(b"\0")[0].decode(),
               "width": w, "height": h, "bit_depth": bd, "crop": [cl, cr, ct, cb]}
        if status == 0:
            def take(n):
                nonlocal o
                b = data[o:o + n]
                o += n
                return b
            pic["params"] = take(psz)
            pic["cus"] = take(ncu * 16)
            pic["pus"] = take(npu * 20)
            pic["tus"] = take(ntu * 28)
            pic["coeffs"] = np.frombuffer(take(ncoef * 2), np.int16)
            pic["sao"] = take(nsao * 24)
            pic["scaling"] = take(hevcr.C.sizeof(hevcr.Scaling)) if has_scaling else b""
        planes = []
        for _st in STAGES:
            st = []
            for c in range(3):
                pw, ph = (w, h) if c == 0 else (w // 2, h // 2)
                st.append(np.frombuffer(data, np.uint16, pw * ph, o).reshape(ph, pw))
                o += pw * ph * 2
            planes.append(st)
        pic["planes"] = planes
        pics.append(pic)
    return pics


def assign_slots(pics: list) -> int:
    """DPB slot per picture: a picture's slot is reused once no later picture
    references it. Writes the slot into each picture and maps the reference
    lists (decoding indices) to slots; returns the number of slots."""
    last_use = {}
    for k, p in enumerate(pics):
        fp = hevcr.FrameParams.from_buffer_copy(p["params"])
        p["fp"] = fp
        for l in range(2):
            for i in range(fp.num_ref[l]):
                last_use[fp.ref_slot[l][i]] = k
    holder = {}  # slot -> picture id it holds
    slot_of = {}
    for k, p in enumerate(pics):
        fp = p["fp"]
        for l in range(2):
            for i in range(fp.num_ref[l]):
                fp.ref_slot[l][i] = slot_of[fp.ref_slot[l][i]]
        free = [s for s in range(MAX_SLOTS) if s not in holder or last_use.get(holder[s], -1) < k]
        if not free:
            raise ValueError("more than %d pictures in use" % MAX_SLOTS)
        slot = free[0]
        holder[slot] = p["pic_id"]
        slot_of[p["pic_id"]] = slot
        p["slot"] = slot
        p["params"] = bytes(fp)
    return max(p["slot"] for p in pics) + 1


def yuv_md5(pics: list) -> str:
    """MD5 of the decoded pictures written as the reference's write_yuv does
    (picture.cpp:42-99): output order (POC within each coded video sequence),
    cropped to the conformance window."""
    import hashlib
    h = hashlib.md5()
    for p in sorted(pics, key=lambda q: (q["cvs"], hevcr.FrameParams.from_buffer_copy(q["params"]).poc
                                         if "params" in q else q["pic_id"])):
        cl, cr, ct, cb = p["crop"]
        for c in range(3):
            s = 1 if c == 0 else 2
            pl = p["planes"][2][c][ct // s:p["height"] // s - cb // s, cl // s:p["width"] // s - cr // s]
            h.update(pl.astype(np.uint8).tobytes() if p["bit_depth"] <= 8 else pl.astype("<u2").tobytes())
    return h.hexdigest()


def main():
    md5 = readme_md5()
    os.makedirs(OUT_DIR, exist_ok=True)
    out = []
    for name in sorted(f for f in os.listdir(FIXTURES) if f.endswith(".265")):
        with tempfile.NamedTemporaryFile(suffix=".tap") as tf:
            subprocess.run([TAP, os.path.join(FIXTURES, name), tf.name], check=True, capture_output=True)
            pics = read_tap(tf.name)
        entry = {"name": name[:-4], "bitstream": "tests/conformance/fixtures/" + name, "pictures": len(pics),
                 "md5_yuv_readme": md5.get(name)}
        bad = sorted({p["reason"] for p in pics if p["status"]})
        first = pics[0]
        entry.update(width=first["width"], height=first["height"], bit_depth=first["bit_depth"], crop=first["crop"])
        if bad:
            entry["unsupported"] = "; ".join(bad)
            out.append(entry)
            print(name, "unsupported:", entry["unsupported"], flush=True)
            continue
        slots = assign_slots(pics)
        entry["md5_yuv_reference"] = yuv_md5(pics)
        entry["dpb_slots"] = slots
        entry["frames"] = []
        for p in pics:
            fp = hevcr.FrameParams.from_buffer_copy(p["params"])
            entry["frames"].append({
                "poc": fp.poc, "cvs": p["cvs"], "slot": p["slot"], "slice_type": fp.slice_type,
                "sha256": {st: hevcr.Picture(p["width"], p["height"], p["bit_depth"], p["planes"][i]).sha256()
                           for i, st in enumerate(STAGES)}})

        def cat(key):
            return np.frombuffer(b"".join(p[key] for p in pics), np.uint8)

        def offs(key, size):
            return np.cumsum([0] + [len(p[key]) // size for p in pics]).astype(np.int64)

        np.savez_compressed(
            os.path.join(OUT_DIR, entry["name"] + ".npz"),
            params=np.stack([np.frombuffer(p["params"], np.uint8) for p in pics]),
            slot=np.array([p["slot"] for p in pics], np.int32),
            cus=cat("cus"), cu_off=offs("cus", 16), pus=cat("pus"), pu_off=offs("pus", 20),
            tus=cat("tus"), tu_off=offs("tus", 28), sao=cat("sao"), sao_off=offs("sao", 24),
            scaling=cat("scaling"), scaling_off=offs("scaling", 1),
            coeffs=np.concatenate([p["coeffs"] for p in pics]),
            coeff_off=np.cumsum([0] + [len(p["coeffs"]) for p in pics]).astype(np.int64))
        if entry["md5_yuv_readme"] is None:
            verdict = "(no README entry)"
        elif entry["md5_yuv_reference"] == entry["md5_yuv_readme"]:
            verdict = "matches the README"
        else:
            verdict = "DIFFERS from the README"
        print(name, len(pics), "pictures,", slots, "slots, reference YUV md5", verdict, flush=True)
        out.append(entry)
    with open(os.path.join(HERE, "streams.json"), "w") as fh:
        json.dump({"generator": "tests/golden/make_stream_goldens.py", "tap": "oracle/ref_tap.cpp",
                   "source": "reference decoder on the reference's conformance bitstreams", "streams": out},
                  fh, indent=1)


if __name__ == "__main__":
    main()
