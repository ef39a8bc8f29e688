"""Batched workloads: independent frames reconstructed in one batch.

  cfg2     32 independent 1920x1080 8-bit frames, 8 intra + 24 B sharing four
           seeded reference pictures;
  inter32  32 independent 1080p B frames sharing four reference pictures;
  cfg3     8 independent 3840x2160 10-bit frames, 2 intra + 6 B with four
           device-resident reference pictures (levels x4, QP 17..37, MV sigma
           24 px).

Member (preset, seed) pairs that also have a golden case name it in
`goldens`; the benchmark and the tests check those frames against the
reference's hashes.
"""
from __future__ import annotations

import hevcr

BATCHES = {
    "cfg2": dict(intra=(2, 8), inter=(1, 24), ref_seed=1, goldens={(2, 1): "intra1080_s1", (1, 1): "b1080_s1"},
                 desc="32 independent 1920x1080 8-bit 4:2:0 frames (8 intra + 24 B sharing 4 reference "
                      "pictures), batched form: one launch per stage over all frames, replayed from a hipGraph"),
    "inter32": dict(intra=(2, 0), inter=(1, 32), ref_seed=1, goldens={(1, 1): "b1080_s1"},
                    desc="32 independent 1920x1080 8-bit 4:2:0 B frames sharing 4 reference pictures, "
                         "batched form, hipGraph replay"),
    "cfg3": dict(intra=(4, 2), inter=(3, 6), ref_seed=1,
                 goldens={(4, 1): "intra2160_10bit_s1", (3, 1): "b2160_10bit_s1"},
                 desc="8 independent 3840x2160 10-bit 4:2:0 frames (2 intra + 6 B with 4 device-resident "
                      "reference pictures; levels x4, QP 17..37, MV sigma 24 px), batched form, hipGraph replay"),
}


def batch_frames(name: str):
    """(frames, refs, members): the batch's frames in order (intra first),
    the reference pictures the inter frames share, and each frame's
    (preset, seed)."""
    wl = BATCHES[name]
    ip, ni = wl["intra"]
    bp, nb = wl["inter"]
    ref_frame = hevcr.Frame(hevcr.preset(bp), wl["ref_seed"])
    refs = ref_frame.refs[:ref_frame.num_ref_pics]
    members = [(ip, s + 1) for s in range(ni)] + [(bp, s + 1) for s in range(nb)]
    frames = [hevcr.Frame(hevcr.preset(p), s, refs=refs if p == bp else None) for p, s in members]
    return frames, refs, members
