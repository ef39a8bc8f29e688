"""Batched form: frames of different resolutions, bit depths and types
reconstructed together (one launch per stage, optionally replayed from a
captured hipGraph) give exactly the reference's golden planes, on every
replay; the batch rejects dependent frames."""
import numpy as np
import pytest

import hevcr
import oracle_lib

pytestmark = pytest.mark.gpu
BY_NAME = {c["name"]: c for c in oracle_lib.load_goldens()}
MIX = ["intra416_s1", "intra416_s2", "intra416_10bit", "intra200x136_ctb16", "intra200x136_ctb16_10bit",
       "b200x136_ctb16_10bit", "b640x360_10bit", "p640x360_all_skip",
       "b1080_s1", "p1080_s1", "b1080_intra_mix", "intra1080_s1", "b2160_10bit_s1"]
# pictures of several slice segments, tiles, PCM and bypass CUs batched with plain ones
MIX_SLICES = ["intra416_slices_tiles_pcm", "intra416_slices_pcm", "intra416_tiles_nolf", "intra416_tiles_slices",
              "b416_slices_types", "b416_tiles_nolf", "p640x360_slices_tiles", "intra200x136_ctb16_10bit_slices_pcm",
              "b200x136_ctb32_10bit_slices", "b1080_slices_tiles", "intra416_s1", "b1080_s1"]


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _setup(engine, names):
    streams, frames, prepared, slots = [], [], [], []
    for name in names:
        f = oracle_lib.make_frame(BY_NAME[name])
        k = f.num_ref_pics
        sid = engine.stream(f.width, f.height, f.bit_depth, k + 1)
        for s in range(k):
            engine.upload(sid, s, f.refs[s])
        streams.append(sid)
        frames.append(f)
        prepared.append(engine.prepare(sid, f.syntax))
        slots.append(k)
    return streams, frames, prepared, slots


def _check_outputs(engine, names, streams, frames, slots, stage="sao"):
    for name, sid, f, slot in zip(names, streams, frames, slots):
        pic = engine.download(sid, slot, f.width, f.height, f.bit_depth)
        assert pic.sha256() == BY_NAME[name]["stages"][stage], f"{name} {stage}"


@pytest.mark.parametrize("graph", [False, True], ids=["direct", "graph"])
def test_mixed_batch_matches_goldens(engine, graph):
    streams, frames, prepared, slots = _setup(engine, MIX)
    try:
        b = engine.batch(prepared, slots)
        for rep in range(3):
            print(f"batch replay {rep} graph={graph}", flush=True)
            b.run(graph)
            engine.sync()
            _check_outputs(engine, MIX, streams, frames, slots)
            # unrelated eager work between replays
            prepared[0].run(slots[0], hevcr.RUN_STOP_RECON)
            engine.sync()
        b.release()
    finally:
        for p in prepared:
            p.release()
        for s in streams:
            engine.destroy_stream(s)


@pytest.mark.parametrize("graph", [False, True], ids=["direct", "graph"])
def test_mixed_batch_with_slices_tiles_pcm(engine, graph):
    streams, frames, prepared, slots = _setup(engine, MIX_SLICES)
    try:
        b = engine.batch(prepared, slots)
        for rep in range(2):
            print(f"slice/tile batch replay {rep} graph={graph}", flush=True)
            b.run(graph)
            engine.sync()
            _check_outputs(engine, MIX_SLICES, streams, frames, slots)
        b.release()
    finally:
        for p in prepared:
            p.release()
        for s in streams:
            engine.destroy_stream(s)


@pytest.mark.parametrize("stop,stage", [(hevcr.RUN_STOP_RECON, "recon"), (hevcr.RUN_STOP_DEBLOCK, "deblock")])
def test_batch_stop_points(engine, stop, stage):
    names = MIX[:6]
    streams, frames, prepared, slots = _setup(engine, names)
    try:
        b = engine.batch(prepared, slots, stop)
        b.run(True)
        engine.sync()
        _check_outputs(engine, names, streams, frames, slots, stage)
        b.release()
    finally:
        for p in prepared:
            p.release()
        for s in streams:
            engine.destroy_stream(s)


def test_batch_rejects_dependent_frames(engine):
    f = oracle_lib.make_frame(BY_NAME["p640x360_all_skip"])
    k = f.num_ref_pics
    sid = engine.stream(f.width, f.height, 8, k + 2)
    try:
        p1 = engine.prepare(sid, f.syntax)
        p2 = engine.prepare(sid, f.syntax)
        with pytest.raises(hevcr.EngineError):
            engine.batch([p1, p2], [k, k])  # same slot twice
        with pytest.raises(hevcr.EngineError):
            engine.batch([p1, p2], [k, 0])  # second frame writes a slot the first reads
        engine.batch([p1, p2], [k, k + 1]).release()  # independent: accepted
        p1.release()
        p2.release()
    finally:
        engine.destroy_stream(sid)
