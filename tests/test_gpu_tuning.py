"""Tuning and the autotuner (hevcr_tuning_*, hevcr_autotune): every way of
cutting frames into work (8x8 PUs four to a wavefront or as tiles of their
own; deblocking + SAO fused or as two kernels; 4x4 intra blocks four to a
step or one) gives the reference's golden planes, in the plain and the
batched form; the autotuner times every candidate, keeps the fastest for
that picture size and bit depth, and frames prepared afterwards use it."""
import pytest

import hevcr
import oracle_lib

pytestmark = pytest.mark.gpu
BY_NAME = {c["name"]: c for c in oracle_lib.load_goldens()}
# 4x4 intra blocks; B frames with 8x8 PUs, intra CUs, SAO on 64x64 CTBs; skip-only P; 10-bit B
CASES = ["intra416_s1", "b1080_intra_mix", "p640x360_all_skip", "b640x360_10bit"]


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


def _teardown(engine, streams, prepared):
    for p in prepared:
        p.release()
    for sid in streams:
        engine.destroy_stream(sid)


def _check(engine, names, streams, frames, slots, what):
    for name, sid, f, slot in zip(names, streams, frames, slots):
        pic = engine.download(sid, slot, f.width, f.height, f.bit_depth)
        assert pic.sha256() == BY_NAME[name]["stages"]["sao"], f"{name}: {what}"


def _reset(engine, names):
    for name in names:
        f = oracle_lib.make_frame(BY_NAME[name])
        engine.set_tuning(f.width, f.height, f.bit_depth)


@pytest.mark.parametrize("cand", range(hevcr.TUNE_CANDIDATES))
def test_every_tuning_gives_the_golden_pictures(engine, cand):
    t = hevcr.Tuning.candidate(cand).as_dict()
    print(f"tuning {t}", flush=True)
    for name in CASES:
        f = oracle_lib.make_frame(BY_NAME[name])
        engine.set_tuning(f.width, f.height, f.bit_depth, **t)
        assert engine.tuning(f.width, f.height, f.bit_depth) == t
    streams, frames, prepared, slots = _setup(engine, CASES)
    try:
        for p, slot in zip(prepared, slots):
            p.run(slot)
        engine.sync()
        _check(engine, CASES, streams, frames, slots, f"plain, tuning {t}")
        b = engine.batch(prepared, slots)
        for rep in range(2):
            b.run(True)
            engine.sync()
            _check(engine, CASES, streams, frames, slots, f"batched replay {rep}, tuning {t}")
        b.release()
    finally:
        _teardown(engine, streams, prepared)
        _reset(engine, CASES)


def test_autotune_picks_the_fastest_and_keeps_it(engine):
    names = ["b1080_s1", "b1080_intra_mix", "p1080_s1"]
    streams, frames, prepared, slots = _setup(engine, names)
    f0 = frames[0]
    try:
        print("autotuning three 1080p frames", flush=True)
        best, ms = engine.autotune(streams, [f.syntax for f in frames], slots, reps=5)
        assert len(ms) == hevcr.TUNE_CANDIDATES and all(x > 0 for x in ms)
        k = min(range(len(ms)), key=lambda i: ms[i])
        assert best == hevcr.Tuning.candidate(k).as_dict()
        assert engine.tuning(f0.width, f0.height, f0.bit_depth) == best
        # frames prepared now use the chosen tuning, and still give the goldens
        again = [engine.prepare(sid, f.syntax) for sid, f in zip(streams, frames)]
        b = engine.batch(again, slots)
        b.run(True)
        engine.sync()
        _check(engine, names, streams, frames, slots, f"after autotune, tuning {best}")
        b.release()
        for p in again:
            p.release()
    finally:
        _teardown(engine, streams, prepared)
        engine.set_tuning(f0.width, f0.height, f0.bit_depth)


def test_tuning_rejects_other_values(engine):
    with pytest.raises(hevcr.EngineError) as ei:
        engine.set_tuning(416, 240, 8, mc8=2)
    assert ei.value.rc == -3  # HEVCR_EINVAL
    with pytest.raises(hevcr.EngineError):
        engine.autotune([12345], [oracle_lib.make_frame(BY_NAME["intra416_s1"]).syntax], [0], reps=1)
