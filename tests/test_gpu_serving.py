"""Served form: streams of mixed resolution and bit depth, GOPs of an intra
frame and P frames that reference the stream's previous decoded frames, run
through the continuous-batching server. Every frame equals the oracle's
reconstruction of the same chain, a stream's frames are the same alone and in
the mix, and admission by DPB footprint queues streams that do not fit."""
import pytest

import hevcr
import oracle_lib
import serving

pytestmark = pytest.mark.gpu

SMALL = [(416, 240, 8), (192, 128, 10), (256, 144, 8)]


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _specs(frames=6):
    return [serving.StreamSpec(w, h, bd, seed=11 + k, frames=frames) for k, (w, h, bd) in enumerate(SMALL)]


def _oracle_chain(spec):
    """SHA-256 per plane of every frame, reconstructed by the oracle with the
    stream's own DPB."""
    dpb = [hevcr.Picture(spec.width, spec.height, spec.bit_depth) for _ in range(serving.DPB_SLOTS)]
    out = []
    for t in range(spec.frames):
        f = serving.gen_frame(spec, t)
        f.refs, f.num_ref_pics = dpb, serving.DPB_SLOTS
        pic = oracle_lib.recon(f)[2]
        dpb[serving.out_slot(t)] = pic
        out.append(pic.sha256())
    return out


def _serve(engine, specs, budget, max_inflight, saturate, collect, lanes=1):
    with hevcr.Server(engine, budget, max_batch=64, max_inflight=max_inflight, lanes=lanes) as srv:
        for s in specs:
            s.sid = srv.open(s.width, s.height, s.bit_depth, serving.DPB_SLOTS)
            s.prepared = []
        serving.prepare_streams(engine, specs, workers=4)

        def on_done(recs):
            for r in recs:
                si, t = divmod(r["tag"], 100000)
                collect(srv, specs, si, t, r)
        records, _ = serving.serve(srv, specs, saturate, on_done)
        for s in specs:
            for p in s.prepared:
                p.release()
        return records


def test_served_frames_match_oracle(engine):
    specs = _specs()
    want = {si: _oracle_chain(s) for si, s in enumerate(specs)}
    got = {}

    def collect(srv, specs, si, t, r):
        s = specs[si]
        pic = engine.download(s.sid, r["out_slot"], s.width, s.height, s.bit_depth)
        got[(si, t)] = pic.sha256()
    print("serving small streams, max_inflight 1", flush=True)
    records = _serve(engine, specs, 1 << 30, 1, True, collect)
    assert len(records) == sum(s.frames for s in specs)
    assert max(r["step_frames"] for r in records) == len(specs)  # frames of all streams shared steps
    for si, s in enumerate(specs):
        for t in range(s.frames):
            assert got[(si, t)] == want[si][t], f"stream {si} frame {t}"


def _digests_served(engine, specs, budget, max_inflight, saturate, lanes=1, steps=None):
    got = {}

    def collect(srv, specs, si, t, r):
        got[(si, t)] = engine.digests([(specs[si].sid, r["out_slot"])])[0]
        if steps is not None:
            steps.append(r["step_frames"])
    _serve(engine, specs, budget, max_inflight, saturate, collect, lanes)
    return got


@pytest.mark.parametrize("lanes,max_inflight", [(2, 2), (2, 4), (3, 6)])
def test_lanes_equal_one_lane(engine, lanes, max_inflight):
    """Steps on several device streams at once: every frame equals the
    one-lane run's, saturated and with Poisson arrivals, and the streams
    split between the lanes (a step holds at most ceil(streams / lanes))."""
    mk = lambda: _specs(frames=8) + _same_size_specs(8)
    ref = _digests_served(engine, mk(), 1 << 30, 1, True)
    print(f"serving on {lanes} lanes, max_inflight {max_inflight}, saturated", flush=True)
    steps = []
    got = _digests_served(engine, mk(), 1 << 30, max_inflight, True, lanes, steps)
    assert got == ref
    assert max(steps) <= -(-len(mk()) // lanes)
    specs = mk()
    serving.poisson_arrivals(specs, rate_fps=600.0, seed=5)
    print(f"serving on {lanes} lanes, Poisson arrivals", flush=True)
    assert _digests_served(engine, specs, 1 << 30, max_inflight, False, lanes) == ref


def test_alone_equals_mixed(engine):
    specs = _specs(frames=8)
    serving.poisson_arrivals(specs, rate_fps=400.0, seed=3)
    print("serving the mix with Poisson arrivals, max_inflight 2", flush=True)
    mixed = _digests_served(engine, specs, 1 << 30, 2, False)
    for si in range(len(specs)):
        alone_spec = _specs(frames=8)[si]
        print(f"serving stream {si} alone", flush=True)
        alone = _digests_served(engine, [alone_spec], 1 << 30, 2, True)
        for t in range(alone_spec.frames):
            assert mixed[(si, t)] == alone[(0, t)], f"stream {si} frame {t}"


def _same_size_specs(frames):
    return [serving.StreamSpec(256, 144, 8, seed=31 + k, frames=frames) for k in range(3)]


def test_admission_queues_streams_over_budget(engine):
    specs = _same_size_specs(5)
    budget = serving.dpb_bytes(specs[0]) + 1  # room for one stream at a time
    print("serving with a one-stream budget", flush=True)
    with hevcr.Server(engine, budget, max_batch=64, max_inflight=2) as srv:
        for s in specs:
            s.sid = srv.open(s.width, s.height, s.bit_depth, serving.DPB_SLOTS)
        assert [srv.admitted(s.sid) for s in specs] == [True, False, False]
        serving.prepare_streams(engine, specs, workers=4)
        order = []
        for si, s in enumerate(specs):
            for t in range(s.frames):
                srv.submit(s.prepared[t], serving.out_slot(t), tag=si * 100000 + t)
            srv.close_stream(s.sid)  # released once its frames are done
        done = 0
        while done < sum(s.frames for s in specs):
            srv.step()
            for r in srv.poll(wait=True):
                order.append(r["tag"] // 100000)
                done += 1
        # streams ran one after another: the second starts only after the first finished
        assert order == sorted(order)
        assert srv.info()["streams"] == 0 and srv.info()["budget_used"] == 0
    # the same streams without a budget limit give the same frames
    ref = _digests_served(engine, _same_size_specs(5), 1 << 30, 2, True)
    got = _digests_served(engine, _same_size_specs(5), budget, 2, True)
    assert got == ref
