"""The benchmark configurations at their full sizes, on the GPU:

  cfg2  32 independent 1080p 8-bit frames (8 intra + 24 B) in one batch,
        replayed from a hipGraph;
  cfg3  8 independent 3840x2160 10-bit frames (2 intra + 6 B);
  cfg4  64 concurrent streams, resolution uniform over 640x360 / 1280x720 /
        1920x1080 / 3840x2160, 25 % 10-bit, GOP 1 intra + 15 P, Poisson
        arrivals, a DPB budget small enough that admission queues streams.

Every frame is compared with the CPU oracle by SHA-256 per plane (computed
on the device for the engine's pictures; for the served mix, the oracle
decodes each stream's own chain), the batches also with the plain form and,
where the frame has a golden case, with the reference's hashes."""
from concurrent.futures import ThreadPoolExecutor

import pytest

import hevcr
import oracle_lib
import serving
import workloads

pytestmark = pytest.mark.gpu
GOLDEN = {c["name"]: c["stages"]["sao"] for c in oracle_lib.load_goldens()}
WORKERS = 16


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _oracle_hashes(frames):
    with ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(lambda f: oracle_lib.recon(f)[2].sha256(), frames))


@pytest.mark.parametrize("name", ["cfg2", "cfg3"])
def test_full_batch_matches_oracle_plain_and_goldens(engine, name):
    wl = workloads.BATCHES[name]
    frames, refs, members = workloads.batch_frames(name)
    n, k = len(frames), len(refs)
    W, H, bd = frames[0].width, frames[0].height, frames[0].bit_depth
    print(f"{name}: {n} frames {W}x{H} {bd}-bit, oracle", flush=True)
    want = _oracle_hashes(frames)
    for i, (p, s) in enumerate(members):
        g = wl["goldens"].get((p, s))
        if g:
            assert want[i] == GOLDEN[g], f"oracle frame {i} vs golden {g}"
    sid = engine.stream(W, H, bd, k + n)
    prepared = []
    try:
        for slot, r in enumerate(refs):
            engine.upload(sid, slot, r)
        slots = [k + i for i in range(n)]
        prepared = [engine.prepare(sid, f.syntax) for f in frames]
        print(f"{name}: plain form", flush=True)
        for pf, slot in zip(prepared, slots):
            pf.run(slot)
        engine.sync()
        plain = engine.sha256([(sid, slot) for slot in slots])
        assert plain == want
        assert [engine.download(sid, slot, W, H, bd).sha256() for slot in slots[:2]] == want[:2]
        plain_d = engine.digests([(sid, slot) for slot in slots])
        b = engine.batch(prepared, slots)
        for rep in range(2):
            print(f"{name}: batched form, graph replay {rep}", flush=True)
            for slot in slots:  # the check sees the batch's own writes
                engine.upload(sid, slot, hevcr.Picture(W, H, bd))
            b.run(True)
            engine.sync()
            assert engine.digests([(sid, slot) for slot in slots]) == plain_d
        assert engine.sha256([(sid, slot) for slot in slots]) == want
        b.release()
    finally:
        for pf in prepared:
            pf.release()
        engine.destroy_stream(sid)


def _oracle_chain(spec):
    dpb = [hevcr.Picture(spec.width, spec.height, spec.bit_depth) for _ in range(serving.DPB_SLOTS)]
    out = []
    for t in range(spec.frames):
        f = serving.gen_frame(spec, t)
        f.refs, f.num_ref_pics = dpb, serving.DPB_SLOTS
        pic = oracle_lib.recon(f)[2]
        dpb[serving.out_slot(t)] = pic
        out.append(pic.sha256())
    return out


def test_cfg4_served_mix_matches_oracle(engine):
    """64 streams of the cfg4 mix through the continuous-batching server
    under Poisson arrivals with a budget for about a third of their DPBs:
    streams queue for admission and every frame equals the oracle's decode of
    its stream."""
    specs = serving.make_specs(64, "cfg4", seed=1)
    kinds = {(s.width, s.height) for s in specs}
    assert kinds == set(serving.RESOLUTIONS) and any(s.bit_depth == 10 for s in specs)
    print("cfg4: oracle chains of 64 streams", flush=True)
    with ThreadPoolExecutor(WORKERS) as ex:
        want = list(ex.map(_oracle_chain, specs))
    budget = sum(serving.dpb_bytes(s) for s in specs) // 3
    got, waited = {}, 0
    with hevcr.Server(engine, budget, max_batch=64, max_inflight=2) as srv:
        for s in specs:
            s.sid = srv.open(s.width, s.height, s.bit_depth, serving.DPB_SLOTS)
            waited += not srv.admitted(s.sid)
        assert waited > 0  # the budget holds back part of the mix
        serving.prepare_streams(engine, specs, workers=WORKERS)
        serving.poisson_arrivals(specs, rate_fps=960.0, seed=11)

        def on_done(recs):  # SHA-256 per plane on the device, one call per completed batch of frames
            hashes = engine.sha256([(specs[r["tag"] // 100000].sid, r["out_slot"]) for r in recs])
            for r, hsh in zip(recs, hashes):
                got[divmod(r["tag"], 100000)] = hsh
        print("cfg4: serving", flush=True)
        records, wall = serving.serve(srv, specs, False, on_done)
        for s in specs:
            for p in s.prepared:
                p.release()
        assert srv.info()["streams"] == 0 and srv.info()["budget_used"] == 0
    assert len(records) == 64 * serving.GOP
    print(f"cfg4: {len(records)} frames in {wall:.2f} s, max step {max(r['step_frames'] for r in records)}",
          flush=True)
    for si, s in enumerate(specs):
        for t in range(s.frames):
            assert got[(si, t)] == want[si][t], f"stream {si} ({s.width}x{s.height}/{s.bit_depth}) frame {t}"
