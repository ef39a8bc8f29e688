// Serving loop: continuous batching of frames from many streams on one engine.
//
// Streams are admitted by the device footprint of their DPB against a budget
// (first come, first served; a stream that does not fit waits until closed
// streams free enough). Each stream's frames run in submission order. A step
// takes the next queued frame of up to max_batch admitted streams (oldest
// arrival first), builds one batch from them and launches it on the engine's
// stream; up to max_inflight steps are queued on the device at once, so frames
// that arrive while earlier steps run are taken by the next step.
//
// Steps run on `lanes` device streams (engine lanes), concurrently with each
// other: a latency-bound stage of one step (intra prediction, with few
// workgroups per dependency wave) leaves the GPU to the other lanes' steps.
// A stream with frames in launched, unfinished steps has its next frames
// launched on the same lane: in-order execution on that lane is what lets
// frame t + 1 of a stream be launched while frame t is still queued (its
// references are complete by the time its kernels start). A stream with no
// such frames may go to any lane. With several lanes a step takes at most
// ceil(admitted streams / lanes) frames, so that the streams split between
// the lanes. Completion is observed with one event per step; steps retire in
// completion order.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <new>
#include <thread>
#include <vector>

#include "runtime.hpp"

using namespace hevcr;

namespace {

// A frame submitted as host syntax: prepared by an ingest worker, run by a
// step once ready, released by the server when its step retires.
struct Ingest {
    const hevcr_frame_syntax* syn = nullptr;
    Stream* st = nullptr;  // stays alive: a stream with outstanding frames is not released
    int stream_id = 0;
    std::atomic<int> state{0};  // 0 queued, 1 prepared, -1 failed
    int rc = HEVCR_OK;
    std::unique_ptr<hevcr_frame> f;
};

struct Job {
    hevcr_frame* f;
    int slot;
    uint64_t tag;
    double t_submit;
    std::shared_ptr<Ingest> in;  // set for frames submitted as syntax
    hevcr_frame* frame() const { return in ? in->f.get() : f; }
    bool ready() const { return !in || in->state.load(std::memory_order_acquire) != 0; }
};

struct ServedStream {
    int id = 0;
    uint64_t footprint = 0;
    bool admitted = false, closing = false;
    std::deque<Job> queue;
    int outstanding = 0;  // queued or in a launched step
    int launched = 0;     // in launched steps that have not retired
    int lane = 0;         // the lane of those steps
};

struct Step {
    std::unique_ptr<hevcr_batch> batch;
    hipEvent_t done = nullptr;
    std::vector<std::pair<int, Job>> jobs;  // (stream id, job)
    double t_launch = 0;
    int lane = 0;
};

}  // namespace

struct hevcr_server {
    hevcr_engine* e = nullptr;
    hevcr_server_cfg cfg{};
    // ingest: syntax submissions prepared by worker threads, oldest first
    std::mutex im;
    std::condition_variable icv;
    std::deque<std::shared_ptr<Ingest>> iq;
    bool stopping = false;
    std::vector<std::thread> workers;
    std::vector<std::unique_ptr<UploadCtx>> uploads;  // one per worker
    std::atomic<uint64_t> ingested{0}, ingest_bytes{0}, ingest_ns{0};
    std::atomic<uint64_t> ingest_phase_ns[4] = {{0}, {0}, {0}, {0}};  // validation, lists, staging, upload
    std::atomic<int> ingesting{0};  // submitted as syntax, not prepared yet
    uint64_t used = 0;
    std::map<int, ServedStream> streams;
    std::deque<int> waiting;  // streams not yet admitted, in open order
    std::deque<Step> inflight;
    std::vector<std::unique_ptr<hevcr_batch>> spare_batches;
    std::vector<hipEvent_t> spare_events;
    std::deque<hevcr_done> done;
    uint64_t frames_done = 0, steps_done = 0;
};

namespace {

int release_stream(hevcr_server* s, int id) {
    auto it = s->streams.find(id);
    if (it == s->streams.end()) return HEVCR_OK;
    if (it->second.admitted) s->used -= it->second.footprint;
    s->streams.erase(it);
    s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), id), s->waiting.end());
    auto es = s->e->streams.find(id);
    if (es != s->e->streams.end()) s->e->streams.erase(es);  // frees the DPB; no queued step uses it
    return HEVCR_OK;
}

int admit_waiting(hevcr_server* s) {
    while (!s->waiting.empty()) {
        ServedStream& ss = s->streams[s->waiting.front()];
        if (s->used + ss.footprint > s->cfg.dpb_budget_bytes) break;
        Stream* st = find_stream(s->e, ss.id);
        if (!st) return fail(HEVCR_ERROR, "served stream %d lost its engine stream", ss.id);
        int rc = allocate_dpb(s->e, *st);
        if (rc) return rc;
        s->used += ss.footprint;
        ss.admitted = true;
        s->waiting.pop_front();
    }
    return HEVCR_OK;
}

int retire(hevcr_server* s, size_t index) {
    Step& step = s->inflight[index];
    const double t = now_seconds();
    const int nframes = (int)step.jobs.size();
    for (auto& sj : step.jobs) {
        if (sj.second.in) sj.second.in->f.reset();  // its step is complete: the frame's block goes back to the pool
        hevcr_done d;
        d.tag = sj.second.tag;
        d.stream_id = sj.first;
        d.out_slot = sj.second.slot;
        d.status = HEVCR_OK;
        d.step_frames = nframes;
        d.t_submit = sj.second.t_submit;
        d.t_launch = step.t_launch;
        d.t_done = t;
        s->done.push_back(d);
        auto it = s->streams.find(sj.first);
        if (it != s->streams.end()) {
            it->second.outstanding--;
            it->second.launched--;
        }
    }
    s->frames_done += (uint64_t)nframes;
    s->steps_done++;
    s->spare_events.push_back(step.done);
    s->spare_batches.push_back(std::move(step.batch));
    std::vector<int> closed;
    for (auto& sj : step.jobs) {
        auto it = s->streams.find(sj.first);
        if (it != s->streams.end() && it->second.closing && it->second.outstanding == 0) closed.push_back(sj.first);
    }
    s->inflight.erase(s->inflight.begin() + (std::ptrdiff_t)index);
    for (int id : closed) release_stream(s, id);
    return closed.empty() ? HEVCR_OK : admit_waiting(s);
}

void ingest_worker(hevcr_server* s, UploadCtx* up) {
    (void)hipSetDevice(s->e->device);
    while (true) {
        std::shared_ptr<Ingest> in;
        {
            std::unique_lock<std::mutex> g(s->im);
            s->icv.wait(g, [&] { return s->stopping || !s->iq.empty(); });
            if (s->stopping) return;
            in = std::move(s->iq.front());
            s->iq.pop_front();
        }
        const double t0 = now_seconds();
        auto f = std::unique_ptr<hevcr_frame>(new (std::nothrow) hevcr_frame());
        int rc = f ? prepare_checked(s->e, *in->st, in->stream_id, in->syn, f.get(), up) : HEVCR_ENOMEM;
        if (!rc) {
            s->ingest_bytes += f->upload_bytes;
            s->ingested++;
            for (int k = 0; k < 4; k++) s->ingest_phase_ns[k] += (uint64_t)(f->prep_s[k] * 1e9);
            in->f = std::move(f);
        }
        in->rc = rc;
        s->ingest_ns += (uint64_t)((now_seconds() - t0) * 1e9);
        in->state.store(rc ? -1 : 1, std::memory_order_release);
        s->ingesting.fetch_sub(1, std::memory_order_release);
    }
}

void stop_workers(hevcr_server* s) {
    {
        std::lock_guard<std::mutex> g(s->im);
        s->stopping = true;
    }
    s->icv.notify_all();
    for (auto& t : s->workers) t.join();
    s->workers.clear();
}

// Frames whose preparation failed leave their stream's queue as completions
// with the error as status.
int drop_failed(hevcr_server* s) {
    std::vector<int> closed;
    for (auto& kv : s->streams) {
        ServedStream& ss = kv.second;
        while (!ss.queue.empty() && ss.queue.front().in && ss.queue.front().in->state.load(std::memory_order_acquire) < 0) {
            const Job& j = ss.queue.front();
            hevcr_done d{};
            d.tag = j.tag;
            d.stream_id = kv.first;
            d.out_slot = j.slot;
            d.status = j.in->rc;
            d.t_submit = j.t_submit;
            d.t_launch = d.t_done = now_seconds();
            s->done.push_back(d);
            ss.queue.pop_front();
            if (--ss.outstanding == 0 && ss.closing) closed.push_back(kv.first);
        }
    }
    for (int id : closed) release_stream(s, id);
    return closed.empty() ? HEVCR_OK : admit_waiting(s);
}

}  // namespace

extern "C" {

double hevcr_server_clock(void) { return now_seconds(); }

int hevcr_server_create(hevcr_engine* e, const hevcr_server_cfg* cfg, hevcr_server** out) {
    if (out) *out = nullptr;
    if (!e || !cfg || !out) return fail(HEVCR_EINVAL, "null argument");
    if (cfg->max_batch < 1 || cfg->max_batch > 1024) return fail(HEVCR_EINVAL, "max_batch %d", cfg->max_batch);
    if (cfg->max_inflight < 1 || cfg->max_inflight > 8) return fail(HEVCR_EINVAL, "max_inflight %d", cfg->max_inflight);
    if (cfg->lanes < 1 || cfg->lanes > kMaxLanes) return fail(HEVCR_EINVAL, "lanes %d (1..%d)", cfg->lanes, kMaxLanes);
    if (cfg->ingest_threads < 0 || cfg->ingest_threads > 64)
        return fail(HEVCR_EINVAL, "ingest_threads %d (0..64)", cfg->ingest_threads);
    HIPCHK(hipSetDevice(e->device));
    int rc = ensure_lanes(e, cfg->lanes);
    if (rc) return rc;
    auto* s = new (std::nothrow) hevcr_server();
    if (!s) return fail(HEVCR_ENOMEM, "server allocation");
    s->e = e;
    s->cfg = *cfg;
    for (int i = 0; i < cfg->ingest_threads; i++) {
        auto up = std::make_unique<UploadCtx>();
        if ((rc = up->init(e->device))) {
            hevcr_server_destroy(s);
            return rc;
        }
        s->uploads.push_back(std::move(up));
    }
    for (auto& up : s->uploads) s->workers.emplace_back(ingest_worker, s, up.get());
    *out = s;
    return HEVCR_OK;
}

void hevcr_server_destroy(hevcr_server* s) {
    if (!s) return;
    stop_workers(s);
    (void)hipSetDevice(s->e->device);
    (void)sync_lanes(s->e);
    for (auto& st : s->inflight) (void)hipEventDestroy(st.done);
    for (auto ev : s->spare_events) (void)hipEventDestroy(ev);
    std::vector<int> ids;
    for (auto& kv : s->streams) ids.push_back(kv.first);
    for (int id : ids) release_stream(s, id);
    delete s;
}

int hevcr_server_open(hevcr_server* s, int width, int height, int bit_depth, int dpb_slots, int* stream_id) {
    if (!s || !stream_id) return fail(HEVCR_EINVAL, "null argument");
    int id = 0;
    int rc = create_stream(s->e, width, height, bit_depth, dpb_slots, false, &id);
    if (rc) return rc;
    Stream* st = find_stream(s->e, id);
    ServedStream ss;
    ss.id = id;
    ss.footprint = st->dpb_bytes();
    if (ss.footprint > s->cfg.dpb_budget_bytes) {
        s->e->streams.erase(id);
        return fail(HEVCR_EINVAL, "a %dx%d stream with %d pictures needs %llu bytes, above the whole budget", width,
                    height, dpb_slots, (unsigned long long)ss.footprint);
    }
    s->streams[id] = std::move(ss);
    s->waiting.push_back(id);
    HIPCHK(hipSetDevice(s->e->device));
    if ((rc = admit_waiting(s))) return rc;
    *stream_id = id;
    return HEVCR_OK;
}

int hevcr_server_close(hevcr_server* s, int stream_id) {
    if (!s) return fail(HEVCR_EINVAL, "null server");
    auto it = s->streams.find(stream_id);
    if (it == s->streams.end()) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    it->second.closing = true;
    if (it->second.outstanding > 0) return HEVCR_OK;  // released when its last frame retires
    HIPCHK(hipSetDevice(s->e->device));
    release_stream(s, stream_id);
    return admit_waiting(s);
}

int hevcr_server_admitted(hevcr_server* s, int stream_id, int* admitted) {
    if (!s || !admitted) return fail(HEVCR_EINVAL, "null argument");
    auto it = s->streams.find(stream_id);
    if (it == s->streams.end()) return fail(HEVCR_EINVAL, "stream %d", stream_id);
    *admitted = it->second.admitted ? 1 : 0;
    return HEVCR_OK;
}

int hevcr_server_submit(hevcr_server* s, hevcr_frame* f, int out_slot, uint64_t tag, double t_arrival) {
    if (!s || !f) return fail(HEVCR_EINVAL, "null argument");
    auto it = s->streams.find(f->stream_id);
    if (it == s->streams.end()) return fail(HEVCR_EINVAL, "frame of stream %d, which this server does not serve", f->stream_id);
    ServedStream& ss = it->second;
    if (ss.closing) return fail(HEVCR_EINVAL, "stream %d is closing", f->stream_id);
    const Stream* st = find_stream(s->e, f->stream_id);
    if (out_slot < 0 || out_slot >= st->nslots) return fail(HEVCR_EINVAL, "slot %d", out_slot);
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < f->p.num_ref[l]; i++)
            if (f->p.ref_slot[l][i] == out_slot) return fail(HEVCR_EINVAL, "reference slot %d is the output slot", out_slot);
    ss.queue.push_back({f, out_slot, tag, t_arrival >= 0 ? t_arrival : now_seconds(), nullptr});
    ss.outstanding++;
    return HEVCR_OK;
}

int hevcr_server_submit_syntax(hevcr_server* s, int stream_id, const hevcr_frame_syntax* syn, int out_slot,
                               uint64_t tag, double t_arrival) {
    if (!s || !syn) return fail(HEVCR_EINVAL, "null argument");
    if (s->cfg.ingest_threads < 1) return fail(HEVCR_EINVAL, "the server has no ingest threads");
    auto it = s->streams.find(stream_id);
    if (it == s->streams.end()) return fail(HEVCR_EINVAL, "stream %d, which this server does not serve", stream_id);
    ServedStream& ss = it->second;
    if (ss.closing) return fail(HEVCR_EINVAL, "stream %d is closing", stream_id);
    Stream* st = find_stream(s->e, stream_id);
    if (out_slot < 0 || out_slot >= st->nslots) return fail(HEVCR_EINVAL, "slot %d", out_slot);
    for (int l = 0; l < 2; l++)
        for (int i = 0; i < syn->p.num_ref[l] && i < HEVCR_MAX_REFS; i++)
            if (syn->p.ref_slot[l][i] == out_slot) return fail(HEVCR_EINVAL, "reference slot %d is the output slot", out_slot);
    auto in = std::make_shared<Ingest>();
    in->syn = syn;
    in->st = st;
    in->stream_id = stream_id;
    Job j{nullptr, out_slot, tag, t_arrival >= 0 ? t_arrival : now_seconds(), in};
    ss.queue.push_back(j);
    ss.outstanding++;
    s->ingesting.fetch_add(1, std::memory_order_relaxed);
    {
        std::lock_guard<std::mutex> g(s->im);
        s->iq.push_back(std::move(in));
    }
    s->icv.notify_one();
    return HEVCR_OK;
}

int hevcr_server_step(hevcr_server* s, int* launched) {
    if (!s || !launched) return fail(HEVCR_EINVAL, "null argument");
    *launched = 0;
    if (int rc = drop_failed(s)) return rc;
    if ((int)s->inflight.size() >= s->cfg.max_inflight) return HEVCR_OK;
    // timed runs (engine profiling) record their events on lane 0 only
    const int nlanes = s->e->profiling ? 1 : s->cfg.lanes;
    // the head frame of every admitted stream, oldest arrival first
    std::vector<std::pair<double, int>> ready;
    int admitted = 0, waiting_streams = 0;  // admitted streams with queued frames, prepared or not
    for (auto& kv : s->streams) {
        admitted += kv.second.admitted ? 1 : 0;
        waiting_streams += kv.second.admitted && !kv.second.queue.empty() ? 1 : 0;
        if (kv.second.admitted && !kv.second.queue.empty() && kv.second.queue.front().ready())
            ready.push_back({kv.second.queue.front().t_submit, kv.first});
    }
    if (ready.empty()) return HEVCR_OK;
    std::sort(ready.begin(), ready.end());
    // the lane with the fewest queued steps among those that can take a frame
    int busy[kMaxLanes] = {0, 0, 0, 0};
    for (auto& st : s->inflight) busy[st.lane]++;
    auto fits = [&](int id, int l) {
        const ServedStream& ss = s->streams[id];
        return ss.launched == 0 ? true : ss.lane == l;
    };
    int lane = -1;
    for (int l = 0; l < nlanes; l++) {
        if (lane >= 0 && busy[l] >= busy[lane]) continue;
        for (auto& r : ready)
            if (fits(r.second, l)) {
                lane = l;
                break;
            }
    }
    if (lane < 0) return HEVCR_OK;  // every ready stream waits for a lane that profiling does not use
    const int cap = std::min(s->cfg.max_batch, nlanes > 1 ? (admitted + nlanes - 1) / nlanes : s->cfg.max_batch);
    std::vector<int> take;
    for (auto& r : ready)
        if ((int)take.size() < cap && fits(r.second, lane)) take.push_back(r.second);
    // While a step runs and frames are still being ingested, a step short of
    // a full one waits: a later, larger step amortises the per-step critical
    // path (the intra waves) that a small one would pay again. With nothing
    // queued on the device it launches at once.
    if (!s->inflight.empty() && (int)take.size() < std::min(cap, waiting_streams) &&
        s->ingesting.load(std::memory_order_acquire) > 0)
        return HEVCR_OK;
    std::vector<hevcr_frame*> frames;
    std::vector<int> slots;
    for (int id : take) {
        const Job& j = s->streams[id].queue.front();
        frames.push_back(j.frame());
        slots.push_back(j.slot);
    }
    HIPCHK(hipSetDevice(s->e->device));
    Step step;
    step.lane = lane;
    if (!s->spare_batches.empty()) {
        step.batch = std::move(s->spare_batches.back());
        s->spare_batches.pop_back();
    } else {
        step.batch.reset(new (std::nothrow) hevcr_batch());
        if (!step.batch) return fail(HEVCR_ENOMEM, "batch allocation");
    }
    int rc = build_batch(s->e, frames.data(), slots.data(), (int)frames.size(), HEVCR_RUN_FULL, step.batch.get(), lane);
    if (!rc) rc = run_batch(s->e, step.batch.get(), false, lane);
    if (rc) {
        s->spare_batches.push_back(std::move(step.batch));
        return rc;
    }
    if (!s->spare_events.empty()) {
        step.done = s->spare_events.back();
        s->spare_events.pop_back();
    } else {
        HIPCHK(hipEventCreateWithFlags(&step.done, hipEventDisableTiming | hipEventBlockingSync));
    }
    HIPCHK(hipEventRecord(step.done, hevcr::lane(s->e, lane).main));
    step.t_launch = now_seconds();
    for (int id : take) {
        ServedStream& ss = s->streams[id];
        step.jobs.push_back({id, ss.queue.front()});
        ss.queue.pop_front();
        ss.launched++;
        ss.lane = lane;
    }
    *launched = (int)step.jobs.size();
    s->inflight.push_back(std::move(step));
    return HEVCR_OK;
}

int hevcr_server_poll(hevcr_server* s, int wait, hevcr_done* out, int cap, int* n) {
    if (!s || !n || (cap > 0 && !out)) return fail(HEVCR_EINVAL, "null argument");
    *n = 0;
    HIPCHK(hipSetDevice(s->e->device));
    // finished steps, in any lane order
    for (size_t i = 0; i < s->inflight.size();) {
        const hipError_t q = hipEventQuery(s->inflight[i].done);
        if (q == hipErrorNotReady) {
            i++;
            continue;
        }
        if (q != hipSuccess) return fail(HEVCR_EDEVICE, "step: %s", hipGetErrorString(q));
        int rc = retire(s, i);
        if (rc) return rc;
    }
    if (wait && s->done.empty() && !s->inflight.empty()) {  // the oldest step
        HIPCHK(hipEventSynchronize(s->inflight.front().done));
        int rc = retire(s, 0);
        if (rc) return rc;
    }
    while (*n < cap && !s->done.empty()) {
        out[(*n)++] = s->done.front();
        s->done.pop_front();
    }
    return HEVCR_OK;
}

int hevcr_server_info(hevcr_server* s, hevcr_server_state* info) {
    if (!s || !info) return fail(HEVCR_EINVAL, "null argument");
    info->streams = (int32_t)s->streams.size();
    info->admitted = 0;
    info->queued_frames = 0;
    for (auto& kv : s->streams) {
        info->admitted += kv.second.admitted ? 1 : 0;
        info->queued_frames += (int32_t)kv.second.queue.size();
    }
    info->inflight_steps = (int32_t)s->inflight.size();
    info->budget_used = s->used;
    info->frames_done = s->frames_done;
    info->steps_done = s->steps_done;
    info->ingested_frames = s->ingested.load();
    info->ingest_bytes = s->ingest_bytes.load();
    info->ingest_busy_s = (double)s->ingest_ns.load() * 1e-9;
    info->ingest_validate_s = (double)s->ingest_phase_ns[0].load() * 1e-9;
    info->ingest_lists_s = (double)s->ingest_phase_ns[1].load() * 1e-9;
    info->ingest_stage_s = (double)s->ingest_phase_ns[2].load() * 1e-9;
    info->ingest_upload_s = (double)s->ingest_phase_ns[3].load() * 1e-9;
    return HEVCR_OK;
}

}  // extern "C"
