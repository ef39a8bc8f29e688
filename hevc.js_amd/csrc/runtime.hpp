// Host-side objects of the reconstruction engine shared by the runtime
// (runtime.cpp) and the serving loop (serve.cpp). Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "engine.hpp"
#include "hevcr.h"

namespace hevcr {

// Kernel launchers (kernels.hip); every launch is asynchronous on `s`.
void launch_expand(hipStream_t s, const FrameDev* fr, const uint32_t* ctb_prefix, int nframes, uint32_t total,
                   int max_ctb_log2);
template <typename T>
void launch_mc(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total);
template <typename T>
void launch_mcb(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total);
template <typename T>
void launch_mc8(hipStream_t s, const FrameDev* fr, const uint32_t* wg_prefix, int nframes, uint32_t total);
template <typename T>
// max_rows: the tallest strip of the launch (8 or 16; bi-predicted strips: 8)
void launch_mcs(hipStream_t s, const FrameDev* fr, const uint32_t* wave_prefix, int nframes, uint32_t total, bool bi,
                int max_rows);
template <typename T>
void launch_residual(hipStream_t s, const FrameDev* fr, const uint32_t* const* wg_prefix, int nframes, const uint32_t* total);
template <typename T>
// max_ctb_log2: the largest CTB of the frames (sizes the workgroups' LDS).
void launch_intra(hipStream_t s, const FrameDev* fr, const uint32_t* diag_prefix, int nframes, int diag, uint32_t total,
                  int max_ctb_log2);
// The intra stage of a batch group in one persistent launch (plus the flag
// initialisation before it): see k_intra_p.
template <typename T>
void launch_intra_persistent(hipStream_t s, const FrameDev* fr, int nframes, int max_nctb, const uint32_t* diag_prefix,
                             const uint32_t* wave_start, int nwaves, uint32_t total_items, uint32_t* ticket,
                             uint32_t* err, uint32_t spin_limit, int max_ctb_log2);
template <typename T>
void launch_deblock(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total);
template <typename T>
void launch_sao(hipStream_t s, const FrameDev* fr, const uint32_t* ctb_prefix, int nframes, uint32_t total);
template <typename T>
void launch_loopfilter(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total);
void launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes);
void launch_strip_expand(hipStream_t s, const McBand* bands, uint32_t n, McStrip* strips);
void launch_digest(hipStream_t s, const PlaneRef* planes, int nplanes, int bpp, uint64_t* out);
// One lane per plane; planes of one sample size (bpp bytes); digest j at out + 8 * planes[j].out.
void launch_sha256(hipStream_t s, const ShaPlane* planes, int nplanes, int bpp, uint32_t* out);
// in_bytes / out_bytes: bytes per stored and per written sample (1, 1), (1, 2) or (2, 2).
void launch_egress(hipStream_t s, const EgressPlane* planes, int nplanes, int in_bytes, int out_bytes,
                   uint64_t max_samples);

// The address of the no-alloc entry point's device scratch array on `device`
// (which must be current), looked up on the first call per device (noalloc.hip).
int noalloc_init(int device, void** ptr);

// Records the message as the calling thread's last error and returns `code`.
int fail(int code, const char* fmt, ...);
const char* last_error();

#define HIPCHK(expr)                                                                                          \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) return ::hevcr::fail(HEVCR_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Device allocation that grows on demand. Growing frees the old block, so a
// caller grows a buffer only when no queued work still reads it.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    int reserve(size_t bytes);
    template <typename U>
    U* as(size_t byte_off = 0) const {
        return reinterpret_cast<U*>(static_cast<char*>(p) + byte_off);
    }
};

// Pinned host staging buffer (hipHostMalloc), so uploads are asynchronous.
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    int reserve(size_t bytes);
};

// Device blocks of prepared frames, reused from frame to frame so that
// preparing and releasing frames in a serving loop neither allocates nor
// frees device memory once the pool has grown (hipFree waits for the device).
// Shared by the engine and its frames, so a frame may outlive the engine.
struct BlobPool {
    int device = 0;
    std::mutex m;
    std::multimap<size_t, void*> free;  // capacity -> block
    std::vector<void*> all;
    ~BlobPool();
    int get(size_t bytes, void** p, size_t* cap);
    void put(void* p, size_t cap);
};

// A block from a BlobPool, returned to it on destruction. The owner returns
// it only when no queued device work reads it any more.
struct PoolBuf {
    std::shared_ptr<BlobPool> pool;
    void* p = nullptr;
    size_t cap = 0;
    PoolBuf() = default;
    PoolBuf(const PoolBuf&) = delete;
    PoolBuf& operator=(const PoolBuf&) = delete;
    ~PoolBuf() { release(); }
    void release() {
        if (p && pool) pool->put(p, cap);
        p = nullptr;
        cap = 0;
    }
    template <typename U>
    U* as(size_t byte_off = 0) const {
        return reinterpret_cast<U*>(static_cast<char*>(p) + byte_off);
    }
};

// What one host thread needs to upload prepared frames without waiting for
// other threads: its own device stream, an event and a pinned staging buffer.
// An ingest thread's uploads: its own stream and two pinned staging buffers
// used in turn. A frame's upload is not waited for by the thread: the frame
// carries the event of its copy (hevcr_frame::upload_ev), which the stream
// that reconstructs it waits on; a staging buffer is reused once the copy
// out of it two frames earlier has finished.
struct UploadCtx {
    int device = 0;
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    void* pinned[2] = {nullptr, nullptr};
    size_t cap[2] = {0, 0};
    int next = 0;  // the staging buffer the next frame uses
    UploadCtx() = default;
    UploadCtx(const UploadCtx&) = delete;
    UploadCtx& operator=(const UploadCtx&) = delete;
    ~UploadCtx();
    int init(int dev);
    int reserve(int k, size_t bytes);
};

struct Picture {
    DevBuf mem;
    void* plane[3] = {nullptr, nullptr, nullptr};
};

struct Stream {
    int w = 0, h = 0, bd = 8, bpp = 1, nslots = 0;
    int stride[3] = {0, 0, 0};
    size_t plane_bytes[3] = {0, 0, 0};
    std::vector<std::unique_ptr<Picture>> slots;  // empty until the DPB is allocated
    size_t picture_bytes() const { return plane_bytes[0] + plane_bytes[1] + plane_bytes[2]; }
    size_t dpb_bytes() const { return picture_bytes() * (size_t)nslots; }
    bool allocated() const { return (int)slots.size() == nslots; }
};

}  // namespace hevcr

struct hevcr_batch;

// A prepared frame: its syntax and derived work lists in one device block,
// plus what batches need to know about it on the host.
struct hevcr_frame {
    int stream_id = 0;
    hevcr_frame_params p{};
    hevcr::FrameConst fc{};
    int bpp = 1;
    bool sao_active = false;
    bool lf_fused = false;   // full runs use the fused deblocking + SAO kernel (SAO on, 32x32 or 64x64 CTBs)
    bool has_intra = false;
    bool intra_dense = false;  // at least half of its CTBs have intra blocks (an intra picture)
    uint32_t ncu = 0, npu = 0, ntu = 0, nctb = 0, ntiles = 0, ntilesb = 0, ntiles8 = 0;  // tiles: uni, bi, 8x8
    uint32_t nstripw = 0, nstripwb = 0;      // waves of 64 strips: uni- and bi-predicted
    int strip_rows = 0;                      // the tallest uni-predicted strip class (8 or 16), 0: none
    uint32_t nblocks[4] = {0, 0, 0, 0};
    std::vector<uint32_t> diag_count;  // CTUs with intra blocks per dependency wave
    hevcr::PoolBuf blob;
    bool compact = false;                    // coefficients in the compact form (FrameDev::sig)
    uint64_t upload_bytes = 0;               // bytes copied to the device when it was prepared
    // byte offsets of the sections of `blob`
    size_t o_cus = 0, o_pus = 0, o_tus = 0, o_coeffs = 0, o_sao = 0, o_first_tu = 0, o_first_pu = 0, o_ctu_ranges = 0, o_tiles = 0, o_tilesb = 0, o_tiles8 = 0,
           o_strips = 0, o_stripsb = 0, o_bands = 0,
           o_blocks[4] = {0, 0, 0, 0}, o_ops = 0, o_op_begin = 0, o_op_span = 0, o_ictu = 0, o_diag_begin = 0, o_ictl = 0,
           o_scale = 0, o_ctl = 0;
    bool scaled = false;                     // scaling lists: factors at o_scale
    bool has_ctl = false;                    // per-CTB control words at o_ctl (slice segments, tiles, PCM, bypass)
    bool seg_weights = false;                // per-segment weights at o_segw, per-CTB segments at o_ctbseg
    double prep_s[4] = {0, 0, 0, 0};         // host time of its preparation: validation, lists, staging, upload
    size_t o_segw = 0, o_ctbseg = 0;
    int nseg_w = 0;                          // slice segments of the o_segw table
    uint32_t nbands = 0;                     // MC bands at o_bands, expanded into the strips on the device
    hipEvent_t upload_ev = nullptr;          // owned: its upload's completion, when not waited for on preparation
    double bytes[HEVCR_NUM_STAGES] = {};     // with deblocking and SAO as separate kernels
    double kbytes[HEVCR_NUM_KERNELS] = {};
    double lf_bytes = 0;                     // the fused kernel's algorithmic bytes
    hevcr_batch* plain = nullptr;  // cached one-frame batch of the plain form
    int plain_slot = -1, plain_stop = -1;
    hevcr_tuning tune{1, 1, 1, 1, 1};  // the tuning it was prepared with
    ~hevcr_frame();
};

namespace hevcr {

// Frames of one pixel type launched together.
struct BatchGroup {
    int bpp = 1;
    std::vector<int> members;       // indices into hevcr_batch::frames
    size_t o_fdev = 0, o_index = 0;  // byte offsets in hevcr_batch::meta
    uint32_t total_cu = 0, total_tiles = 0, total_tilesb = 0, total_tiles8 = 0, total_res[4] = {0, 0, 0, 0}, total_dbk = 0, total_sao = 0,
             total_lf = 0, total_stripw = 0, total_stripwb = 0;
    int strip_rows = 0;                // tallest uni-predicted strip class of the members
    std::vector<uint32_t> diag_total;  // CTUs per intra wave over all members
    int max_ctb_log2 = 4;              // largest CTB of the members (sizes k_expand's workgroups)
    int max_nctb = 0;                  // most CTBs of a member
    bool intra_persistent = true;      // the intra stage as one persistent launch (every member tuned so)
    size_t o_wave = 0;                 // byte offset in meta: wave_start (diag_total.size() + 1 entries, items)
    uint32_t intra_items = 0;          // (frame, CTU, component) items of the intra stage
    int n() const { return (int)members.size(); }
    // prefix arrays inside the index block: CTBs (expansion), tiles, res x4, deblock tiles, sao CTBs, 8x8-tile
    // groups, fused loop-filter CTBs, bi-predicted tiles, uni- and bi-predicted strip waves, then one per
    // intra wave
    size_t prefix_off(int which) const { return o_index + (size_t)which * (members.size() + 1) * 4; }
};
enum { kPreCu = 0, kPreTiles = 1, kPreRes = 2, kPreDbk = 6, kPreSao = 7, kPreTiles8 = 8, kPreLf = 9, kPreTilesB = 10, kPreStrips = 11, kPreStripsB = 12,
       kPreDiag = 13 };

// Algorithmic bytes per stage and per kernel of a frame run up to `stop`.
void run_bytes(const hevcr_frame* f, int stop, double* stage_bytes, double* kernel_bytes);

}  // namespace hevcr

struct hevcr_batch {
    std::vector<hevcr_frame*> frames;
    std::vector<int> slots;
    int stop = HEVCR_RUN_FULL;
    hevcr::BatchGroup groups[2];
    hevcr::DevBuf meta;     // FrameDev arrays and prefix arrays
    hevcr::DevBuf arena;    // per-frame work memory: side information, motion, residual planes, pre-SAO picture
    size_t o_ctl = 0;       // byte offset in arena: the persistent intra tickets of the two groups
    hevcr::PinnedBuf staging;
    hipEvent_t staged = nullptr;  // the last upload from `staging` has finished
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double bytes[HEVCR_NUM_STAGES] = {};
    double kbytes[HEVCR_NUM_KERNELS] = {};
    ~hevcr_batch();
};

namespace hevcr {

// A device stream that runs batches in order, with its own side stream for
// the syntax expansion and the events that fork and join it.
struct Lane {
    hipStream_t main = nullptr, side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int kMaxLanes = 4;

}  // namespace hevcr

struct hevcr_engine {
    int device = 0;
    void* scratch = nullptr;       // this device's hevcr_residual_blocks_noalloc scratch (noalloc_init)
    hevcr::DevBuf intra_err;       // set by a persistent intra launch whose wait gave up (hevcr_engine_sync reports it)
    uint32_t intra_spin_limit = 1u << 24;  // polls before a persistent intra wait gives up
    hipStream_t stream = nullptr;  // all reconstruction work of lane 0, in order
    hipStream_t upload = nullptr;  // syntax uploads of hevcr_frame_prepare, DPB clears
    hipStream_t side = nullptr;    // syntax expansion, concurrent with prediction (untimed runs)
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<hevcr::Lane> more_lanes;  // lanes 1.. (created by ensure_lanes)
    std::map<std::tuple<int, int, int>, hevcr_tuning> tuning;  // (width, height, bit depth)
    std::map<int, std::unique_ptr<hevcr::Stream>> streams;
    int next_stream = 1;
    bool profiling = false;
    struct Pending {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    double timer_ms[HEVCR_NUM_STAGES + HEVCR_NUM_KERNELS] = {};  // per stage, then per kernel
    uint64_t timer_n[HEVCR_NUM_STAGES + HEVCR_NUM_KERNELS] = {};
    hevcr::DevBuf kat_coeffs, kat_out, kat_blocks, kat_meta, digest_buf, egress_desc, egress_stage, sha_buf;
    std::shared_ptr<hevcr::BlobPool> blobs;  // device blocks of prepared frames
};

namespace hevcr {

Stream* find_stream(hevcr_engine* e, int id);
// Creates a stream; its DPB is allocated now or, with allocate == false, by a
// later allocate_dpb (the serving loop admits streams that way).
int create_stream(hevcr_engine* e, int width, int height, int bit_depth, int dpb_slots, bool allocate, int* id);
int allocate_dpb(hevcr_engine* e, Stream& st);
// Lane 0 is the engine's own stream (e->stream, e->side); lanes 1 .. n - 1
// are created here on first use. Batches on different lanes run concurrently.
int ensure_lanes(hevcr_engine* e, int n);
Lane lane(const hevcr_engine* e, int i);
// Waits for every lane.
int sync_lanes(hevcr_engine* e);
// (Re)builds `b` for these frames; reuses its buffers where they are large
// enough. Its upload is queued on the lane the batch will run on.
int build_batch(hevcr_engine* e, hevcr_frame* const* frames, const int* slots, int n, int stop, hevcr_batch* b,
                int lane_index = 0);
// Timed runs (e->profiling) only on lane 0.
int run_batch(hevcr_engine* e, hevcr_batch* b, bool graph, int lane_index = 0);
// Monotonic clock in seconds.
double now_seconds();
// Validates `syn` against the stream and prepares it into `f` (host work
// lists, one device block from the engine's pool). up: the calling thread's
// upload context (its stream, pinned staging); null: the engine's upload
// stream through pageable memory. Thread-safe for distinct frames as long as
// the stream is not destroyed meanwhile.
int prepare_checked(hevcr_engine* e, Stream& st, int stream_id, const hevcr_frame_syntax* syn, hevcr_frame* f,
                    UploadCtx* up);

}  // namespace hevcr
