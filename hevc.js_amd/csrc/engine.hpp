// Internal device-side descriptors shared by the host runtime and the HIP
// kernels of the reconstruction engine.
//
// Every launch covers a batch of frames (one frame in the plain form): kernels
// receive an array of FrameDev descriptors and, per stage, a prefix array over
// the batch's frames that maps each workgroup to its frame. A frame's work
// lists are built once, when its syntax is prepared, and stay in device memory;
// a batch adds only the descriptors and prefix arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hevcr_syntax.h"

namespace hevcr {

constexpr int kMaxRefs = HEVCR_MAX_REFS;

// Per-4x4 side information word (the reference keeps these as separate
// grids filled while parsing, coding_tree.cpp:956-1054 and 1233-1247).
//   bits  0..2  log2 of the TU size covering the cell
//   bit   3     coded luma in that TU
//   bit   4     vertical edge at the cell's left boundary
//   bit   5     horizontal edge at the cell's top boundary
//   bits  6..7  prediction mode of the CU (HEVCR_MODE_*)
//   bits  8..15 QpY of the CU (int8)
//   bits 16..19 slice_beta_offset_div2 + 6 of the cell's slice segment
//   bits 20..23 slice_tc_offset_div2 + 6
//   bit  24     the filters keep the cell's samples: a PCM CU with
//               pcm_loop_filter_disabled_flag, or a transquant-bypass CU
// The edge flags are already cleared where the reference's
// is_boundary_excluded (deblocking.cpp:50-120) drops the edge.
namespace info {
constexpr uint32_t kLog2Mask = 7u;
constexpr uint32_t kCbf = 1u << 3;
constexpr uint32_t kEdgeV = 1u << 4;
constexpr uint32_t kEdgeH = 1u << 5;
constexpr int kModeShift = 6;
constexpr int kQpShift = 8;
constexpr int kBetaShift = 16;
constexpr int kTcShift = 20;
constexpr uint32_t kNoFilter = 1u << 24;
}  // namespace info

// Per-CTB control word of pictures with several slice segments or tiles, or
// with PCM / transquant-bypass CUs (FrameDev::ctb_ctl; null otherwise).
//   bits  0..7  SAO edge-offset neighbours this CTB's samples may not read
//               (sao.cpp:142-185): bit kNb* per neighbouring CTB
//   bit   8     its slice segment has slice_deblocking_filter_disabled_flag
//   bit   9     deblocking skips the edges on its left CTB boundary
//   bit  10     and on its top CTB boundary (tile / slice boundaries,
//               deblocking.cpp:69-113)
//   bit  11     it holds samples the filters keep (info::kNoFilter)
//   bits 16..23 beta / tc offsets of its slice segment, as in the info word
namespace ctl {
constexpr int kNbL = 0, kNbR = 1, kNbT = 2, kNbB = 3, kNbTL = 4, kNbTR = 5, kNbBL = 6, kNbBR = 7;
constexpr uint32_t kSaoMask = 0xffu;
constexpr uint32_t kDbkOff = 1u << 8;
constexpr uint32_t kLeftOff = 1u << 9;
constexpr uint32_t kTopOff = 1u << 10;
constexpr uint32_t kNoFilter = 1u << 11;
constexpr uint32_t kOffsets = 0xffu << 16;
}  // namespace ctl

// Motion of one 4x4 cell (PUMotionInfo, inter_prediction.h:16-20).
struct Motion4 {
    int16_t mv[2][2];
    int8_t ref[2];
    uint8_t flags;
    uint8_t pad;
};
static_assert(sizeof(Motion4) == 12, "Motion4 is three dwords");

// One dequant + inverse-transform block (one component of one TU).
struct ResBlock {
    uint32_t coeff_off;  // into the frame's coefficient pool
    uint16_t x, y;       // component sample position
    uint8_t log2;        // 2..5
    uint8_t cidx;
    uint8_t qp;          // qP' (QpBdOffset included)
    uint8_t flags;       // kResIntra | kResDst | kResTskip | kResToPlane
};
constexpr uint8_t kResIntra = 1, kResDst = 2, kResTskip = 4, kResToPlane = 8;
// PCM samples written to the picture as they are; transquant bypass: the
// levels are the residual (no dequantisation, no transform)
constexpr uint8_t kResPcm = 16, kResBypass = 32, kResRaw = kResPcm | kResBypass;

// One motion-compensation tile: a part of a PU at most 16x16 luma, carrying
// the PU's motion.
struct McTile {
    uint16_t x, y;       // luma position of the tile
    uint8_t w, h;        // luma size (multiples of 4)
    int16_t mv[2][2];
    int8_t ref[2];
    uint8_t flags;       // HEVCR_PRED_L0 | HEVCR_PRED_L1
    uint8_t pad[3];
};
static_assert(sizeof(McTile) == 20, "McTile is five dwords");

// One motion-compensation strip: luma columns x .. x + 3 of a PU over h rows
// (4, 8, 12 or 16) from y, with the chroma columns x / 2 .. x / 2 + 1 over
// h / 2 rows, carrying the PU's motion.
struct McStrip {
    uint16_t x, y;
    int16_t mv[2][2];
    int8_t ref[2];
    uint8_t flags;       // HEVCR_PRED_L0 | HEVCR_PRED_L1; none: a padding entry
    uint8_t h;
};
static_assert(sizeof(McStrip) == 16, "McStrip is four dwords");

// A band of a PU as the host hands it over: its first strip, where its n
// strips go in the frame's strip lists (strip k is the first moved 4k luma
// columns right), expanded on the device (k_strip_expand). A band with no
// prediction flags pads a height class to whole waves.
struct McBand {
    McStrip s;
    uint32_t dst, n;
};
static_assert(sizeof(McBand) == 24, "McBand is six dwords");

// One intra block of one component, in decoding order within its CTU, with
// its neighbour availability precomputed on the host. The reference samples
// form one line of 4n + 1 entries, index 0 the bottom of the bottom-left run,
// 2n the corner, 4n the end of the top-right run. The available entries are
// a run [lo, hi) with at most one hole: a side of the block is missing at the
// picture edge or when its neighbouring CTB lies in another slice segment or
// tile, the bottom-left and top-right extensions are prefixes of their runs
// (z-order is monotone along a row and a column), and a CTB above-left
// (above) that is available makes the left and above (left) ones available
// too; so the only hole is the corner alone (kIntraGapCorner: [2n, 2n + 1))
// or the corner and the top row (kIntraGapTop: [2n, 3n + 1)) between an
// available left column and an available top-right run. Substitution
// (intra_prediction.cpp:177-197) then clamps an index into [lo, hi - 1] and
// maps the hole to entry 2n - 1; lo == hi means no sample is available.
struct IntraOp {
    uint16_t x, y;       // component position
    uint8_t log2;        // component block size
    uint8_t mode;        // intra prediction mode
    uint8_t lo;          // first available reference index
    uint8_t hi;          // one past the last available reference index
    uint8_t flags;       // kIntraFilter | kIntraStrong | kIntraCbf
    int8_t ang;          // intraPredAngle of the mode
    uint16_t inv;        // |invAngle| of the mode (used for negative angles)
};
constexpr uint8_t kIntraFilter = 1;  // [1,2,1] reference smoothing applies (intra_prediction.cpp:217-230)
constexpr uint8_t kIntraStrong = 2;  // and the strong bilinear test is made (32x32 luma, SPS flag on)
constexpr uint8_t kIntraCbf = 4;     // residual present
constexpr uint8_t kIntraGapCorner = 8;  // entry 2n unavailable inside [lo, hi)
constexpr uint8_t kIntraGapTop = 16;    // entries 2n .. 3n unavailable inside [lo, hi)
constexpr uint8_t kIntraGroupStep = 0x80;  // a step of up to four 4x4 blocks side by side
constexpr uint8_t kIctlWaits = 1;    // a CTU with intra blocks has an intra neighbour above or left
constexpr uint8_t kIctlSignals = 2;  // and one below or right that waits for it
static_assert(sizeof(IntraOp) == 12, "IntraOp is three dwords");

// intraPredAngle and |invAngle| per mode, Tables 8-4/8-5 (reference: intra_prediction.cpp:13-29).
struct IntraAngles {
    int8_t ang[35];
    uint16_t inv[35];
};
constexpr IntraAngles kIntraAngles = {
    {0, 0, 32, 26, 21, 17, 13, 9, 5, 2, 0, -2, -5, -9, -13, -17, -21, -26,
     -32, -26, -21, -17, -13, -9, -5, -2, 0, 2, 5, 9, 13, 17, 21, 26, 32},
    {0, 0, 256, 315, 390, 482, 630, 910, 1638, 4096, 0, 4096, 1638, 910, 630, 482, 390, 315,
     256, 315, 390, 482, 630, 910, 1638, 4096, 0, 4096, 1638, 910, 630, 482, 390, 315, 256}};

// Frame constants.
struct FrameConst {
    int W, H, bd, maxv;
    int ctb_log2, wctb, hctb;
    int gw, gh;          // 4x4 grid
    int stride[3];       // plane strides in samples
    int slice_type;
    int weighted;        // explicit weighted prediction
    int luma_log2_wd, chroma_log2_wd;
    int num_ref[2];
    int ref_poc[2][kMaxRefs];
    int16_t wt[2][kMaxRefs][3];
    int16_t off[2][kMaxRefs][3];
    int pps_cb_qp_offset, pps_cr_qp_offset;
    int beta_offset_div2, tc_offset_div2;
    int strong_intra_smoothing, intra_smoothing_disabled;
    int deblock;         // deblocking runs (slice_deblocking_filter_disabled_flag = 0 in some slice segment)
    int pcm_lf_off;      // pcm_loop_filter_disabled_flag
};

// Residual blocks handled by one workgroup of the residual kernel, per size.
constexpr int kResPerGroup[4] = {64, 32, 16, 8};

// Scaling factors of a frame with scaling lists, expanded per block: for block
// size N = 4 << q and matrixId k (cIdx + 3 for non-intra CUs) the NxN factors
// m(x, y) sit row-major at kScaleOff[q] + k * N * N.
constexpr int kScaleOff[4] = {0, 6 * 16, 6 * 16 + 6 * 64, 6 * 16 + 6 * 64 + 6 * 256};
constexpr int kScaleBytes = kScaleOff[3] + 6 * 1024;

// Everything a kernel needs to know about one frame of a batch.
struct FrameDev {
    FrameConst fc;
    void* rec[3];                       // reconstruction / deblocking target planes
    void* out[3];                       // SAO destination planes (a DPB slot)
    int16_t* res[3];                    // intra residual planes (stride fc.stride)
    const void* ref[2][kMaxRefs][3];    // reference picture planes per list entry
    uint32_t* info4;                    // per-4x4 side information
    Motion4* mot;                       // per-4x4 motion
    // syntax, as handed over
    const hevcr_cu* cus;
    const uint32_t* tuw;                // per transform unit: x / 4, y / 4, log2 size - 2, coded luma (bits 0, 12, 24, 26)
    const hevcr_pu* pus;
    const int16_t* coeffs;              // dense coefficient pool, or
    const uint32_t* sig;                // compact coefficient lists (hevcr_frame_syntax::sig); null: dense
    const hevcr_sao* sao;
    // derived work lists (built when the frame is prepared)
    const uint32_t* cu_first_tu;        // ncu + 1 entries
    const uint32_t* cu_first_pu;        // ncu + 1 entries
    const uint32_t* ctu_ranges;         // per CTB: first / end CU, TU, PU (6 entries)
    const McTile* tiles;                // uni-predicted motion-compensation tiles (all but 8x8)
    const McTile* tilesb;               // bi-predicted tiles (all but 8x8)
    const McTile* tiles8;               // 8x8 tiles (8x8 PUs)
    const McStrip* strips;              // uni-predicted strips, by height class, each padded to 64 entries
    const McStrip* stripsb;             // bi-predicted strips, likewise
    const ResBlock* blocks[4];          // coded residual blocks per size 4..32
    const IntraOp* ops;                 // intra blocks per (CTU, component), in dependency-level steps
    const uint32_t* op_begin;           // 3 * nctb + 1 entries
    const uint8_t* op_span;             // per block: the size of the step it starts, or 0
    const uint32_t* ictu;               // CTUs with intra blocks, by dependency wave
    const uint32_t* diag_begin;         // waves + 1 entries into ictu
    uint32_t* iflags;                   // persistent intra: per CTU and component (3 * ctu + c), 1 once final
    const uint8_t* ictl;                // persistent intra: per CTU, kIctlWaits | kIctlSignals
    const uint8_t* scale_m;             // scaling factors (kScaleOff layout), or null: flat m = 16
    const uint32_t* ctb_ctl;            // per CTB control word (ctl::), or null: one slice segment, one tile,
                                        // no PCM / bypass CUs
    const uint8_t* ctb_seg;             // per CTB: its slice segment (with seg_w)
    const hevcr_weight* seg_w;          // per slice segment and list entry [seg][list][entry]: the weights
                                        // motion compensation applies, or null: fc.wt / fc.off
    const hevcr_weight* seg_wd;         // per slice segment (with seg_w): w[0], w[1] = its luma and chroma
                                        // log2 weight denominators
    uint32_t nblocks[4];
    uint32_t ntiles, ntiles8;           // uni-predicted tiles, 8x8 tiles
};

// One plane for the digest kernel.
struct PlaneRef {
    const void* p;
    int w, h, stride;
    int pad;
    uint64_t key;
};

// One output plane of an egress: w x h samples from src (row stride in
// samples, already offset to the crop window), written row after row without
// padding to dst (device memory or mapped pinned host memory).
struct EgressPlane {
    const void* src;
    void* dst;
    int w, h, stride, pad;
};

// One plane for the SHA-256 kernel: the message is the w x h samples from src
// (row stride in samples, already offset to the crop window), rows one after
// another, each sample as its bytes in memory (one byte at 8 bits, a 16-bit
// little-endian word above): the bytes write_yuv writes for the plane.
struct ShaPlane {
    const void* src;
    int w, h, stride, out;  // out: index of the plane's 8-word digest in the output array
};

// Work-range prefix for kernels whose grid spans several frames: frame f owns
// workgroups [prefix[f], prefix[f+1]). Workgroups never straddle frames, so
// the lookup is uniform across a workgroup.
__host__ __device__ inline int frame_of(const uint32_t* prefix, int nframes, uint32_t g) {
    int lo = 0, hi = nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

}  // namespace hevcr
