// Internal device-side descriptors shared by the host runtime and the HIP
// kernels of the reconstruction engine.
//
// Every launch covers a batch of frames (one frame in the plain form): kernels
// receive an array of FrameDev descriptors and work items tagged with their
// frame's index in that array.
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
namespace info {
constexpr uint32_t kLog2Mask = 7u;
constexpr uint32_t kCbf = 1u << 3;
constexpr uint32_t kEdgeV = 1u << 4;
constexpr uint32_t kEdgeH = 1u << 5;
constexpr int kModeShift = 6;
constexpr int kQpShift = 8;
}  // namespace info

// Motion of one 4x4 cell (PUMotionInfo, inter_prediction.h:16-20).
struct Motion4 {
    int16_t mv[2][2];
    int8_t ref[2];
    uint8_t flags;
    uint8_t pad;
};

// One dequant + inverse-transform block (one component of one TU).
struct ResBlock {
    uint32_t coeff_off;  // into the frame's coefficient pool
    uint16_t x, y;       // component sample position
    uint8_t log2;        // 2..5
    uint8_t cidx;
    uint8_t qp;          // qP' (QpBdOffset included)
    uint8_t flags;       // kResIntra | kResDst | kResTskip | kResToPlane
    uint16_t frame;
    uint16_t pad;
};
constexpr uint8_t kResIntra = 1, kResDst = 2, kResTskip = 4, kResToPlane = 8;

// One motion-compensation tile: a part of a PU at most 16x16 luma, carrying
// the PU's motion.
struct McTile {
    uint16_t frame;
    uint16_t x, y;       // luma position of the tile
    uint8_t w, h;        // luma size (multiples of 4)
    int16_t mv[2][2];
    int8_t ref[2];
    uint8_t flags;       // HEVCR_PRED_L0 | HEVCR_PRED_L1
    uint8_t pad;
};

// One intra block of one component, in decoding order within its CTU, with
// its neighbour availability precomputed on the host. Availability along
// each side of a block is a prefix (z-order is monotone along a row and a
// column), so five counts describe it exactly: nbl bottom-left samples next
// to the left column, the left column (0 or n), the corner (0/1), the top
// row (0 or n), ntr top-right samples.
struct IntraOp {
    uint16_t x, y;       // component position
    uint8_t log2;        // component block size
    uint8_t mode;        // intra prediction mode
    uint8_t cbf;         // residual present
    uint8_t nbl;
    uint8_t left, corner, top, ntr;
};

struct IntraItem {
    uint16_t frame;
    uint16_t ctu;
};

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
};

// Everything a kernel needs to know about one frame of a batch.
struct FrameDev {
    FrameConst fc;
    void* rec[3];                       // reconstruction / deblocking target planes
    void* out[3];                       // SAO destination planes (a DPB slot)
    int16_t* res[3];                    // intra residual planes (stride fc.stride)
    const void* ref[2][kMaxRefs][3];    // reference picture planes per list entry
    uint32_t* info4;                    // per-4x4 side information
    Motion4* mot;                       // per-4x4 motion
    const hevcr_cu* cus;
    const hevcr_tu* tus;
    const hevcr_pu* pus;
    const uint32_t* cu_first_tu;        // ncu + 1 entries
    const uint32_t* cu_first_pu;        // ncu + 1 entries
    const int16_t* coeffs;
    const hevcr_sao* sao;
    const IntraOp* ops;
    const uint32_t* op_begin;           // 3 * nctb + 1 entries
};

// Work-range prefix for kernels whose grid spans several frames: frame f owns
// global items [prefix[f], prefix[f+1]).
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
