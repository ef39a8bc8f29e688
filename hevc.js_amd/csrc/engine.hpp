// Internal device-side descriptors shared by the host runtime and the HIP
// kernels of the reconstruction engine.
#pragma once

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
//   bits 16..21 luma intra mode
//   bits 22..27 chroma intra mode
namespace info {
constexpr uint32_t kLog2Mask = 7u;
constexpr uint32_t kCbf = 1u << 3;
constexpr uint32_t kEdgeV = 1u << 4;
constexpr uint32_t kEdgeH = 1u << 5;
constexpr int kModeShift = 6;
constexpr int kQpShift = 8;
constexpr int kLumaModeShift = 16;
constexpr int kChromaModeShift = 22;
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
    uint32_t coeff_off;  // into the coefficient pool (and the residual pool)
    uint16_t x, y;       // component sample position
    uint8_t log2;        // 2..5
    uint8_t cidx;
    uint8_t qp;          // qP' (QpBdOffset included)
    uint8_t flags;       // kResIntra | kResDst | kResTskip | kResToPlane
};
constexpr uint8_t kResIntra = 1, kResDst = 2, kResTskip = 4, kResToPlane = 8;

// One motion-compensation tile: a part of a PU at most 16x16 luma.
struct McTile {
    uint32_t pu;
    uint16_t x, y;  // luma position of the tile
    uint8_t w, h;   // luma size (multiples of 4)
    uint8_t pad[2];
};

// One intra block of one component, in decoding order within its CTU.
struct IntraOp {
    uint16_t x, y;       // component position
    uint8_t log2;        // component block size
    uint8_t mode;        // intra prediction mode
    uint8_t cbf;         // residual present
    uint8_t pad;
    uint32_t resid_off;  // into the residual pool
};

// Frame constants passed by value to every kernel.
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

// Reference picture planes per list entry.
struct RefPlanes {
    const void* p[2][kMaxRefs][3];
};

}  // namespace hevcr
