// Host front end: slice data (CTU syntax) -> reconstruction records.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "cabac.hpp"
#include "params.hpp"

namespace hevcf {

// Motion of one 4x4 block (or of one 16x16 block of a stored picture).
struct Mv4 {
    int16_t mv[2][2] = {{0, 0}, {0, 0}};
    int8_t ref[2] = {-1, -1};
    uint8_t flags = 0;  // bit 0: L0 used, bit 1: L1 used, bit 7: intra (no motion)
};

// A decoded picture as later pictures see it: POC, long-term marking, and
// the motion kept for temporal motion vector prediction (8.5.3.2.8).
struct DpbPic {
    int id = -1;                    // decoding index
    int slot = -1;                  // engine DPB slot
    int poc = 0;
    int cvs = 0;
    bool short_term = false, long_term = false, output_needed = false;
    int width = 0, height = 0;
    std::vector<Mv4> col;           // per 16x16 block, raster
    int col_stride = 0;
    int ref_poc[2][16] = {};        // POCs of this picture's reference lists when it was decoded
    bool ref_lt[2][16] = {};
    int num_ref[2] = {0, 0};
};

// The reference lists of the slice being parsed.
struct RefLists {
    int num[2] = {0, 0};
    DpbPic* pic[2][16] = {};
    int poc[2][16] = {};
    bool lt[2][16] = {};
};

// Records of one picture, in decoding order (the layout hevcr_frame_syntax points at).
struct PictureRecords {
    std::vector<hevcr_cu> cus;
    std::vector<hevcr_pu> pus;
    std::vector<hevcr_tu> tus;
    std::vector<int16_t> coeffs;      // coefficient levels; a PCM CU's samples
    std::vector<hevcr_sao> sao;
    bool has_scaling = false;
    hevcr_scaling scaling{};
    std::vector<hevcr_slice> slices;  // one per slice segment
    std::vector<uint8_t> ctb_slice;   // per CTB: its slice segment
    bool has_tiles = false;
    hevcr_tiles tiles{};
    int pcm_loop_filter_disabled = 0;
};

// Parses the slice segments of one picture, in decoding order, into records
// and fills the picture's temporal motion store when finished.
class PictureDecoder {
public:
    // `first`: the picture's first slice segment header (kept alive by the
    // caller for the decoder's lifetime, as are refs, cur and out).
    PictureDecoder(const Sps& sps, const Pps& pps, const SliceHeader& first, const RefLists& refs, DpbPic& cur,
                   PictureRecords& out);
    ~PictureDecoder();
    // One slice segment's data; `sh` stays alive until the next call.
    void segment(const SliceHeader& sh, const std::vector<uint8_t>& rbsp, const std::vector<uint32_t>& epb);
    bool complete() const;
    void finish();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace hevcf
