// Host front end: one picture's slice segments through the slice parser
// (slice_decoder.hpp) with the CABAC decoder.
#include "slice_data.hpp"

#include "slice_decoder.hpp"

namespace hevcf {

using sd::SliceDecoder;

struct PictureDecoder::Impl {
    Cabac cab;
    SliceDecoder<Cabac> dec;
    Impl(const Sps& sps, const Pps& pps, const SliceHeader& first, const RefLists& refs, DpbPic& cur, PictureRecords& out)
        : dec(sps, pps, first, refs, cur, out, cab) {}
};

PictureDecoder::PictureDecoder(const Sps& sps, const Pps& pps, const SliceHeader& first, const RefLists& refs, DpbPic& cur,
                               PictureRecords& out) {
    out.cus.clear();
    out.pus.clear();
    out.tus.clear();
    out.coeffs.clear();
    out.sao.clear();
    out.slices.clear();
    out.ctb_slice.clear();
    out.has_tiles = false;
    impl_ = std::make_unique<Impl>(sps, pps, first, refs, cur, out);
}

PictureDecoder::~PictureDecoder() = default;

void PictureDecoder::segment(const SliceHeader& sh, const std::vector<uint8_t>& rbsp, const std::vector<uint32_t>& epb) {
    impl_->dec.segment(sh, rbsp.data(), rbsp.size(), epb);
}

bool PictureDecoder::complete() const { return impl_->dec.complete(); }

void PictureDecoder::finish() { impl_->dec.finish(); }

}  // namespace hevcf
