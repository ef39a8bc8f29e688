// Host front end: bit reader over an RBSP and Annex-B NAL splitting
// (H.265 7.3.1 / B.2; the reference's counterparts are
// src/bitstream/bitstream_reader.cpp and src/bitstream/nal_parser.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace hevcf {

// Raised on a malformed or unsupported stream; the C-ABI turns it into an
// error code and keeps the message.
struct ParseError : std::runtime_error {
    int code;
    ParseError(int c, const char* what) : std::runtime_error(what), code(c) {}
};

class Bits {
public:
    Bits() = default;
    Bits(const uint8_t* p, size_t n) : p_(p), n_(n) {}

    uint32_t u(int n) {
        uint32_t v = 0;
        for (int i = 0; i < n; i++) v = (v << 1) | bit();
        return v;
    }
    bool flag() { return bit() != 0; }
    uint32_t ue() {
        int z = 0;
        while (!bit()) {
            if (++z > 31) throw ParseError(-1, "exp-Golomb code too long");
        }
        return z ? ((1u << z) - 1) + u(z) : 0;
    }
    int32_t se() {
        const uint32_t k = ue();
        return (k & 1) ? static_cast<int32_t>((k + 1) / 2) : -static_cast<int32_t>(k / 2);
    }
    void skip(size_t nbits) {
        pos_ += nbits;
        if (pos_ > 8 * n_) throw ParseError(-1, "read past the end of a NAL unit");
    }
    void align() { pos_ = (pos_ + 7) & ~static_cast<size_t>(7); }
    bool aligned() const { return (pos_ & 7) == 0; }
    size_t pos() const { return pos_; }
    size_t bytes() const { return n_; }
    // more_rbsp_data(): anything but the trailing stop bit and zero bits left
    bool more_rbsp_data() const {
        if (pos_ >= 8 * n_) return false;
        size_t last = n_;
        while (last > 0 && p_[last - 1] == 0) last--;
        if (last == 0) return false;
        const uint8_t b = p_[last - 1];
        const size_t stop = 8 * (last - 1) + (7 - static_cast<size_t>(__builtin_ctz(b)));
        return pos_ < stop;
    }

private:
    uint32_t bit() {
        if (pos_ >= 8 * n_) throw ParseError(-1, "read past the end of a NAL unit");
        const uint32_t b = (p_[pos_ >> 3] >> (7 - (pos_ & 7))) & 1u;
        pos_++;
        return b;
    }
    const uint8_t* p_ = nullptr;
    size_t n_ = 0;
    size_t pos_ = 0;
};

// NAL unit types used here (Table 7-1).
enum NalType : int {
    NAL_TRAIL_N = 0, NAL_TRAIL_R = 1, NAL_TSA_N = 2, NAL_TSA_R = 3, NAL_STSA_N = 4, NAL_STSA_R = 5,
    NAL_RADL_N = 6, NAL_RADL_R = 7, NAL_RASL_N = 8, NAL_RASL_R = 9,
    NAL_BLA_W_LP = 16, NAL_BLA_W_RADL = 17, NAL_BLA_N_LP = 18, NAL_IDR_W_RADL = 19, NAL_IDR_N_LP = 20,
    NAL_CRA = 21, NAL_VPS = 32, NAL_SPS = 33, NAL_PPS = 34, NAL_AUD = 35, NAL_EOS = 36, NAL_EOB = 37,
};

struct Nal {
    int type = 0, layer = 0, tid = 0;
    std::vector<uint8_t> rbsp;         // payload after the 2-byte header, emulation prevention removed
    std::vector<uint32_t> epb;         // payload offsets (before removal) of the removed 0x03 bytes
};

// Payload of one NAL unit (header included) -> Nal. Returns false for an
// empty or forbidden-bit unit.
inline bool make_nal(const uint8_t* p, size_t n, Nal& out) {
    if (n < 2 || (p[0] & 0x80)) return false;
    out.type = (p[0] >> 1) & 0x3f;
    out.layer = ((p[0] & 1) << 5) | (p[1] >> 3);
    out.tid = (p[1] & 7) - 1;
    out.rbsp.clear();
    out.epb.clear();
    out.rbsp.reserve(n);
    int zeros = 0;
    for (size_t i = 2; i < n; i++) {
        const uint8_t b = p[i];
        if (zeros >= 2 && b == 3) {  // emulation_prevention_three_byte
            out.epb.push_back(static_cast<uint32_t>(i - 2));
            zeros = 0;
            continue;
        }
        out.rbsp.push_back(b);
        zeros = b == 0 ? zeros + 1 : 0;
    }
    return true;
}

// Finds Annex-B start codes in buf[from, n): returns the offset of the first
// byte after the next 0x000001, or n.
inline size_t next_start(const uint8_t* buf, size_t from, size_t n) {
    for (size_t i = from; i + 2 < n; i++)
        if (buf[i] == 0 && buf[i + 1] == 0 && buf[i + 2] == 1) return i + 3;
    return n;
}

}  // namespace hevcf
