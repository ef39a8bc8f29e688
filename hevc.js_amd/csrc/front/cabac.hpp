// Host front end: CABAC arithmetic decoder and context models (H.265 9.3).
// The decoder keeps the 9-bit range and offset of 9.3.4.3 with the offset
// followed by up to 8 bits read ahead, refilled a byte at a time.
#pragma once

#include <cstdint>
#include <cstring>

namespace hevcf {

// Context model offsets (one block per syntax element; sizes from Tables 9-5..9-37).
enum Ctx : int {
    C_SAO_MERGE = 0, C_SAO_TYPE = 1, C_SPLIT_CU = 2, C_TQ_BYPASS = 5, C_SKIP = 6, C_PRED_MODE = 9,
    C_PART_MODE = 10, C_PREV_INTRA = 14, C_CHROMA_MODE = 15, C_RQT_ROOT_CBF = 16, C_MERGE_FLAG = 17,
    C_MERGE_IDX = 18, C_INTER_PRED = 19, C_REF_IDX = 24, C_MVP = 26, C_SPLIT_TF = 27, C_CBF_LUMA = 30,
    C_CBF_CHROMA = 32, C_MVD_G0 = 36, C_MVD_G1 = 37, C_QP_DELTA = 38, C_TSKIP = 40, C_LAST_X = 42,
    C_LAST_Y = 60, C_CSBF = 78, C_SIG = 82, C_GT1 = 124, C_GT2 = 148, C_NUM = 154
};

// initValue per initType (0: I; 1: P, or B with cabac_init_flag; 2: B, or P
// with cabac_init_flag), in the order of Ctx. Values of H.265 Tables 9-5..9-37;
// 154 stands where a table has no entry for that initType.
inline const uint8_t* cabac_init_values(int init_type) {
    static const uint8_t kInit[3][C_NUM] = {
        {153, 200, 139, 141, 157, 154, 154, 154, 154, 154, 184, 154, 154, 154, 184, 63,
         154, 154, 154, 154, 154, 154, 154, 154, 154, 154, 154, 153, 138, 138, 111, 141,
         94, 138, 182, 154, 154, 154, 154, 154, 139, 139,
         110, 110, 124, 125, 140, 153, 125, 127, 140, 109, 111, 143, 127, 111, 79, 108, 123, 63,
         110, 110, 124, 125, 140, 153, 125, 127, 140, 109, 111, 143, 127, 111, 79, 108, 123, 63,
         91, 171, 134, 141,
         111, 111, 125, 110, 110, 94, 124, 108, 124, 107, 125, 141, 179, 153, 125, 107, 125, 141, 179, 153, 125,
         107, 125, 141, 179, 153, 125, 140, 139, 182, 182, 152, 136, 152, 136, 153, 136, 139, 111, 136, 139, 111,
         140, 92, 137, 138, 140, 152, 138, 139, 153, 74, 149, 92, 139, 107, 122, 152, 140, 179, 166, 182, 140, 227,
         122, 197,
         138, 153, 136, 167, 152, 152},
        {153, 185, 107, 139, 126, 154, 197, 185, 201, 149, 154, 139, 154, 154, 154, 152,
         79, 110, 122, 95, 79, 63, 31, 31, 153, 153, 168, 124, 138, 94, 153, 111,
         149, 107, 167, 154, 140, 198, 154, 154, 139, 139,
         125, 110, 94, 110, 95, 79, 125, 111, 110, 78, 110, 111, 111, 95, 94, 108, 123, 108,
         125, 110, 94, 110, 95, 79, 125, 111, 110, 78, 110, 111, 111, 95, 94, 108, 123, 108,
         121, 140, 61, 154,
         155, 154, 139, 153, 139, 123, 123, 63, 153, 166, 183, 140, 136, 153, 154, 166, 183, 140, 136, 153, 154,
         166, 183, 140, 136, 153, 154, 170, 153, 123, 123, 107, 121, 107, 121, 167, 151, 183, 140, 151, 183, 140,
         154, 196, 196, 167, 154, 152, 167, 182, 182, 134, 149, 136, 153, 121, 136, 137, 169, 194, 166, 167, 154, 167,
         137, 182,
         107, 167, 91, 122, 107, 167},
        {153, 160, 107, 139, 126, 154, 197, 185, 201, 134, 154, 139, 154, 154, 183, 152,
         79, 154, 137, 95, 79, 63, 31, 31, 153, 153, 168, 224, 167, 122, 153, 111,
         149, 92, 167, 154, 169, 198, 154, 154, 139, 139,
         125, 110, 124, 110, 95, 94, 125, 111, 111, 79, 125, 126, 111, 111, 79, 108, 123, 93,
         125, 110, 124, 110, 95, 94, 125, 111, 111, 79, 125, 126, 111, 111, 79, 108, 123, 93,
         121, 140, 61, 154,
         170, 154, 139, 153, 139, 123, 123, 63, 124, 166, 183, 140, 136, 153, 154, 166, 183, 140, 136, 153, 154,
         166, 183, 140, 136, 153, 154, 170, 153, 138, 138, 122, 121, 122, 121, 167, 151, 183, 140, 151, 183, 140,
         154, 196, 167, 167, 154, 152, 167, 182, 182, 134, 149, 136, 153, 121, 136, 122, 169, 208, 166, 167, 154, 152,
         167, 182,
         107, 167, 91, 107, 107, 167},
    };
    return kInit[init_type];
}

class Cabac {
public:
    // Context state: pStateIdx << 1 | valMps.
    uint8_t ctx[C_NUM];

    void init_contexts(int init_type, int slice_qp) {
        const uint8_t* iv = cabac_init_values(init_type);
        const int qp = slice_qp < 0 ? 0 : (slice_qp > 51 ? 51 : slice_qp);
        for (int i = 0; i < C_NUM; i++) {
            const int slope = (iv[i] >> 4) * 5 - 45, offset = ((iv[i] & 15) << 3) - 16;
            int pre = ((slope * qp) >> 4) + offset;
            pre = pre < 1 ? 1 : (pre > 126 ? 126 : pre);
            ctx[i] = static_cast<uint8_t>(pre <= 63 ? (63 - pre) << 1 : ((pre - 64) << 1) | 1);
        }
    }

    // Starts the arithmetic decoder on bytes [p, end) (9.3.2.5). value_ holds
    // the 9-bit ivlOffset above 7 bits read ahead (so ranges compare shifted
    // by 7); need_ counts up from -8 to the next byte refill.
    void start(const uint8_t* p, const uint8_t* end) {
        base_ = p;
        p_ = p;
        end_ = end;
        range_ = 510;
        value_ = (byte() << 8) | byte();
        need_ = -8;
    }

    int decode(int ci) {
        uint8_t& s = ctx[ci];
        const int state = s >> 1, mps = s & 1;
        const uint32_t lps = kLps[state][(range_ >> 6) & 3];
        range_ -= lps;
        const uint32_t scaled = range_ << 7;
        if (value_ < scaled) {  // most probable symbol
            s = static_cast<uint8_t>((kNextMps[state] << 1) | mps);
            if (range_ < 256) {
                range_ <<= 1;
                shift1();
            }
            return mps;
        }
        const int n = __builtin_clz(lps) - 23;  // renormalisation shifts: lps << n lands in [256, 510]
        value_ = (value_ - scaled) << n;
        range_ = lps << n;
        need_ += n;
        if (need_ >= 0) {
            value_ |= byte() << need_;
            need_ -= 8;
        }
        s = static_cast<uint8_t>((kNextLps[state] << 1) | (state == 0 ? 1 - mps : mps));
        return 1 - mps;
    }

    int bypass() {
        shift1();
        const uint32_t scaled = range_ << 7;
        if (value_ >= scaled) {
            value_ -= scaled;
            return 1;
        }
        return 0;
    }

    // A bypass bin of an escape code's unary prefix (exp-Golomb, the
    // prefix of coeff_abs_level_remaining): a plain bypass bin here; an
    // engine that chooses bins keeps such runs short.
    int bypass_prefix() { return bypass(); }

    uint32_t bypass_bits(int n) {
        uint32_t v = 0;
        for (int i = 0; i < n; i++) v = (v << 1) | static_cast<uint32_t>(bypass());
        return v;
    }

    // Terminating bins: end_of_slice_segment_flag, end_of_subset_one_bit,
    // pcm_flag (the role matters only to an engine that chooses the bins).
    enum Term { kEndOfSlice = 0, kEndOfSubset = 1, kPcmFlag = 2 };
    int terminate(int role = kEndOfSlice) {
        (void)role;
        range_ -= 2;
        const uint32_t scaled = range_ << 7;
        if (value_ >= scaled) return 1;
        if (range_ < 256) {
            range_ <<= 1;
            shift1();
        }
        return 0;
    }

    void save(uint8_t* out) const { std::memcpy(out, ctx, C_NUM); }
    void load(const uint8_t* in) { std::memcpy(ctx, in, C_NUM); }

    // After a terminating bin of 1: the byte after the one holding the next
    // bit of the codeword, which is the alignment (or stop) bit equal to one
    // that ends the byte. The decoder has taken 9 bits at its start and one
    // per renormalisation shift: 8 per byte read less the 7 - (need_ + 8)
    // read ahead.
    const uint8_t* aligned_end() const {
        const size_t taken = 8 * static_cast<size_t>(p_ - base_) - static_cast<size_t>(-need_ - 1);
        return base_ + taken / 8 + 1;
    }
    // The next substream, read on from the byte after the aligned end (tiles
    // without entry points; the reference reads tiles this way,
    // coding_tree.cpp:628-639).
    void restart_aligned() { start(aligned_end(), end_); }
    // PCM samples (7.3.8.7): raw bits from the byte after pcm_flag's
    // alignment, then the arithmetic decoder starts again where they end.
    void raw_begin() {
        raw_ = aligned_end();
        raw_bit_ = 0;
    }
    uint32_t raw_bits(int n) {
        uint32_t v = 0;
        for (int i = 0; i < n; i++) {
            const uint32_t byte = raw_ < end_ ? *raw_ : 0u;
            v = (v << 1) | ((byte >> (7 - raw_bit_)) & 1u);
            if (++raw_bit_ == 8) {
                raw_bit_ = 0;
                raw_++;
            }
        }
        return v;
    }
    void raw_end() { start(raw_ + (raw_bit_ ? 1 : 0), end_); }

    static constexpr uint8_t kLps[64][4] = {
        {128, 176, 208, 240}, {128, 167, 197, 227}, {128, 158, 187, 216}, {123, 150, 178, 205}, {116, 142, 169, 195},
        {111, 135, 160, 185}, {105, 128, 152, 175}, {100, 122, 144, 166}, {95, 116, 137, 158},  {90, 110, 130, 150},
        {85, 104, 123, 142},  {81, 99, 117, 135},   {77, 94, 111, 128},   {73, 89, 105, 122},   {69, 85, 100, 116},
        {66, 80, 95, 110},    {62, 76, 90, 104},    {59, 72, 86, 99},     {56, 69, 81, 94},     {53, 65, 77, 89},
        {51, 62, 73, 85},     {48, 59, 69, 80},     {46, 56, 66, 76},     {43, 53, 63, 72},     {41, 50, 59, 69},
        {39, 48, 56, 65},     {37, 45, 54, 62},     {35, 43, 51, 59},     {33, 41, 48, 56},     {32, 39, 46, 53},
        {30, 37, 43, 50},     {29, 35, 41, 48},     {27, 33, 39, 45},     {26, 31, 37, 43},     {24, 30, 35, 41},
        {23, 28, 33, 39},     {22, 27, 32, 37},     {21, 26, 30, 35},     {20, 24, 29, 33},     {19, 23, 27, 31},
        {18, 22, 26, 30},     {17, 21, 25, 28},     {16, 20, 23, 27},     {15, 19, 22, 25},     {14, 18, 21, 24},
        {14, 17, 20, 23},     {13, 16, 19, 22},     {12, 15, 18, 21},     {12, 14, 17, 20},     {11, 14, 16, 19},
        {11, 13, 15, 18},     {10, 12, 15, 17},     {10, 12, 14, 16},     {9, 11, 13, 15},      {9, 11, 12, 14},
        {8, 10, 12, 14},      {8, 9, 11, 13},       {7, 9, 11, 12},       {7, 9, 10, 12},       {7, 8, 10, 11},
        {6, 8, 9, 11},        {6, 7, 9, 10},        {6, 7, 8, 9},         {2, 2, 2, 2}};
    static constexpr uint8_t kNextLps[64] = {0,  0,  1,  2,  2,  4,  4,  5,  6,  7,  8,  9,  9,  11, 11, 12,
                                             13, 13, 15, 15, 16, 16, 18, 18, 19, 19, 21, 21, 22, 22, 23, 24,
                                             24, 25, 26, 26, 27, 27, 28, 29, 29, 30, 30, 30, 31, 32, 32, 33,
                                             33, 33, 34, 34, 35, 35, 35, 36, 36, 36, 37, 37, 37, 38, 38, 63};
    static constexpr uint8_t kNextMps[64] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16,
                                             17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                             33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48,
                                             49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60, 61, 62, 62, 63};

private:
    uint32_t byte() { return p_ < end_ ? *p_++ : 0u; }
    void shift1() {
        value_ <<= 1;
        if (++need_ == 0) {
            need_ = -8;
            value_ |= byte();
        }
    }

    const uint8_t* base_ = nullptr;  // start of the current codeword
    const uint8_t* p_ = nullptr;
    const uint8_t* end_ = nullptr;
    uint32_t range_ = 510, value_ = 0;
    int need_ = -8;
    const uint8_t* raw_ = nullptr;   // PCM sample reading
    int raw_bit_ = 0;
};

}  // namespace hevcf
