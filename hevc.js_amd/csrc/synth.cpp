// Seeded syntax and reference-picture generator (integer arithmetic only).
//
// The generated syntax obeys the structural rules the reference's parser
// enforces, so that the same records can be fed to the reference's
// reconstruction functions:
//   - CU quadtree with forced split across the picture edge
//     (src/decoding/coding_tree.cpp:692-715)
//   - intra NxN only at the minimum CB size, inter partitions as in
//     src/decoding/coding_tree.cpp:925-952
//   - TU quadtree with implicit split above the max TB size and for intra NxN
//     (src/decoding/coding_tree.cpp:1143-1174), chroma of 4x4 luma quads
//     carried by blkIdx 3 at the parent position (coding_tree.cpp:1335-1347)
//   - chroma intra mode derived from 5 candidates with the mode-34 substitution
//     (src/decoding/coding_tree.cpp:298-309)
//   - SaoOffsetVal sign convention of src/decoding/coding_tree.cpp:146-164
//   - reference lists cycling RPS pictures when more entries are active than
//     pictures (src/decoding/dpb.cpp:255-266, 282-295)
//   - slice segments covering runs of the CTB scan, tile by tile with tiles
//     (src/syntax/pps.cpp derive_tile_scan, coding_tree.cpp:536-671); every
//     segment predicts from the picture's lists, built from the first
//     segment (decoder.cpp:108-115)
//   - PCM only for intra 2Nx2N CUs of the PCM size range
//     (coding_tree.cpp:852-875)
#include "hevcr_synth.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// round(0.6 * exp(-0.15 * s) * 2^32), s = x + y in 0..62: significance thresholds.
const uint32_t kSigThr[63] = {
    2576980378u, 2218027566u, 1909074018u, 1643155235u, 1414276817u, 1217279337u,
    1047722035u, 901782713u,  776171574u,  668057065u,  575002044u,  494908846u,
    425971991u,  366637491u,  315567813u,  271611734u,  233778386u,  201214921u,
    173187288u,  149063680u,  128300298u,  110429090u,  95047199u,   81807882u,
    70412697u,   60604770u,   52163009u,   44897118u,   38643307u,   33260603u,
    28627666u,   24640061u,   21207897u,   18253806u,   15711196u,   13522752u,
    11639140u,   10017901u,   8622487u,    7421444u,    6387696u,    5497941u,
    4732121u,    4072975u,    3505642u,    3017334u,    2597043u,    2235296u,
    1923937u,    1655948u,    1425288u,    1226756u,    1055879u,    908803u,
    782214u,     673258u,     579479u,     498762u,     429288u,     369492u,
    318025u,     273726u,     235598u};

// round(sin(2*pi*i/256) * 1024)
const int16_t kSin[256] = {
    0, 25, 50, 75, 100, 125, 150, 175, 200, 224, 249, 273, 297, 321, 345, 369, 392, 415, 438,
    460, 483, 505, 526, 548, 569, 590, 610, 630, 650, 669, 688, 706, 724, 742, 759, 775, 792,
    807, 822, 837, 851, 865, 878, 891, 903, 915, 926, 936, 946, 955, 964, 972, 980, 987, 993,
    999, 1004, 1009, 1013, 1016, 1019, 1021, 1023, 1024, 1024, 1024, 1023, 1021, 1019, 1016,
    1013, 1009, 1004, 999, 993, 987, 980, 972, 964, 955, 946, 936, 926, 915, 903, 891, 878,
    865, 851, 837, 822, 807, 792, 775, 759, 742, 724, 706, 688, 669, 650, 630, 610, 590, 569,
    548, 526, 505, 483, 460, 438, 415, 392, 369, 345, 321, 297, 273, 249, 224, 200, 175, 150,
    125, 100, 75, 50, 25, 0, -25, -50, -75, -100, -125, -150, -175, -200, -224, -249, -273,
    -297, -321, -345, -369, -392, -415, -438, -460, -483, -505, -526, -548, -569, -590, -610,
    -630, -650, -669, -688, -706, -724, -742, -759, -775, -792, -807, -822, -837, -851, -865,
    -878, -891, -903, -915, -926, -936, -946, -955, -964, -972, -980, -987, -993, -999, -1004,
    -1009, -1013, -1016, -1019, -1021, -1023, -1024, -1024, -1024, -1023, -1021, -1019, -1016,
    -1013, -1009, -1004, -999, -993, -987, -980, -972, -964, -955, -946, -936, -926, -915,
    -903, -891, -878, -865, -851, -837, -822, -807, -792, -775, -759, -742, -724, -706, -688,
    -669, -650, -630, -610, -590, -569, -548, -526, -505, -483, -460, -438, -415, -392, -369,
    -345, -321, -297, -273, -249, -224, -200, -175, -150, -125, -100, -75, -50, -25};

constexpr uint32_t P(double p) { return static_cast<uint32_t>(p * 65536.0 + 0.5); }

class Rng {
public:
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& v : s_) v = splitmix(z);
    }
    uint64_t next() {
        const uint64_t result = rotl(s_[1] * 5, 7) * 9;
        const uint64_t t = s_[1] << 17;
        s_[2] ^= s_[0];
        s_[3] ^= s_[1];
        s_[1] ^= s_[2];
        s_[0] ^= s_[3];
        s_[2] ^= t;
        s_[3] = rotl(s_[3], 45);
        return result;
    }
    uint32_t u32() { return static_cast<uint32_t>(next() >> 32); }
    // Uniform integer in [lo, hi] (multiply-high range reduction).
    int uniform(int lo, int hi) {
        const uint64_t span = static_cast<uint64_t>(hi - lo) + 1;
        return lo + static_cast<int>((static_cast<uint64_t>(u32()) * span) >> 32);
    }
    bool bern16(uint32_t p16) { return (u32() >> 16) < p16; }
    bool bern32(uint32_t p32) { return u32() < p32; }
    // Irwin-Hall(12) approximation of N(0, sigma^2).
    int gauss(int sigma) {
        int64_t s = 0;
        for (int i = 0; i < 12; i++) s += static_cast<int64_t>(u32() >> 16);
        s -= 6 * 65536;  // mean 0, std 65536
        const int64_t v = s * sigma;
        return static_cast<int>(v >= 0 ? (v + 32768) >> 16 : -((-v + 32768) >> 16));
    }

private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint64_t s_[4];
};

}  // namespace

struct hevcr_synth_frame {
    hevcr_frame_syntax syn;
    std::vector<hevcr_slice> slices;
    std::vector<uint8_t> ctb_slice;
    hevcr_tiles tiles;
    std::vector<hevcr_cu> cus;
    std::vector<hevcr_pu> pus;
    std::vector<hevcr_tu> tus;
    std::vector<int16_t> coeffs;
    std::vector<hevcr_sao> sao;
    hevcr_scaling scaling;
    int num_ref_pics = 0;
};

namespace {

struct Gen {
    const hevcr_synth_cfg& c;
    Rng rng;
    hevcr_synth_frame& f;
    int n_active[2] = {0, 0};
    int slice_type;  // of the slice segment being generated

    Gen(const hevcr_synth_cfg& cfg, uint64_t seed, hevcr_synth_frame& out)
        : c(cfg), rng(seed), f(out), slice_type(cfg.slice_type) {}

    // Samples of one PCM block (pcm_bits bits each) at the end of the pool.
    uint32_t pcm_block(int log2n) {
        const int n = 1 << log2n;
        const uint32_t off = static_cast<uint32_t>(f.coeffs.size());
        f.coeffs.resize(f.coeffs.size() + static_cast<size_t>(n) * n, 0);
        for (int i = 0; i < n * n; i++) f.coeffs[off + i] = static_cast<int16_t>(rng.uniform(0, (1 << c.pcm_bits) - 1));
        return off;
    }

    // Dense NxN coefficient block at the end of the pool.
    uint32_t coeff_block(int log2n) {
        const int n = 1 << log2n;
        const uint32_t off = static_cast<uint32_t>(f.coeffs.size());
        f.coeffs.resize(f.coeffs.size() + static_cast<size_t>(n) * n, 0);
        int16_t* blk = f.coeffs.data() + off;
        bool any = false;
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                if (!rng.bern32(kSigThr[x + y])) continue;
                // |level| = 1 + Geometric(1/2): mean 2
                uint64_t bits = rng.next();
                int mag = 1;
                while ((bits & 1) && mag < 32) {
                    mag++;
                    bits >>= 1;
                }
                const int sign = (rng.u32() & 1) ? -1 : 1;
                blk[y * n + x] = static_cast<int16_t>(sign * mag * c.level_mul);
                any = true;
            }
        if (!any) blk[0] = static_cast<int16_t>(((rng.u32() & 1) ? -1 : 1) * c.level_mul);
        return off;
    }

    // Transform tree (coding_tree.cpp:1143-1248 structure). cbf_c_parent carries the
    // chroma cbf decided at the 8x8 node for its four 4x4 luma leaves.
    void tu_tree(uint32_t cu_idx, bool intra, bool nxn, int x0, int y0, int xb, int yb,
                 int log2, int depth, int blk, unsigned cbf_c_parent) {
        bool split;
        if (log2 > 5 || (nxn && depth == 0))
            split = true;
        else if (log2 > 2)
            split = rng.bern16(c.p_tu_split);
        else
            split = false;

        unsigned cbf_c = cbf_c_parent;
        if (log2 > 2) {
            cbf_c = 0;
            if (rng.bern16(c.p_cbf_chroma)) cbf_c |= HEVCR_CBF_CB;
            if (rng.bern16(c.p_cbf_chroma)) cbf_c |= HEVCR_CBF_CR;
        }
        if (split) {
            const int h = 1 << (log2 - 1);
            tu_tree(cu_idx, intra, nxn, x0, y0, x0, y0, log2 - 1, depth + 1, 0, cbf_c);
            tu_tree(cu_idx, intra, nxn, x0 + h, y0, x0, y0, log2 - 1, depth + 1, 1, cbf_c);
            tu_tree(cu_idx, intra, nxn, x0, y0 + h, x0, y0, log2 - 1, depth + 1, 2, cbf_c);
            tu_tree(cu_idx, intra, nxn, x0 + h, y0 + h, x0, y0, log2 - 1, depth + 1, 3, cbf_c);
            return;
        }
        hevcr_tu t{};
        t.x = static_cast<uint16_t>(x0);
        t.y = static_cast<uint16_t>(y0);
        t.xbase = static_cast<uint16_t>(xb);
        t.ybase = static_cast<uint16_t>(yb);
        t.log2_size = static_cast<uint8_t>(log2);
        t.blk_idx = static_cast<uint8_t>(blk);
        t.cu_idx = cu_idx;
        unsigned cbf = cbf_c;
        // An inter root TU without chroma residual always codes luma
        // (cbf_luma is inferred, coding_tree.cpp:1226-1230).
        if (!intra && depth == 0 && !(cbf_c & (HEVCR_CBF_CB | HEVCR_CBF_CR)))
            cbf |= HEVCR_CBF_Y;
        else if (rng.bern16(c.p_cbf_luma))
            cbf |= HEVCR_CBF_Y;
        t.cbf = static_cast<uint8_t>(cbf);
        if (cbf & HEVCR_CBF_Y) t.coeff_off[0] = coeff_block(log2);
        const bool chroma_here = (log2 > 2) || (blk == 3);
        const int log2c = log2 > 2 ? log2 - 1 : 2;
        if (chroma_here) {
            if (cbf & HEVCR_CBF_CB) t.coeff_off[1] = coeff_block(log2c);
            if (cbf & HEVCR_CBF_CR) t.coeff_off[2] = coeff_block(log2c);
        }
        f.tus.push_back(t);
    }

    static int derive_chroma(int coded, int luma) {
        if (coded == 4) return luma;
        static const int cand[4] = {0, 26, 10, 1};
        const int m = cand[coded];
        return m == luma ? 34 : m;
    }

    void make_pu(int x, int y, int w, int h) {
        hevcr_pu p{};
        p.x = static_cast<uint16_t>(x);
        p.y = static_cast<uint16_t>(y);
        p.w = static_cast<uint16_t>(w);
        p.h = static_cast<uint16_t>(h);
        unsigned flags;
        if (slice_type == HEVCR_SLICE_P) {
            flags = HEVCR_PRED_L0;
        } else {
            const bool small = (w + h) == 12;  // 8x4 / 4x8 are never bi-predicted
            if (!small && rng.bern16(c.p_bi))
                flags = HEVCR_PRED_L0 | HEVCR_PRED_L1;
            else
                flags = (rng.u32() & 1) ? HEVCR_PRED_L1 : HEVCR_PRED_L0;
        }
        p.pred_flags = static_cast<uint8_t>(flags);
        p.ref_idx[0] = p.ref_idx[1] = -1;
        for (int l = 0; l < 2; l++) {
            if (!(flags & (1u << l))) continue;
            int ri;
            if (c.weighted) {
                const bool wp = rng.bern16(c.p_wp);
                ri = rng.uniform(0, c.num_refs - 1) + (wp ? c.num_refs : 0);
            } else {
                ri = rng.uniform(0, n_active[l] - 1);
            }
            p.ref_idx[l] = static_cast<int8_t>(ri);
            for (int k = 0; k < 2; k++) {
                int v = rng.gauss(c.mv_sigma_qpel);
                if (v > c.mv_clamp_qpel) v = c.mv_clamp_qpel;
                if (v < -c.mv_clamp_qpel) v = -c.mv_clamp_qpel;
                p.mv[l][k] = static_cast<int16_t>(v);
            }
        }
        f.pus.push_back(p);
    }

    void make_cu(int x0, int y0, int log2) {
        const int size = 1 << log2;
        hevcr_cu cu{};
        cu.x = static_cast<uint16_t>(x0);
        cu.y = static_cast<uint16_t>(y0);
        cu.log2_size = static_cast<uint8_t>(log2);
        cu.qp_y = static_cast<int8_t>(rng.uniform(c.qp_min, c.qp_max));
        int mode = HEVCR_MODE_INTRA;
        if (slice_type != HEVCR_SLICE_I && !rng.bern16(c.p_intra_cu))
            mode = rng.bern16(c.p_skip) ? HEVCR_MODE_SKIP : HEVCR_MODE_INTER;
        cu.pred_mode = static_cast<uint8_t>(mode);
        const uint32_t cu_idx = static_cast<uint32_t>(f.cus.size());
        if (c.p_bypass && rng.bern16(c.p_bypass)) cu.flags |= HEVCR_CU_BYPASS;  // skip CUs too (coding_tree.cpp:757-760)

        if (mode == HEVCR_MODE_INTRA && c.p_pcm && log2 <= 5 && rng.bern16(c.p_pcm)) {
            cu.flags = HEVCR_CU_PCM;  // the reference keeps no bypass flag for PCM CUs
            cu.part_mode = HEVCR_PART_2Nx2N;
            cu.intra_luma[0] = 1;
            cu.intra_chroma = 1;
            f.cus.push_back(cu);
            hevcr_tu t{};
            t.x = t.xbase = static_cast<uint16_t>(x0);
            t.y = t.ybase = static_cast<uint16_t>(y0);
            t.log2_size = static_cast<uint8_t>(log2);
            t.cu_idx = cu_idx;
            t.cbf = HEVCR_CBF_Y | HEVCR_CBF_CB | HEVCR_CBF_CR;
            t.coeff_off[0] = pcm_block(log2);
            t.coeff_off[1] = pcm_block(log2 - 1);
            t.coeff_off[2] = pcm_block(log2 - 1);
            f.tus.push_back(t);
            return;
        }
        if (mode == HEVCR_MODE_INTRA) {
            const bool nxn = (log2 == 3) && rng.bern16(c.p_intra_nxn);
            cu.part_mode = nxn ? HEVCR_PART_NxN : HEVCR_PART_2Nx2N;
            const int npu = nxn ? 4 : 1;
            for (int i = 0; i < npu; i++) cu.intra_luma[i] = static_cast<uint8_t>(rng.uniform(0, 34));
            cu.intra_chroma = static_cast<uint8_t>(derive_chroma(rng.uniform(0, 4), cu.intra_luma[0]));
            f.cus.push_back(cu);
            tu_tree(cu_idx, true, nxn, x0, y0, x0, y0, log2, 0, 0, 0);
            return;
        }

        int part = HEVCR_PART_2Nx2N;
        if (mode == HEVCR_MODE_INTER && log2 > 3) {
            const int nparts = c.amp ? 7 : 3;
            static const int kParts[7] = {HEVCR_PART_2Nx2N, HEVCR_PART_2NxN, HEVCR_PART_Nx2N,
                                          HEVCR_PART_2NxnU, HEVCR_PART_2NxnD, HEVCR_PART_nLx2N,
                                          HEVCR_PART_nRx2N};
            part = kParts[rng.uniform(0, nparts - 1)];
        }
        cu.part_mode = static_cast<uint8_t>(part);
        const int s = size, h = size / 2, q = size / 4;
        const size_t pu0 = f.pus.size();
        switch (part) {
            case HEVCR_PART_2Nx2N: make_pu(x0, y0, s, s); break;
            case HEVCR_PART_2NxN: make_pu(x0, y0, s, h); make_pu(x0, y0 + h, s, h); break;
            case HEVCR_PART_Nx2N: make_pu(x0, y0, h, s); make_pu(x0 + h, y0, h, s); break;
            case HEVCR_PART_2NxnU: make_pu(x0, y0, s, q); make_pu(x0, y0 + q, s, s - q); break;
            case HEVCR_PART_2NxnD: make_pu(x0, y0, s, s - q); make_pu(x0, y0 + s - q, s, q); break;
            case HEVCR_PART_nLx2N: make_pu(x0, y0, q, s); make_pu(x0 + q, y0, s - q, s); break;
            case HEVCR_PART_nRx2N: make_pu(x0, y0, s - q, s); make_pu(x0 + s - q, y0, q, s); break;
        }
        cu.num_pu = static_cast<uint8_t>(f.pus.size() - pu0);
        f.cus.push_back(cu);
        if (mode == HEVCR_MODE_INTER && !rng.bern16(c.p_root_cbf_zero))
            tu_tree(cu_idx, false, false, x0, y0, x0, y0, log2, 0, 0, 0);
    }

    void cu_tree(int x0, int y0, int log2) {
        const int size = 1 << log2;
        bool split;
        if (x0 + size > c.width || y0 + size > c.height)
            split = log2 > 3;
        else
            split = log2 > 3 && rng.bern16(c.p_cu_split);
        if (!split) {
            make_cu(x0, y0, log2);
            return;
        }
        const int h = size / 2;
        cu_tree(x0, y0, log2 - 1);
        if (x0 + h < c.width) cu_tree(x0 + h, y0, log2 - 1);
        if (y0 + h < c.height) cu_tree(x0, y0 + h, log2 - 1);
        if (x0 + h < c.width && y0 + h < c.height) cu_tree(x0 + h, y0 + h, log2 - 1);
    }

    void sao_ctb(hevcr_sao& s) {
        std::memset(&s, 0, sizeof(s));
        if (!c.sao_enabled) return;
        const int maxoff = (1 << ((c.bit_depth < 10 ? c.bit_depth : 10) - 5)) - 1;
        for (int ci = 0; ci < 3; ci++) {
            int type;
            if (ci == 2) {
                type = s.type[1];  // SaoTypeIdx[2] = SaoTypeIdx[1]
            } else {
                type = rng.bern16(c.p_sao_off) ? 0 : rng.uniform(1, 2);
            }
            s.type[ci] = static_cast<uint8_t>(type);
            if (!type) continue;
            if (type == 2) {
                if (ci == 2)
                    s.eo_class[2] = s.eo_class[1];  // Cr shares Cb's class
                else
                    s.eo_class[ci] = static_cast<uint8_t>(rng.uniform(0, 3));
                const int a0 = rng.uniform(0, maxoff), a1 = rng.uniform(0, maxoff);
                const int a2 = rng.uniform(0, maxoff), a3 = rng.uniform(0, maxoff);
                s.offset[ci][0] = static_cast<int8_t>(a0);
                s.offset[ci][1] = static_cast<int8_t>(a1);
                s.offset[ci][2] = 0;
                s.offset[ci][3] = static_cast<int8_t>(-a2);
                s.offset[ci][4] = static_cast<int8_t>(-a3);
            } else {
                s.band_pos[ci] = static_cast<uint8_t>(rng.uniform(0, 31));
                for (int k = 0; k < 4; k++)
                    s.offset[ci][k] = static_cast<int8_t>(rng.uniform(-maxoff, maxoff));
                s.offset[ci][4] = 0;
            }
        }
    }

    void ref_lists() {
        hevcr_frame_params& p = f.syn.p;
        if (c.slice_type == HEVCR_SLICE_I) return;
        const int nb = c.num_refs_before, na = c.num_refs - c.num_refs_before;
        // ref picture k: k < nb has POC poc-1-k, else poc+1+(k-nb)
        std::vector<int> before, after;
        for (int k = 0; k < nb; k++) before.push_back(k);
        for (int k = 0; k < na; k++) after.push_back(nb + k);
        std::vector<int> temp0 = before, temp1 = after;
        temp0.insert(temp0.end(), after.begin(), after.end());
        temp1.insert(temp1.end(), before.begin(), before.end());
        const int active = c.weighted ? 2 * c.num_refs : c.num_refs;
        const int nl = c.slice_type == HEVCR_SLICE_B ? 2 : 1;
        for (int l = 0; l < nl; l++) {
            const std::vector<int>& temp = l == 0 ? temp0 : temp1;
            p.num_ref[l] = active;
            n_active[l] = active;
            for (int i = 0; i < active; i++) {
                const int k = temp[i % c.num_refs];
                p.ref_slot[l][i] = k;
                p.ref_poc[l][i] = k < nb ? c.poc - 1 - k : c.poc + 1 + (k - nb);
            }
        }
        if (c.weighted) {
            p.weighted_pred = 1;
            p.luma_log2_wd = 6;
            p.chroma_log2_wd = 6;
            for (int l = 0; l < nl; l++)
                for (int i = 0; i < active; i++) {
                    hevcr_weight& w = p.weights[l][i];
                    for (int ci = 0; ci < 3; ci++) {
                        if (i < c.num_refs) {
                            w.w[ci] = 64;
                            w.o[ci] = 0;
                        } else {
                            w.w[ci] = static_cast<int16_t>(rng.uniform(32, 96));
                            w.o[ci] = static_cast<int16_t>(rng.uniform(-16, 16));
                        }
                    }
                }
        }
        f.num_ref_pics = c.num_refs;
    }
};

bool cfg_valid(const hevcr_synth_cfg* c) {
    if (!c) return false;
    if (c->width < 8 || c->height < 8 || (c->width & 7) || (c->height & 7)) return false;
    if (c->width > 8192 || c->height > 4352) return false;
    if (c->bit_depth < 8 || c->bit_depth > 10) return false;
    if (c->ctb_log2 < 4 || c->ctb_log2 > 6) return false;
    if (c->slice_type < 0 || c->slice_type > 2) return false;
    if (c->qp_min < -6 * (c->bit_depth - 8) || c->qp_max > 51 || c->qp_min > c->qp_max) return false;
    if (c->level_mul < 1 || c->level_mul > 64) return false;
    if (c->num_slices < 0 || c->num_slices > HEVCR_MAX_SLICES) return false;
    if (c->tile_cols < 0 || c->tile_cols > HEVCR_MAX_TILE_COLS || c->tile_rows < 0 || c->tile_rows > HEVCR_MAX_TILE_ROWS)
        return false;
    {
        const int ctb = 1 << c->ctb_log2;
        const int wc = (c->width + ctb - 1) / ctb, hc = (c->height + ctb - 1) / ctb;
        if (c->tile_cols > wc || c->tile_rows > hc || c->num_slices > wc * hc) return false;
    }
    if (c->p_pcm && (c->pcm_bits < 1 || c->pcm_bits > c->bit_depth)) return false;
    if (c->slice_type != HEVCR_SLICE_I) {
        if (c->num_refs < 1 || c->num_refs > 8) return false;
        if (c->num_refs_before < 0 || c->num_refs_before > c->num_refs) return false;
        if (c->mv_clamp_qpel < 0 || c->mv_clamp_qpel > 4096) return false;
    }
    return true;
}

}  // namespace

extern "C" {

int hevcr_synth_preset(int preset, hevcr_synth_cfg* c) {
    if (!c) return -1;
    std::memset(c, 0, sizeof(*c));
    c->ctb_log2 = 6;
    c->poc = 8;
    c->level_mul = 1;
    c->bit_depth = 8;
    c->p_cu_split = P(0.5);
    c->p_tu_split = P(0.5);
    c->p_intra_nxn = P(0.5);
    c->sao_enabled = 1;
    c->strong_intra_smoothing = 1;
    c->amp = 1;
    c->p_bi = P(0.4);
    c->p_wp = P(0.2);
    c->p_root_cbf_zero = P(0.3);
    c->num_refs = 4;
    c->num_refs_before = 2;
    switch (preset) {
        case 0:
        case 2:
        case 4:
            c->width = preset == 0 ? 416 : (preset == 2 ? 1920 : 3840);
            c->height = preset == 0 ? 240 : (preset == 2 ? 1080 : 2160);
            c->slice_type = HEVCR_SLICE_I;
            c->qp_min = 22;
            c->qp_max = 37;
            c->p_cbf_luma = P(0.85);
            c->p_cbf_chroma = P(0.7);
            if (preset == 4) {
                c->bit_depth = 10;
                c->level_mul = 4;
                c->qp_min = 17;
            }
            return 0;
        case 1:
        case 3:
        case 5:
            c->width = preset == 3 ? 3840 : 1920;
            c->height = preset == 3 ? 2160 : 1080;
            c->slice_type = preset == 5 ? HEVCR_SLICE_P : HEVCR_SLICE_B;
            c->qp_min = 27;
            c->qp_max = 37;
            c->p_cbf_luma = P(0.7);
            c->p_cbf_chroma = P(0.5);
            c->p_skip = P(0.2);
            c->mv_sigma_qpel = 32;
            c->mv_clamp_qpel = 512;
            c->weighted = 1;
            if (preset == 5) c->num_refs_before = 4;
            if (preset == 3) {
                c->bit_depth = 10;
                c->level_mul = 4;
                c->qp_min = 17;
                c->mv_sigma_qpel = 96;
                c->mv_clamp_qpel = 1024;
            }
            return 0;
        default:
            return -1;
    }
}

hevcr_synth_frame* hevcr_synth_generate(const hevcr_synth_cfg* cfg, uint64_t seed) {
    if (!cfg_valid(cfg)) return nullptr;
    auto* f = new hevcr_synth_frame();
    std::memset(&f->syn, 0, sizeof(f->syn));
    hevcr_frame_params& p = f->syn.p;
    p.width = cfg->width;
    p.height = cfg->height;
    p.bit_depth = cfg->bit_depth;
    p.ctb_log2 = cfg->ctb_log2;
    p.min_cb_log2 = 3;
    p.slice_type = cfg->slice_type;
    p.poc = cfg->poc;
    p.strong_intra_smoothing = cfg->strong_intra_smoothing;
    p.sao_enabled = cfg->sao_enabled;
    p.deblocking_disabled = cfg->deblocking_disabled;

    Gen g(*cfg, seed, *f);
    g.ref_lists();
    const int ctb = 1 << cfg->ctb_log2;
    const int wc = (cfg->width + ctb - 1) / ctb, hc = (cfg->height + ctb - 1) / ctb;
    f->sao.resize(static_cast<size_t>(wc) * hc);
    // CTB scan: raster, or tile by tile (each tile in raster order)
    std::vector<int> scan;
    scan.reserve(static_cast<size_t>(wc) * hc);
    const int tcols = cfg->tile_cols > 1 ? cfg->tile_cols : 1, trows = cfg->tile_rows > 1 ? cfg->tile_rows : 1;
    if (tcols > 1 || trows > 1) {
        std::memset(&f->tiles, 0, sizeof(f->tiles));
        f->tiles.cols = tcols;
        f->tiles.rows = trows;
        f->tiles.loop_filter_across_tiles = cfg->loop_filter_across_tiles;
        // seeded widths: each tile column / row at least one CTB
        auto split = [&](int total, int parts, uint16_t* out) {
            std::vector<int> cut;  // parts - 1 distinct cut points in 1..total-1
            while ((int)cut.size() < parts - 1) {
                const int v = g.rng.uniform(1, total - 1);
                bool dup = false;
                for (int u : cut) dup = dup || u == v;
                if (!dup) cut.push_back(v);
            }
            std::sort(cut.begin(), cut.end());
            int prev = 0;
            for (int i = 0; i < parts - 1; i++) {
                out[i] = static_cast<uint16_t>(cut[static_cast<size_t>(i)] - prev);
                prev = cut[static_cast<size_t>(i)];
            }
            out[parts - 1] = static_cast<uint16_t>(total - prev);
        };
        split(wc, tcols, f->tiles.col_width);
        split(hc, trows, f->tiles.row_height);
        int y0 = 0;
        for (int tr = 0; tr < trows; tr++) {
            int x0 = 0;
            for (int tcl = 0; tcl < tcols; tcl++) {
                for (int y = y0; y < y0 + f->tiles.row_height[tr]; y++)
                    for (int x = x0; x < x0 + f->tiles.col_width[tcl]; x++) scan.push_back(y * wc + x);
                x0 += f->tiles.col_width[tcl];
            }
            y0 += f->tiles.row_height[tr];
        }
        f->syn.tiles = &f->tiles;
    } else {
        for (int a = 0; a < wc * hc; a++) scan.push_back(a);
    }
    // slice segments: seeded boundaries in the scan and parameters per segment
    const int nseg = cfg->num_slices > 1 ? cfg->num_slices : 1;
    std::vector<int> seg_start(1, 0);
    if (nseg > 1) {
        while ((int)seg_start.size() < nseg) {
            const int v = g.rng.uniform(1, wc * hc - 1);
            bool dup = false;
            for (int u : seg_start) dup = dup || u == v;
            if (!dup) seg_start.push_back(v);
        }
        std::sort(seg_start.begin(), seg_start.end());
        f->slices.resize(static_cast<size_t>(nseg));
        for (int k = 0; k < nseg; k++) {
            hevcr_slice& sl = f->slices[static_cast<size_t>(k)];
            if (k > 0 && g.rng.bern16(cfg->p_dependent)) {
                int ind = k - 1;  // the previous independent segment's parameters
                sl = f->slices[static_cast<size_t>(ind)];
                continue;
            }
            std::memset(&sl, 0, sizeof(sl));
            sl.slice_type = cfg->slice_type;
            if (k > 0 && cfg->slice_type != HEVCR_SLICE_I && g.rng.bern16(cfg->p_slice_type_change))
                sl.slice_type = cfg->slice_type == HEVCR_SLICE_B && (g.rng.u32() & 1) ? HEVCR_SLICE_P : HEVCR_SLICE_I;
            sl.slice_cb_qp_offset = g.rng.uniform(-12, 12);
            sl.slice_cr_qp_offset = g.rng.uniform(-12, 12);
            sl.deblocking_disabled = (g.rng.u32() & 3) == 0 ? 1 : cfg->deblocking_disabled;
            sl.beta_offset_div2 = g.rng.uniform(-6, 6);
            sl.tc_offset_div2 = g.rng.uniform(-6, 6);
            sl.loop_filter_across_slices = static_cast<int32_t>(g.rng.u32() & 1);
            if (p.weighted_pred && sl.slice_type != HEVCR_SLICE_I) {
                sl.weighted_pred = 1;
                sl.luma_log2_wd = p.luma_log2_wd;
                sl.chroma_log2_wd = p.chroma_log2_wd;
                for (int l = 0; l < 2; l++)
                    for (int i = 0; i < p.num_ref[l]; i++)
                        for (int ci = 0; ci < 3; ci++) {
                            const bool id = i < cfg->num_refs;  // the first num_refs entries keep identity weights
                            sl.weights[l][i].w[ci] = static_cast<int16_t>(id ? 64 : g.rng.uniform(32, 96));
                            sl.weights[l][i].o[ci] = static_cast<int16_t>(id ? 0 : g.rng.uniform(-16, 16));
                        }
            }
        }
        // segment 0 is the picture's: its parameters are p's too
        f->slices[0].slice_type = cfg->slice_type;
        const hevcr_slice& s0 = f->slices[0];
        p.slice_cb_qp_offset = s0.slice_cb_qp_offset;
        p.slice_cr_qp_offset = s0.slice_cr_qp_offset;
        p.deblocking_disabled = s0.deblocking_disabled;
        p.beta_offset_div2 = s0.beta_offset_div2;
        p.tc_offset_div2 = s0.tc_offset_div2;
        if (p.weighted_pred) std::memcpy(p.weights, s0.weights, sizeof(p.weights));
        f->ctb_slice.assign(static_cast<size_t>(wc) * hc, 0);
        f->syn.slices = f->slices.data();
        f->syn.num_slices = static_cast<uint32_t>(nseg);
        f->syn.ctb_slice = f->ctb_slice.data();
    }
    int seg = 0;
    for (size_t i = 0; i < scan.size(); i++) {
        while (seg + 1 < nseg && (int)i >= seg_start[static_cast<size_t>(seg) + 1]) seg++;
        const int a = scan[i], rx = a % wc, ry = a / wc;
        if (nseg > 1) {
            f->ctb_slice[static_cast<size_t>(a)] = static_cast<uint8_t>(seg);
            g.slice_type = f->slices[static_cast<size_t>(seg)].slice_type;
        }
        g.sao_ctb(f->sao[static_cast<size_t>(a)]);
        g.cu_tree(rx * ctb, ry * ctb, cfg->ctb_log2);
    }
    f->syn.pcm_loop_filter_disabled = cfg->pcm_loop_filter_disabled;

    f->syn.cus = f->cus.data();
    f->syn.num_cus = static_cast<uint32_t>(f->cus.size());
    f->syn.pus = f->pus.data();
    f->syn.num_pus = static_cast<uint32_t>(f->pus.size());
    f->syn.tus = f->tus.data();
    f->syn.num_tus = static_cast<uint32_t>(f->tus.size());
    f->syn.coeffs = f->coeffs.data();
    f->syn.sig = nullptr;  // dense coefficients (hevcr_syntax_compact makes the compact form)
    f->syn.num_sig = 0;
    f->syn.num_coeffs = f->coeffs.size();
    f->syn.sao = f->sao.data();
    if (cfg->scaling_lists) {
        // Factors rising away from DC with seeded jitter, 1..255, on their own
        // stream so that the rest of the syntax is the same with and without.
        Rng r(seed ^ 0x5CA11157ull);
        for (int sz = 0; sz < 4; sz++)
            for (int m = 0; m < 6; m++) {
                const int n = sz ? 8 : 4;
                for (int i = 0; i < n * n; i++) {
                    const int v = 8 + 3 * ((i % n) + (i / n)) + r.uniform(-6, 24);
                    f->scaling.list[sz][m][i] = static_cast<uint8_t>(v < 1 ? 1 : (v > 255 ? 255 : v));
                }
                for (int i = n * n; i < 64; i++) f->scaling.list[sz][m][i] = 0;
                if (sz >= 2) f->scaling.dc[sz - 2][m] = static_cast<uint8_t>(r.uniform(1, 40));
            }
        f->syn.scaling = &f->scaling;
    }
    return f;
}

const hevcr_frame_syntax* hevcr_synth_syntax(const hevcr_synth_frame* f) { return f ? &f->syn : nullptr; }

void hevcr_synth_free(hevcr_synth_frame* f) { delete f; }

int hevcr_synth_num_ref_pictures(const hevcr_synth_frame* f) { return f ? f->num_ref_pics : 0; }

int hevcr_synth_ref_picture(int width, int height, int bd, uint64_t seed, uint16_t* y, uint16_t* cb,
                            uint16_t* cr) {
    if (width < 8 || height < 8 || (width & 1) || (height & 1) || bd < 8 || bd > 12) return -1;
    if (!y || !cb || !cr) return -1;
    Rng rng(seed);
    const int maxv = (1 << bd) - 1;
    uint16_t* planes[3] = {y, cb, cr};
    for (int ci = 0; ci < 3; ci++) {
        const int w = ci ? width / 2 : width, h = ci ? height / 2 : height;
        // Uniform noise, then a 9x9 box mean via a summed-area table (edge-clamped).
        std::vector<int32_t> noise(static_cast<size_t>(w) * h);
        for (auto& v : noise) v = static_cast<int32_t>(rng.uniform(0, maxv));
        std::vector<int64_t> sat(static_cast<size_t>(w + 1) * (h + 1), 0);
        for (int j = 0; j < h; j++) {
            int64_t row = 0;
            for (int i = 0; i < w; i++) {
                row += noise[static_cast<size_t>(j) * w + i];
                sat[static_cast<size_t>(j + 1) * (w + 1) + i + 1] = sat[static_cast<size_t>(j) * (w + 1) + i + 1] + row;
            }
        }
        const int kx = 1 + rng.uniform(0, 3), ky = 1 + rng.uniform(0, 3);
        const int amp = (maxv + 1) / 4;
        for (int j = 0; j < h; j++) {
            const int y0 = j - 4 < 0 ? 0 : j - 4, y1 = j + 5 > h ? h : j + 5;
            for (int i = 0; i < w; i++) {
                const int x0 = i - 4 < 0 ? 0 : i - 4, x1 = i + 5 > w ? w : i + 5;
                const int64_t s = sat[static_cast<size_t>(y1) * (w + 1) + x1] - sat[static_cast<size_t>(y0) * (w + 1) + x1] -
                                  sat[static_cast<size_t>(y1) * (w + 1) + x0] + sat[static_cast<size_t>(y0) * (w + 1) + x0];
                const int64_t n = static_cast<int64_t>(y1 - y0) * (x1 - x0);
                const int mean = static_cast<int>((s + n / 2) / n);
                const int phase = ((i * kx * 256) / w + (j * ky * 256) / h) & 255;
                int v = mean + ((amp * kSin[phase]) >> 10);
                v = v < 0 ? 0 : (v > maxv ? maxv : v);
                planes[ci][static_cast<size_t>(j) * w + i] = static_cast<uint16_t>(v);
            }
        }
    }
    return 0;
}

}  // extern "C"
