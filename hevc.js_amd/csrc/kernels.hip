// HIP kernels of the reconstruction engine (gfx950, 64-lane wavefronts).
//
// Every launch covers a batch of frames (the plain form is a batch of one).
// Stage order (reference frame driver: decoder.cpp:52-304, re-ordered in a
// way that is bit-exact because of the dependency structure):
//   expand  : CU / TU / PU records -> per-4x4 side information and motion
//   mc      : inter prediction of every PU into the picture
//   residual: dequantisation + inverse transform of every coded block; inter
//             blocks are added to the prediction in place, intra residuals go
//             to a residual plane for the intra pass
//   intra   : CTU wavefront (dependency waves) of intra prediction + reconstruction
//   deblock : vertical then horizontal edges, fused per 64x64 tile in LDS
//   sao     : deblocked picture -> output picture, per CTB
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"
#include "tables.hpp"

namespace hevcr {

// Clip3 (H.265 5.8) for lo <= hi, which every call here has (coordinate
// ranges, sample ranges, +-tc): as max then min, which the compiler emits as
// v_max / v_min (or v_med3) rather than compare-and-select pairs.
__device__ __forceinline__ int clip3(int lo, int hi, int v) { return min(max(v, lo), hi); }
__device__ __forceinline__ int clip16(int v) { return clip3(-32768, 32767, v); }
__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

// Pointers read out of frame descriptors are generic (flat) pointers to the
// compiler, which then emits flat loads and stores. Every picture plane, table
// and staging buffer they point to is device global memory: the cast makes
// the accesses global_load / global_store. (The host pass parses these
// functions but emits none of them; it has no global address space.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using Glob = __attribute__((address_space(1))) T;
#else
template <typename T>
using Glob = T;
#endif
template <typename T>
__device__ __forceinline__ Glob<T>* gp(T* p) {
    return (Glob<T>*)p;
}

// Four consecutive samples (one 4-byte load for 8-bit pictures, 8 bytes above).
template <typename T>
struct Quad;
template <>
struct Quad<uint8_t> {
    static __device__ __forceinline__ void load(const uint8_t* p, int* v) {
        const uint32_t w = *gp(reinterpret_cast<const uint32_t*>(p));
        v[0] = w & 255;
        v[1] = (w >> 8) & 255;
        v[2] = (w >> 16) & 255;
        v[3] = w >> 24;
    }
    static __device__ __forceinline__ void store(uint8_t* p, const int* v) {
        *gp(reinterpret_cast<uint32_t*>(p)) = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
    }
};
template <>
struct Quad<uint16_t> {
    static __device__ __forceinline__ void load(const uint16_t* p, int* v) {
        const uint2 w = *gp(reinterpret_cast<const uint2*>(p));
        v[0] = w.x & 0xffff;
        v[1] = w.x >> 16;
        v[2] = w.y & 0xffff;
        v[3] = w.y >> 16;
    }
    static __device__ __forceinline__ void store(uint16_t* p, const int* v) {
        uint2 w;
        w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
        w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
        *gp(reinterpret_cast<uint2*>(p)) = w;
    }
};
// Two consecutive samples, packed as 16-bit halves of one word (low = first).
template <typename T>
struct Pair;
template <>
struct Pair<uint8_t> {
    static __device__ __forceinline__ uint32_t load(const uint8_t* p) {
        const uint32_t w = *gp(reinterpret_cast<const uint16_t*>(p));
        return (w & 255) | ((w >> 8) << 16);
    }
    static __device__ __forceinline__ void store(uint8_t* p, uint32_t v) {
        *gp(reinterpret_cast<uint16_t*>(p)) = (uint16_t)((v & 255) | (((v >> 16) & 255) << 8));
    }
};
template <>
struct Pair<uint16_t> {
    static __device__ __forceinline__ uint32_t load(const uint16_t* p) { return *gp(reinterpret_cast<const uint32_t*>(p)); }
    static __device__ __forceinline__ void store(uint16_t* p, uint32_t v) { *gp(reinterpret_cast<uint32_t*>(p)) = v; }
};

__device__ __forceinline__ void load4_i16(const int16_t* p, int* v) {
    const uint2 w = *gp(reinterpret_cast<const uint2*>(p));
    v[0] = (int16_t)(w.x & 0xffff);
    v[1] = (int16_t)(w.x >> 16);
    v[2] = (int16_t)(w.y & 0xffff);
    v[3] = (int16_t)(w.y >> 16);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2). Remap so that every XCD gets one contiguous run of the work list:
// neighbouring items then share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    const uint32_t q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------------
// Syntax expansion: the reference's grid writes done while parsing
// (coding_tree.cpp:956-1054 for CUs, 1233-1247 for TUs,
// inter_prediction.cpp:17-27 for motion). One workgroup per CTU with a lane
// per 4x4 cell of the launch's largest CTB (256 lanes for 64x64, 64 for 32x32,
// 16 for 16x16; a smaller CTB leaves the rest idle). The CTU's CU, TU and PU
// records (at most one per lane at each CTB size) are loaded into LDS
// together; CUs and transform leaves each cover one contiguous run of the
// CTU's z-order, so a cell finds its CU and its leaf by binary search over
// the records' first z-index, and its PU among its CU's (at most 4) PUs.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int zorder16(int x, int y) {  // 4-bit x, y
    int z = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) z |= (((x >> b) & 1) << (2 * b)) | (((y >> b) & 1) << (2 * b + 1));
    return z;
}

// Last index k in [0, n) with z0[k] <= z (z0 ascending, z0[0] <= z).
__device__ __forceinline__ int run_of(const uint16_t* z0, int n, int z) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (z0[mid] <= z)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_expand(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ ctb_prefix,
                                                int nframes) {
    __shared__ hevcr_cu scu[64];
    __shared__ uint16_t cu_z[64];
    __shared__ uint16_t cu_pu[65];  // first PU of each CU, relative to the CTU's first PU
    __shared__ uint16_t tu_z[256];
    __shared__ uint16_t tu_w[256];  // log2 size | cbf luma << 8
    __shared__ hevcr_pu spu[256];
    const uint32_t g = blockIdx.x;
    const int f = frame_of(ctb_prefix, nframes, g);
    const FrameDev& d = fr[f];
    const int ctb = (int)(g - ctb_prefix[f]);
    const int tid = threadIdx.x;
    const int ctb_log2 = d.fc.ctb_log2, wctb = d.fc.wctb, gw = d.fc.gw, gh = d.fc.gh;
    const int cx0 = (ctb % wctb) << ctb_log2, cy0 = (ctb / wctb) << ctb_log2;
    const Glob<const uint32_t>* rg = gp(d.ctu_ranges + 6 * ctb);  // CU, TU, PU ranges of this CTU
    const uint32_t c0 = rg[0], c1 = rg[1], t0 = rg[2], t1 = rg[3], p0 = rg[4], p1 = rg[5];
    const int ncu = (int)(c1 - c0), ntu = (int)(t1 - t0), npu = (int)(p1 - p0);
    // all record loads of the workgroup at once
    if (tid < ncu) {
        const hevcr_cu cu = gp(d.cus)[c0 + tid];
        scu[tid] = cu;
        cu_z[tid] = (uint16_t)zorder16((cu.x - cx0) >> 2, (cu.y - cy0) >> 2);
    }
    if (tid <= ncu) cu_pu[tid] = (uint16_t)(gp(d.cu_first_pu)[c0 + tid] - p0);
    if (tid < ntu) {
        const uint32_t t = gp(d.tuw)[t0 + tid];
        tu_z[tid] = (uint16_t)zorder16((int)(t & 0xfffu) - (cx0 >> 2), (int)((t >> 12) & 0xfffu) - (cy0 >> 2));
        tu_w[tid] = (uint16_t)((((t >> 24) & 3u) + 2) | ((t >> 26) & 1u) << 8);
    }
    if (tid < npu) spu[tid] = gp(d.pus)[p0 + tid];
    __syncthreads();
    const int i = tid & ((1 << (ctb_log2 - 2)) - 1), j = tid >> (ctb_log2 - 2);  // cell column / row in the CTU
    const int gx = (cx0 >> 2) + i, gy = (cy0 >> 2) + j;
    if ((i << 2) >= (1 << ctb_log2) || (j << 2) >= (1 << ctb_log2) || gx >= gw || gy >= gh) return;
    const int z = zorder16(i, j);
    const int k = run_of(cu_z, ncu, z);
    const hevcr_cu& cu = scu[k];
    const int cb = 1 << cu.log2_size;
    const int dx = (i << 2) + cx0 - cu.x, dy = (j << 2) + cy0 - cu.y;  // offset inside the CU
    int vx = -1, hy = -1;  // internal PU edge offsets
    switch (cu.part_mode) {
        case HEVCR_PART_Nx2N: vx = cb / 2; break;
        case HEVCR_PART_NxN: vx = cb / 2; hy = cb / 2; break;
        case HEVCR_PART_nLx2N: vx = cb / 4; break;
        case HEVCR_PART_nRx2N: vx = 3 * cb / 4; break;
        case HEVCR_PART_2NxN: hy = cb / 2; break;
        case HEVCR_PART_2NxnU: hy = cb / 4; break;
        case HEVCR_PART_2NxnD: hy = 3 * cb / 4; break;
        default: break;
    }
    // slice segment's deblocking offsets and boundary rules (ctl word), else the frame's
    const uint32_t cw = d.ctb_ctl ? gp(d.ctb_ctl)[ctb]
                                  : ((uint32_t)(d.fc.beta_offset_div2 + 6) << info::kBetaShift) |
                                        ((uint32_t)(d.fc.tc_offset_div2 + 6) << info::kTcShift);
    // A PCM CU writes no filter grids (coding_tree.cpp:852-875 returns before
    // 955-1054): its cells keep the picture's defaults, TU size = CTB size,
    // no coded luma and no edge flags of their own.
    const bool pcm = cu.flags & HEVCR_CU_PCM;
    uint32_t w = (uint32_t)(pcm ? ctb_log2 : cu.log2_size) | ((uint32_t)cu.pred_mode << info::kModeShift) |
                 ((uint32_t)(uint8_t)cu.qp_y << info::kQpShift) | (cw & ctl::kOffsets);
    if (pcm ? d.fc.pcm_lf_off != 0 : (cu.flags & HEVCR_CU_BYPASS) != 0) w |= info::kNoFilter;
    if (!pcm && (dx == 0 || dx == vx)) w |= info::kEdgeV;
    if (!pcm && (dy == 0 || dy == hy)) w |= info::kEdgeH;
    if (ntu > 0 && !pcm) {  // the transform leaf covering the cell, if any (none in CUs without residual)
        const int t = run_of(tu_z, ntu, z);
        const int l2 = tu_w[t] & 255, cells = 1 << (2 * (l2 - 2));
        if (tu_z[t] <= z && z < tu_z[t] + cells) {
            w = (w & ~(info::kLog2Mask | info::kCbf)) | (uint32_t)l2 | ((tu_w[t] & 256) ? info::kCbf : 0u);
            const int tz = tu_z[t];
            // the leaf's first cell: de-interleave its z-index
            int tx = 0, ty = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                tx |= ((tz >> (2 * b)) & 1) << b;
                ty |= ((tz >> (2 * b + 1)) & 1) << b;
            }
            if (i == tx) w |= info::kEdgeV;
            if (j == ty) w |= info::kEdgeH;
        }
    }
    // edges the reference excludes, decided by the Q side (this CTB): its
    // segment's deblocking off, or its left / top CTB boundary one that
    // filtering may not cross (deblocking.cpp:50-120)
    if (cw & ctl::kDbkOff) w &= ~(info::kEdgeV | info::kEdgeH);
    if (i == 0 && (cw & ctl::kLeftOff)) w &= ~info::kEdgeV;
    if (j == 0 && (cw & ctl::kTopOff)) w &= ~info::kEdgeH;
    gp(d.info4)[(size_t)gy * gw + gx] = w;
    if (d.mot) {
        const int q0 = cu_pu[k], q1 = cu_pu[k + 1];
        for (int q = q0; q < q1; q++) {
            const hevcr_pu& pu = spu[q];
            const int px = (i << 2) + cx0 - pu.x, py = (j << 2) + cy0 - pu.y;
            if (px >= 0 && px < pu.w && py >= 0 && py < pu.h) {
                Motion4 m;
                m.mv[0][0] = pu.mv[0][0];
                m.mv[0][1] = pu.mv[0][1];
                m.mv[1][0] = pu.mv[1][0];
                m.mv[1][1] = pu.mv[1][1];
                m.ref[0] = pu.ref_idx[0];
                m.ref[1] = pu.ref_idx[1];
                m.flags = pu.pred_flags;
                m.pad = 0;
                gp(d.mot)[(size_t)gy * gw + gx] = m;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Motion compensation (interpolation.cpp:45-411): one 64-lane workgroup per
// tile of at most 16x16 luma (+ two 8x8 chroma). All reference footprints of
// the tile (both lists, three components) are fetched into LDS as int16 with
// clamp-on-load (the reference's edge clamping); the horizontal pass writes
// its int16 results row-major, and the vertical pass produces 4 consecutive
// samples of a row per lane (vertical tap pairs gathered by byte permutes), so
// the weighted prediction is stored 4 samples at a time. Both passes multiply
// pairs of samples by pairs of taps with v_dot2_i32_i16 (exact: products and
// sums fit in int32); the reference stores both intermediates as int16, and
// so does this kernel. The separable 2-D form with the identity tap for a zero
// fraction equals the reference's H-only / V-only / copy paths exactly:
// 64 * s >> shift1 and (64 * h) >> 6 lose nothing.
// ---------------------------------------------------------------------------

typedef short hevcr_short2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int c) {
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(hevcr_short2, a), __builtin_bit_cast(hevcr_short2, b), c, false);
}

// Word of two 16-bit halves (lo, hi).
__device__ __forceinline__ uint32_t pack16(int lo, int hi) {
    return __builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, 0x05040100u);
}
__host__ __device__ constexpr uint32_t pack16c(int lo, int hi) {
    return (uint32_t)(lo & 0xffff) | ((uint32_t)(hi & 0xffff) << 16);
}

// Four consecutive outputs of an NT-tap filter from the 16-bit samples held
// in W (pairs: W[j] = (v[2j], v[2j + 1])): o[k] = sum_q tap[q] * v[k + q].
template <int NT>
__device__ __forceinline__ void filter4(const uint32_t* W, const uint32_t* T, int* o) {
    constexpr int NP = NT / 2;
    uint32_t Q[NP + 1];  // pairs shifted by one sample: (v[2j + 1], v[2j + 2])
#pragma unroll
    for (int j = 0; j <= NP; j++) Q[j] = __builtin_amdgcn_alignbit(W[j + 1], W[j], 16);
    o[0] = o[1] = o[2] = o[3] = 0;
#pragma unroll
    for (int j = 0; j < NP; j++) {
        o[0] = dot2(W[j], T[j], o[0]);
        o[1] = dot2(Q[j], T[j], o[1]);
        o[2] = dot2(W[j + 1], T[j], o[2]);
        o[3] = dot2(Q[j + 1], T[j], o[3]);
    }
}

// Four consecutive outputs of one row from the vertical NT-tap filter over
// NT rows of 4 columns, read row-major from t (row stride ld, 8-byte aligned
// rows): o[k] = sum_q tap[q] * t[q][k]. Vertical pairs (rows 2i, 2i + 1) of a
// column are gathered from the row words with byte permutes.
template <int NT>
__device__ __forceinline__ void vfilter4(const int16_t* t, int ld, const uint32_t* T, int* o) {
    uint2 r[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) r[q] = *reinterpret_cast<const uint2*>(t + q * ld);
    o[0] = o[1] = o[2] = o[3] = 0;
#pragma unroll
    for (int i = 0; i < NT / 2; i++) {
        const uint2 u = r[2 * i], v = r[2 * i + 1];
        o[0] = dot2(__builtin_amdgcn_perm(v.x, u.x, 0x05040100u), T[i], o[0]);
        o[1] = dot2(__builtin_amdgcn_perm(v.x, u.x, 0x07060302u), T[i], o[1]);
        o[2] = dot2(__builtin_amdgcn_perm(v.y, u.y, 0x05040100u), T[i], o[2]);
        o[3] = dot2(__builtin_amdgcn_perm(v.y, u.y, 0x07060302u), T[i], o[3]);
    }
}

// Weighted-prediction parameters of one component of a tile (interpolation.cpp:
// 215-300, 392-410), read from the frame constants when the tile starts: read
// where they are used, after the filters, their load latency would come on
// top of the tile's time.
struct McOut {
    int stride, maxv, shw, weighted, lwd, obd, w0, w1, o0, o1;  // o0, o1 in 8-bit units (scaled by 2^obd in use)
    void* plane;
};

// A 16-bit field through its aligned 32-bit word: with a uniform address this
// is a scalar load, whose value needs no transfer from a vector register.
__device__ __forceinline__ int load_i16_word(const int16_t* p) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint32_t w = *gp(reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3));
    return (a & 2) ? (int)(int16_t)(w >> 16) : (int)(int16_t)(w & 0xffff);
}

// lists: which lists' weights are loaded (bit l for list l; the others stay 0);
// (x, y): the luma position of the work item, whose CTB's slice segment gives
// the weights when segments weight the lists differently (FrameDev::seg_w).
__device__ __forceinline__ McOut mc_out(const FrameDev* dp, int c, int r0, int r1, int lists, int x, int y) {
    const FrameConst& fc = dp->fc;
    McOut o;
    o.plane = dp->rec[c];
    o.stride = fc.stride[c];
    o.maxv = fc.maxv;
    o.shw = 14 - fc.bd > 2 ? 14 - fc.bd : 2;  // shift1 of the weighting
    o.weighted = fc.weighted;
    o.lwd = (c ? fc.chroma_log2_wd : fc.luma_log2_wd) + o.shw;
    o.obd = fc.bd - 8;
    // loaded now, used only after the filters: no arithmetic on them here, so that
    // nothing waits for these loads before the footprint loads are issued
    o.w0 = o.w1 = o.o0 = o.o1 = 0;
    if (dp->seg_w) {
        const int seg = gp(dp->ctb_seg)[(y >> fc.ctb_log2) * fc.wctb + (x >> fc.ctb_log2)];
        const hevcr_weight* sw = dp->seg_w + seg * 2 * kMaxRefs;
        o.lwd = load_i16_word(&dp->seg_wd[seg].w[c ? 1 : 0]) + o.shw;
        if (lists & 1) {
            o.w0 = load_i16_word(&sw[r0].w[c]);
            o.o0 = load_i16_word(&sw[r0].o[c]);
        }
        if (lists & 2) {
            o.w1 = load_i16_word(&sw[kMaxRefs + r1].w[c]);
            o.o1 = load_i16_word(&sw[kMaxRefs + r1].o[c]);
        }
        return o;
    }
    if (lists & 1) {
        o.w0 = load_i16_word(&fc.wt[0][r0][c]);
        o.o0 = load_i16_word(&fc.off[0][r0][c]);
    }
    if (lists & 2) {
        o.w1 = load_i16_word(&fc.wt[1][r1][c]);
        o.o1 = load_i16_word(&fc.off[1][r1][c]);
    }
    return o;
}

// Cb's parameters where first, else Cr's (per lane: lanes of both chroma
// components run together).
__device__ __forceinline__ McOut mc_out_pick(const McOut& cb, const McOut& cr, bool first) {
    McOut o;
    o.plane = first ? cb.plane : cr.plane;
    o.stride = first ? cb.stride : cr.stride;
    o.maxv = cb.maxv;
    o.shw = cb.shw;
    o.weighted = cb.weighted;
    o.lwd = cb.lwd;
    o.obd = cb.obd;
    o.w0 = first ? cb.w0 : cr.w0;
    o.w1 = first ? cb.w1 : cr.w1;
    o.o0 = first ? cb.o0 : cr.o0;
    o.o1 = first ? cb.o1 : cr.o1;
    return o;
}

// Weighted sample prediction of 4 consecutive samples of one row (2 when
// !four) at (px, py) from the vertical-pass outputs a (list 0) and b (list 1),
// stored together; `pairs`: stored as two 2-sample stores (chroma rows start
// at even, not 4-aligned, columns).
template <typename T>
__device__ __forceinline__ void mc_weight_row(const McOut& p, int f0, int f1, const int* a, const int* b, int px, int py,
                                              bool four, bool pairs) {
    const int shw = p.shw;
    int v[4];
    if (!p.weighted) {
        if (f0 && f1) {
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = (a[k] + b[k] + (1 << shw)) >> (shw + 1);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = ((f0 ? a[k] : b[k]) + (1 << (shw - 1))) >> shw;
        }
    } else {
        const int lwd = p.lwd, o0 = p.o0 * (1 << p.obd), o1 = p.o1 * (1 << p.obd);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (f0 && f1)
                v[k] = (a[k] * p.w0 + b[k] * p.w1 + (o0 + o1 + 1) * (1 << lwd)) >> (lwd + 1);
            else
                v[k] = f0 ? ((a[k] * p.w0 + (1 << (lwd - 1))) >> lwd) + o0 : ((b[k] * p.w1 + (1 << (lwd - 1))) >> lwd) + o1;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = clip3(0, p.maxv, v[k]);
    T* dst = static_cast<T*>(p.plane) + py * p.stride + px;
    if (!pairs && four) {
        Quad<T>::store(dst, v);
    } else {
        Pair<T>::store(dst, (uint32_t)v[0] | ((uint32_t)v[1] << 16));
        if (four) Pair<T>::store(dst + 2, (uint32_t)v[2] | ((uint32_t)v[3] << 16));
    }
}

// Footprint rows starting at reference column xi, rows yi .. yi + ROWS - 1, all
// inside the picture, as int16 LDS rows of `ld` elements. Each task produces 4
// (8-bit) or 2 (16-bit) consecutive samples from two aligned 32-bit loads and
// a byte shift, so a row of 23 8-bit samples takes 6 tasks instead of 23
// single-sample loads. fetch() issues the loads into registers, commit() shifts
// and stores them (after the loads have landed). Reads go up to 8 bytes past
// the footprint's last sample: the pictures are padded.
template <typename T, int ROWS, int PER_ROW, int LANES>
struct FpWords {
    static constexpr int S = 4 / sizeof(T);  // samples per 32-bit word
    static constexpr int TASKS = ROWS * PER_ROW, ITERS = (TASKS + LANES - 1) / LANES;
    uint32_t lo[ITERS], hi[ITERS];
    __device__ __forceinline__ void fetch(const T* ref, int st, int xi, int yi, int lane) {
        const int xa = xi & ~(S - 1);
#pragma unroll
        for (int k = 0; k < ITERS; k++) {
            const int i = min(lane + LANES * k, TASKS - 1);
            const int r = i / PER_ROW, q = i - r * PER_ROW;
            const Glob<const uint32_t>* p = gp(reinterpret_cast<const uint32_t*>(ref + (yi + r) * st + xa + S * q));
            lo[k] = p[0];
            hi[k] = p[1];
        }
    }
    __device__ __forceinline__ void commit(int xi, int16_t* f, int ld, int lane) const {
        const int sh = (xi & (S - 1)) * (int)sizeof(T) * 8;  // bit shift from the aligned start
#pragma unroll
        for (int k = 0; k < ITERS; k++) {
            const int i = lane + LANES * k;
            if (i >= TASKS) break;
            const int r = i / PER_ROW, q = i - r * PER_ROW;
            const uint32_t w = (uint32_t)(((uint64_t)hi[k] << 32 | lo[k]) >> sh);  // samples S*q .. S*q + S - 1
            if (sizeof(T) == 1) {
                uint2 o;
                o.x = __builtin_amdgcn_perm(0u, w, 0x0c010c00u);  // bytes 0, 1 -> 16-bit lanes
                o.y = __builtin_amdgcn_perm(0u, w, 0x0c030c02u);  // bytes 2, 3
                *reinterpret_cast<uint2*>(f + r * ld + 4 * q) = o;
            } else {
                *reinterpret_cast<uint32_t*>(f + r * ld + 2 * q) = w;
            }
        }
    }
};

// LDS regions (int16 elements). Luma: footprint 24 x 24 per list (23 x 23
// used), H output 16 columns x 24 rows. Chroma, per list and
// component: footprint 12 x 12 (11 x 11 used), H output 8 columns x 12 rows.
constexpr int kFpY = 24 * 24, kFpC = 12 * 12, kTmpY = 16 * 24, kTmpC = 8 * 12;

// A workgroup works on two slots, each one tile and one list of its motion:
// a bi-predicted tile with its two lists (MC_BI), or two uni-predicted tiles,
// one list each (MC_UNI: twice the footprint loads in flight per wave, for the
// same LDS).
constexpr bool MC_UNI = true, MC_BI = false;

template <typename T, bool UNI>
__global__ void __launch_bounds__(64) k_mc(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ wg_prefix,
                                           int nframes) {
    __shared__ __align__(16) int16_t fp[2 * kFpY + 4 * kFpC];
    __shared__ __align__(16) int16_t tmp[2 * kTmpY + 4 * kTmpC];
    const uint32_t gi = xcd_remap(blockIdx.x, gridDim.x);
    const int fi = frame_of(wg_prefix, nframes, gi);
    const FrameDev* dp = &fr[fi];
    // the tile records, made wave-uniform so that everything derived from them is scalar
    const uint32_t first = (gi - wg_prefix[fi]) * (UNI ? 2 : 1);
    const int nt = UNI ? (first + 1 < dp->ntiles ? 2 : 1) : 1;  // tiles in this workgroup
    McTile tl[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const McTile* list = UNI ? dp->tiles : dp->tilesb;
        const Glob<const uint32_t>* tw32 = gp(reinterpret_cast<const uint32_t*>(list + first + (t < nt ? t : 0)));
        uint32_t w[5];
#pragma unroll
        for (int k = 0; k < 5; k++) w[k] = __builtin_amdgcn_readfirstlane(tw32[k]);
        __builtin_memcpy(&tl[t], w, sizeof(McTile));
    }
    // slot s: tile ts(s), list ls(s)
    auto ts = [&](int s) -> const McTile& { return tl[UNI ? s : 0]; };
    auto ls = [&](int s) -> int { return UNI ? ((tl[s].flags & 1) ? 0 : 1) : s; };
    auto on = [&](int s) -> bool { return UNI ? s < nt : ((tl[0].flags >> s) & 1) != 0; };
    // motion of list l (selects, not indexing: l is not a compile-time constant for uni tiles)
    auto mv = [](const McTile& t, int l, int k) -> int { return l ? t.mv[1][k] : t.mv[0][k]; };
    auto refi = [](const McTile& t, int l) -> int { return l ? t.ref[1] : t.ref[0]; };
    const int lane = threadIdx.x;
    const int W = dp->fc.W, H = dp->fc.H, bd = dp->fc.bd;
    const int sh1 = bd - 8 < 4 ? bd - 8 : 4;
    auto fp_at = [&](int l, int c) -> int16_t* { return c ? fp + 2 * kFpY + (l * 2 + c - 1) * kFpC : fp + l * kFpY; };
    auto tmp_at = [&](int l, int c) -> int16_t* { return c ? tmp + 2 * kTmpY + (l * 2 + c - 1) * kTmpC : tmp + l * kTmpY; };
    // lane -> (row, column) of the first footprint element it loads on the clamped
    // path; each further element is 64 positions on (luma: 2 rows + 16 columns,
    // chroma: 5 rows + 4 columns)
    const int ry0 = (lane * 2731) >> 16, cy0 = lane - 24 * ry0;  // lane / 24, lane % 24
    const int rc0 = (lane * 5462) >> 16, cc0 = lane - 12 * rc0;  // lane / 12, lane % 12
    // weighting parameters per slot (uni) or of the tile's two lists (bi)
    McOut po[2][3];
#pragma unroll
    for (int t = 0; t < (UNI ? 2 : 1); t++) {
        const int r0 = tl[t].ref[0] >= 0 ? tl[t].ref[0] : 0, r1 = tl[t].ref[1] >= 0 ? tl[t].ref[1] : 0;
        const int lists = UNI ? 1 << ls(t) : 3;
#pragma unroll
        for (int c = 0; c < 3; c++) po[t][c] = mc_out(dp, c, r0, r1, lists, tl[t].x, tl[t].y);
    }

    // phase A: every footprint of both slots. Interior footprints (the common case)
    // as aligned words, all components in flight together; the others element by
    // element with clamped addresses (the reference's edge clamping). Elements
    // outside the (w+7) x (h+7) footprint are never read later.
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (!on(s)) continue;
        const McTile& tile = ts(s);
        const int l = ls(s), tw = tile.w, th = tile.h;
        const int mvx = mv(tile, l, 0), mvy = mv(tile, l, 1), ri = refi(tile, l);
        {
            const int xy = tile.x + (mvx >> 2) - 3, yy = tile.y + (mvy >> 2) - 3;
            const int xc = (tile.x >> 1) + (mvx >> 3) - 1, yc = (tile.y >> 1) + (mvy >> 3) - 1;
            if (xy >= 0 && yy >= 0 && xy + tw + 7 <= W && yy + th + 7 <= H && xc >= 0 && yc >= 0 &&
                xc + (tw >> 1) + 3 <= (W >> 1) && yc + (th >> 1) + 3 <= (H >> 1)) {
                constexpr int S = 4 / (int)sizeof(T);
                FpWords<T, 23, 24 / S, 64> wy;
                FpWords<T, 11, 12 / S, 64> wc[2];
                wy.fetch(static_cast<const T*>(dp->ref[l][ri][0]), dp->fc.stride[0], xy, yy, lane);
#pragma unroll
                for (int c = 1; c < 3; c++)
                    wc[c - 1].fetch(static_cast<const T*>(dp->ref[l][ri][c]), dp->fc.stride[c], xc, yc, lane);
                wy.commit(xy, fp_at(s, 0), 24, lane);
#pragma unroll
                for (int c = 1; c < 3; c++) wc[c - 1].commit(xc, fp_at(s, c), 12, lane);
                continue;
            }
        }
        int vy[9], vc[2][3];
        {   // luma
            const int st = dp->fc.stride[0];
            const int xi = tile.x + (mvx >> 2) - 3, yi = tile.y + (mvy >> 2) - 3;
            const Glob<const T>* ref = gp(static_cast<const T*>(dp->ref[l][ri][0]));
            int r = ry0, col = cy0;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                vy[k] = ref[(uint32_t)(clip3(0, H - 1, yi + r) * st + clip3(0, W - 1, xi + col))];
                col += 16;
                r += 2;
                if (col >= 24) {
                    col -= 24;
                    r += 1;
                }
            }
        }
#pragma unroll
        for (int c = 1; c < 3; c++) {
            const int st = dp->fc.stride[c];
            const int xi = (tile.x >> 1) + (mvx >> 3) - 1, yi = (tile.y >> 1) + (mvy >> 3) - 1;
            const Glob<const T>* ref = gp(static_cast<const T*>(dp->ref[l][ri][c]));
            int r = rc0, col = cc0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                vc[c - 1][k] = ref[(uint32_t)(clip3(0, (H >> 1) - 1, yi + r) * st + clip3(0, (W >> 1) - 1, xi + col))];
                col += 4;
                r += 5;
                if (col >= 12) {
                    col -= 12;
                    r += 1;
                }
            }
        }
        {
            int16_t* f = fp_at(s, 0);
#pragma unroll
            for (int k = 0; k < 9; k++) f[lane + 64 * k] = (int16_t)vy[k];  // element i = r * 24 + col
        }
#pragma unroll
        for (int c = 1; c < 3; c++) {
            int16_t* f = fp_at(s, c);
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k < 2 || lane < kFpC - 128) f[lane + 64 * k] = (int16_t)vc[c - 1][k];
        }
    }
    __syncthreads();
    // phase B: horizontal passes; an item filters two rows x 4 columns, stored
    // row-major as 16-bit values
#pragma unroll
    for (int s = 0; s < 2; s++) {
        if (!on(s)) continue;
        const McTile& tile = ts(s);
        const int mvx = mv(tile, ls(s), 0), tw = tile.w, th = tile.h;
        if (lane < 48) {  // luma: 12 row pairs x 4 column groups
            const int rp = lane >> 2, g = (lane & 3) << 2;
            if (2 * rp < th + 7 && g < tw) {
                uint32_t tp[4];
#pragma unroll
                for (int j = 0; j < 4; j++) tp[j] = kLumaPairs[mvx & 3][j];
                const int16_t* f = fp_at(s, 0);
                uint2* t = reinterpret_cast<uint2*>(tmp_at(s, 0) + 2 * rp * 16 + g);
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint2* src = reinterpret_cast<const uint2*>(f + (2 * rp + h) * 24 + g);
                    const uint2 a = src[0], b = src[1], c = src[2];
                    const uint32_t Wd[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
                    int o[4];
                    filter4<8>(Wd, tp, o);
                    t[h * 4] = make_uint2(pack16(o[0] >> sh1, o[1] >> sh1), pack16(o[2] >> sh1, o[3] >> sh1));
                }
            }
        }
        if (lane < 24) {  // chroma: 2 components x 6 row pairs x 2 column groups
            const int c = 1 + (lane >= 12), j = lane % 12, rp = j >> 1, g = (j & 1) << 2;
            if (2 * rp < (th >> 1) + 3 && g < (tw >> 1)) {
                const uint32_t tp[2] = {kChromaPairs[mvx & 7][0], kChromaPairs[mvx & 7][1]};
                const int16_t* f = fp_at(s, c);
                uint2* t = reinterpret_cast<uint2*>(tmp_at(s, c) + 2 * rp * 8 + g);
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint2* src = reinterpret_cast<const uint2*>(f + (2 * rp + h) * 12 + g);
                    const uint2 a = src[0], b = src[1];
                    const uint32_t Wd[4] = {a.x, a.y, b.x, b.y};
                    int o[4];
                    filter4<4>(Wd, tp, o);
                    t[h * 2] = make_uint2(pack16(o[0] >> sh1, o[1] >> sh1), pack16(o[2] >> sh1, o[3] >> sh1));
                }
            }
        }
    }
    __syncthreads();
    // phase C: vertical passes + weighted sample prediction (interpolation.cpp:215-300,
    // 392-410). An item produces 4 consecutive samples of one row, stored together.
#pragma unroll
    for (int t = 0; t < (UNI ? 2 : 1); t++) {
        if (UNI && t >= nt) break;
        const McTile& tile = tl[t];
        const int tw = tile.w, th = tile.h;
        const int f0 = UNI ? ls(t) == 0 : tile.flags & 1, f1 = UNI ? ls(t) == 1 : (tile.flags >> 1) & 1;
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            // luma: item = (row, 4-column group); chroma: 2 components x 8 rows x 2 groups
            int c, y, xg;
            if (pass == 0) {
                c = 0;
                y = lane >> 2;
                xg = (lane & 3) << 2;
            } else {
                if (lane >= 32) break;
                c = 1 + (lane >> 4);
                y = (lane >> 1) & 7;
                xg = (lane & 1) << 2;
            }
            const int sub = pass;
            const int bw = tw >> sub, bh = th >> sub;
            if (y >= bh || xg >= bw) continue;
            int a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 2; s++) {  // the slots of this tile: its own (uni) or both lists (bi)
                if (UNI ? s != t : !on(s)) continue;
                const int l = ls(s), mvy = mv(tile, l, 1);
                int o[4];
                if (pass == 0) {
                    uint32_t tp[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) tp[j] = kLumaPairs[mvy & 3][j];
                    vfilter4<8>(tmp_at(s, 0) + y * 16 + xg, 16, tp, o);
                } else {
                    const uint32_t tp[2] = {kChromaPairs[mvy & 7][0], kChromaPairs[mvy & 7][1]};
                    vfilter4<4>(tmp_at(s, c) + y * 8 + xg, 8, tp, o);
                }
#pragma unroll
                for (int k = 0; k < 4; k++) (l ? b : a)[k] = (int16_t)(o[k] >> 6);
            }
            // chroma blocks 2 or 6 samples wide end inside a group
            mc_weight_row<T>(pass ? mc_out_pick(po[t][1], po[t][2], c == 1) : po[t][0], f0, f1, a, b,
                             (tile.x >> sub) + xg, (tile.y >> sub) + y, bw - xg >= 4, pass == 1);
        }
    }
}

// Four 8x8 tiles per 64-lane wave, 16 lanes each: 8x8 PUs are the most common
// ones, and a 16x16 tile's 64 lanes would be three-quarters idle for them.
// Same passes as k_mc; per slot and list the LDS holds the 15 x 15 luma
// footprint in 16 x 16, the H output as 8 columns x 16 rows, and per
// chroma component a 7 x 7 footprint in 8 x 8 and 4 columns x 8 rows.
constexpr int kF8Y = 16 * 16, kT8Y = 8 * 16, kF8C = 8 * 8, kT8C = 4 * 8;
constexpr int kMc8Slot = 2 * (kF8Y + kT8Y + 2 * (kF8C + kT8C));

template <typename T>
__global__ void __launch_bounds__(64) k_mc8(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ wg_prefix,
                                            int nframes) {
    __shared__ __align__(16) int16_t lds[4 * kMc8Slot];
    const uint32_t gi = xcd_remap(blockIdx.x, gridDim.x);
    const int fi = frame_of(wg_prefix, nframes, gi);
    const FrameDev* dp = &fr[fi];
    const int lane = threadIdx.x, slot = lane >> 4, u = lane & 15;
    const uint32_t first = (gi - wg_prefix[fi]) * 4, ti = first + slot;
    const bool active = ti < dp->ntiles8;  // a partial last group repeats its first tile and stores nothing
    const McTile tile = *gp(dp->tiles8 + (active ? ti : first));
    const int W = dp->fc.W, H = dp->fc.H, bd = dp->fc.bd;
    const int sh1 = bd - 8 < 4 ? bd - 8 : 4;
    const int flags = tile.flags;
    const int f0 = flags & 1, f1 = (flags >> 1) & 1;
    const int r0 = tile.ref[0] >= 0 ? tile.ref[0] : 0, r1 = tile.ref[1] >= 0 ? tile.ref[1] : 0;
    const McOut py = mc_out(dp, 0, r0, r1, 3, tile.x, tile.y), pcb = mc_out(dp, 1, r0, r1, 3, tile.x, tile.y),
                pcr = mc_out(dp, 2, r0, r1, 3, tile.x, tile.y);
    int16_t* base = lds + slot * kMc8Slot;
    auto fy = [&](int l) { return base + l * kF8Y; };
    auto ty = [&](int l) { return base + 2 * kF8Y + l * kT8Y; };
    auto fcm = [&](int l, int c) { return base + 2 * (kF8Y + kT8Y) + (l * 2 + c - 1) * kF8C; };
    auto tcm = [&](int l, int c) { return base + 2 * (kF8Y + kT8Y) + 4 * kF8C + (l * 2 + c - 1) * kT8C; };

    // phase A: footprints (aligned word loads inside the picture, else clamped
    // single-sample loads; loads issued together)
#pragma unroll
    for (int l = 0; l < 2; l++) {
        if (!((flags >> l) & 1)) continue;
        const int mvx = tile.mv[l][0], mvy = tile.mv[l][1];
        {
            const int xy = tile.x + (mvx >> 2) - 3, yy = tile.y + (mvy >> 2) - 3;
            const int xc = (tile.x >> 1) + (mvx >> 3) - 1, yc = (tile.y >> 1) + (mvy >> 3) - 1;
            if (xy >= 0 && yy >= 0 && xy + 15 <= W && yy + 15 <= H && xc >= 0 && yc >= 0 && xc + 7 <= (W >> 1) &&
                yc + 7 <= (H >> 1)) {
                constexpr int S = 4 / (int)sizeof(T);
                // the three components' loads are all in flight before the first wait
                FpWords<T, 15, 16 / S, 16> wy;
                FpWords<T, 7, 8 / S, 16> wc[2];
                wy.fetch(static_cast<const T*>(dp->ref[l][tile.ref[l]][0]), dp->fc.stride[0], xy, yy, u);
#pragma unroll
                for (int c = 1; c < 3; c++)
                    wc[c - 1].fetch(static_cast<const T*>(dp->ref[l][tile.ref[l]][c]), dp->fc.stride[c], xc, yc, u);
                wy.commit(xy, fy(l), 16, u);
#pragma unroll
                for (int c = 1; c < 3; c++) wc[c - 1].commit(xc, fcm(l, c), 8, u);
                continue;
            }
        }
        int vy[16], vc[2][4];
        {   // luma: lane u loads footprint column u, rows 0..15
            const int st = dp->fc.stride[0];
            const int xi = tile.x + (mvx >> 2) - 3, yi = tile.y + (mvy >> 2) - 3;
            const Glob<const T>* ref = gp(static_cast<const T*>(dp->ref[l][tile.ref[l]][0]));
            const int xx = clip3(0, W - 1, xi + u);
#pragma unroll
            for (int k = 0; k < 16; k++) vy[k] = ref[(uint32_t)(clip3(0, H - 1, yi + k) * st + xx)];
        }
#pragma unroll
        for (int c = 1; c < 3; c++) {  // chroma: element i = u + 16k of an 8 x 8 region
            const int st = dp->fc.stride[c];
            const int xi = (tile.x >> 1) + (mvx >> 3) - 1, yi = (tile.y >> 1) + (mvy >> 3) - 1;
            const Glob<const T>* ref = gp(static_cast<const T*>(dp->ref[l][tile.ref[l]][c]));
            const int xx = clip3(0, (W >> 1) - 1, xi + (u & 7));
#pragma unroll
            for (int k = 0; k < 4; k++)
                vc[c - 1][k] = ref[(uint32_t)(clip3(0, (H >> 1) - 1, yi + 2 * k + (u >> 3)) * st + xx)];
        }
#pragma unroll
        for (int k = 0; k < 16; k++) fy(l)[k * 16 + u] = (int16_t)vy[k];
#pragma unroll
        for (int c = 1; c < 3; c++)
#pragma unroll
            for (int k = 0; k < 4; k++) fcm(l, c)[u + 16 * k] = (int16_t)vc[c - 1][k];
    }
    __syncthreads();
    // phase B: horizontal passes (2 rows x 4 columns per item, stored row-major)
#pragma unroll
    for (int l = 0; l < 2; l++) {
        if (!((flags >> l) & 1)) continue;
        const int mvx = tile.mv[l][0];
        {   // luma: 8 row pairs x 2 column groups
            const int rp = u >> 1, g = (u & 1) << 2;
            uint32_t tp[4];
#pragma unroll
            for (int j = 0; j < 4; j++) tp[j] = kLumaPairs[mvx & 3][j];
            uint2* t = reinterpret_cast<uint2*>(ty(l) + 2 * rp * 8 + g);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const uint2* src = reinterpret_cast<const uint2*>(fy(l) + (2 * rp + h) * 16 + g);
                const uint2 p = src[0], q = src[1], r = src[2];
                const uint32_t Wd[6] = {p.x, p.y, q.x, q.y, r.x, r.y};
                int o[4];
                filter4<8>(Wd, tp, o);
                t[h * 2] = make_uint2(pack16(o[0] >> sh1, o[1] >> sh1), pack16(o[2] >> sh1, o[3] >> sh1));
            }
        }
        if (u < 8) {  // chroma: 2 components x 4 row pairs
            const int c = 1 + (u >> 2), rp = u & 3;
            const uint32_t tp[2] = {kChromaPairs[mvx & 7][0], kChromaPairs[mvx & 7][1]};
            uint2* t = reinterpret_cast<uint2*>(tcm(l, c) + 2 * rp * 4);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const uint2* src = reinterpret_cast<const uint2*>(fcm(l, c) + (2 * rp + h) * 8);
                const uint2 p = src[0], q = src[1];
                const uint32_t Wd[4] = {p.x, p.y, q.x, q.y};
                int o[4];
                filter4<4>(Wd, tp, o);
                t[h] = make_uint2(pack16(o[0] >> sh1, o[1] >> sh1), pack16(o[2] >> sh1, o[3] >> sh1));
            }
        }
    }
    __syncthreads();
    // phase C: vertical passes + weighting, 4 consecutive samples of one row per item
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
        int c, y, xg;
        if (pass == 0) {  // luma: 8 rows x 2 column groups
            c = 0;
            y = u >> 1;
            xg = (u & 1) << 2;
        } else {  // chroma: 2 components x 4 rows
            if (u >= 8) break;
            c = 1 + (u >> 2);
            y = u & 3;
            xg = 0;
        }
        int a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
#pragma unroll
        for (int l = 0; l < 2; l++) {
            if (!((flags >> l) & 1)) continue;
            const int mvy = tile.mv[l][1];
            int o[4];
            if (pass == 0) {
                uint32_t tp[4];
#pragma unroll
                for (int j = 0; j < 4; j++) tp[j] = kLumaPairs[mvy & 3][j];
                vfilter4<8>(ty(l) + y * 8 + xg, 8, tp, o);
            } else {
                const uint32_t tp[2] = {kChromaPairs[mvy & 7][0], kChromaPairs[mvy & 7][1]};
                vfilter4<4>(tcm(l, c) + y * 4, 4, tp, o);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) (l ? b : a)[k] = (int16_t)(o[k] >> 6);
        }
        if (active)  // 8x8 blocks: luma and chroma columns are 4-aligned
            mc_weight_row<T>(pass ? mc_out_pick(pcb, pcr, c == 1) : py, f0, f1, a, b, (tile.x >> pass) + xg,
                             (tile.y >> pass) + y, true, false);
    }
}

// ---------------------------------------------------------------------------
// Motion compensation by strips (interpolation.cpp:215-300, 392-410, the same
// arithmetic as k_mc). A lane predicts one strip: 4 luma columns of one PU
// over 4, 8, 12 or 16 rows and the matching 2 chroma columns of both
// components. Reference rows stream through registers: the horizontal pass
// of a row (8 taps at 4 outputs: v_dot4 over four 8-bit samples, v_dot2 over
// two 16-bit ones), rows paired per column as 16-bit halves, and the vertical
// pass (v_dot2 over those pairs) as soon as its rows exist. No LDS and no
// barriers: a wave is 64 independent strips of one height, and a lane's
// arithmetic is a fixed sequence whatever its fractions, so all 64 lanes do
// useful work.
//
// With 8-bit samples the v_dot4 operands are signed bytes: s ^ 0x80 is s - 128,
// and since every filter's taps sum to 64, the horizontal sums come out as
// H' = H - 8192 (in int16 range, as H is), which the vertical pass adds back
// as 64 * 8192. The vertical taps are scaled by 2^(10 - w) for the weighting
// shift w = 14 - bitDepth (at most 2^8 * 58: still int16), so that without
// weighting the reference's two roundings become one at bit 16:
// floor((floor(V / 64) + 2^(w-1)) / 2^w) = floor((V 2^(10-w) + 2^15) / 2^16),
// the high half of the vertical sum, which v_sat_pk_u8_i16 (8-bit) or packed
// 16-bit min / max (above) clamp two at a time.
// ---------------------------------------------------------------------------

// D = A . B + C over signed bytes (dot4) or signed 16-bit halves (dot2). With
// the clamp bit the compiler takes the three-operand encodings (the
// accumulating two-operand forms it otherwise picks need a copy of the start
// value for every sum); the sums here never reach the int32 limits, so the
// clamp never acts.
__device__ __forceinline__ int dot4_0(uint32_t a, uint32_t b) { return __builtin_amdgcn_sdot4((int)a, (int)b, 0, true); }
__device__ __forceinline__ int dot4_3(uint32_t a, uint32_t b, int c) {
    return __builtin_amdgcn_sdot4((int)a, (int)b, c, true);
}
__device__ __forceinline__ int dot2_0(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(hevcr_short2, a), __builtin_bit_cast(hevcr_short2, b), 0, true);
}
__device__ __forceinline__ int dot2_3(uint32_t a, uint32_t b, int c) {
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(hevcr_short2, a), __builtin_bit_cast(hevcr_short2, b), c, true);
}
// Two signed 16-bit halves saturated to 0..255 in bytes 0 and 1 (the upper
// half is not relied on).
__device__ __forceinline__ uint32_t sat_pk_u8(uint32_t v) {
    uint32_t d;
    asm("v_sat_pk_u8_i16 %0, %1" : "=v"(d) : "v"(v));
    return d;
}
// Both 16-bit halves of v times 2^s.
__device__ __forceinline__ uint32_t pk_shl16(uint32_t v, int s) {
    return ((v & 0xffffu) << s & 0xffffu) | ((v >> 16) << s << 16);
}

// Words of one reference row of a strip's window: NS samples from the
// window's first column on (8-bit: 4 a word, 16-bit: 2).
template <typename T, int NS>
struct StripRow {
    static constexpr int S = 4 / (int)sizeof(T);
    static constexpr int NW = (NS + S - 1) / S;        // words holding the window
    static constexpr int NR = (NS + 2 * S - 2) / S;    // aligned words covering it from any offset
};

// Loads rows [0, ROWS) of a window whose sample (0, 0) is at (xi, yi), all
// inside the picture: NR aligned words a row, shifted into place.
template <typename T, int NS, int ROWS>
__device__ __forceinline__ void strip_load(const T* plane, int stride, int xi, int yi, uint32_t (*e)[StripRow<T, NS>::NW]) {
    using R = StripRow<T, NS>;
    const int bits = (xi & (R::S - 1)) * 8 * (int)sizeof(T);
    const Glob<const uint32_t>* p = gp(reinterpret_cast<const uint32_t*>(plane + (size_t)yi * stride + (xi & ~(R::S - 1))));
    const int ws = stride / R::S;  // row stride in words (rows are word aligned)
    uint32_t raw[ROWS][R::NR];
#pragma unroll
    for (int r = 0; r < ROWS; r++)
#pragma unroll
        for (int j = 0; j < R::NR; j++) raw[r][j] = p[r * ws + j];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
#pragma unroll
        for (int j = 0; j < R::NW; j++) {
            const uint32_t hi = raw[r][j + 1 < R::NR ? j + 1 : j];
            e[r][j] = j + 1 < R::NR ? __builtin_amdgcn_alignbit(hi, raw[r][j], bits) : raw[r][j] >> bits;
        }
    }
}

// Horizontal pass of one row: NO outputs of an NT-tap filter (taps: 8-bit as
// signed bytes in th, giving H - 8192; 16-bit as pairs in tp, giving H >> sh1).
template <typename T, int NT, int NO>
__device__ __forceinline__ void strip_hrow(const uint32_t* e, const uint32_t* th, const uint32_t* tp, int sh1, int* o) {
    if constexpr (sizeof(T) == 1) {
        constexpr uint32_t kFlip = 0x80808080u;
        if constexpr (NT == 8) {
            const uint32_t f0 = e[0] ^ kFlip, f1 = e[1] ^ kFlip, f2 = e[2] ^ kFlip;
            o[0] = dot4_3(f0, th[0], dot4_0(f1, th[1]));
#pragma unroll
            for (int k = 1; k < NO; k++)
                o[k] = dot4_3(__builtin_amdgcn_alignbyte(f1, f0, k), th[0], dot4_0(__builtin_amdgcn_alignbyte(f2, f1, k), th[1]));
        } else {
            const uint32_t f0 = e[0] ^ kFlip, f1 = e[1] ^ kFlip;
            o[0] = dot4_0(f0, th[0]);
#pragma unroll
            for (int k = 1; k < NO; k++) o[k] = dot4_0(__builtin_amdgcn_alignbyte(f1, f0, k), th[0]);
        }
    } else {
        if constexpr (NT == 8) {
            // output k: pairs (v[k + 2j], v[k + 2j + 1]), j < 4: words from e, or shifted by one sample
            uint32_t q[5];
#pragma unroll
            for (int j = 0; j < 5; j++) q[j] = __builtin_amdgcn_alignbit(e[j + 1], e[j], 16);
            o[0] = dot2_3(e[0], tp[0], dot2_3(e[1], tp[1], dot2_3(e[2], tp[2], dot2_0(e[3], tp[3]))));
            o[1] = dot2_3(q[0], tp[0], dot2_3(q[1], tp[1], dot2_3(q[2], tp[2], dot2_0(q[3], tp[3]))));
            o[2] = dot2_3(e[1], tp[0], dot2_3(e[2], tp[1], dot2_3(e[3], tp[2], dot2_0(e[4], tp[3]))));
            o[3] = dot2_3(q[1], tp[0], dot2_3(q[2], tp[1], dot2_3(q[3], tp[2], dot2_0(q[4], tp[3]))));
        } else {
            const uint32_t q0 = __builtin_amdgcn_alignbit(e[1], e[0], 16), q1 = __builtin_amdgcn_alignbit(e[2], e[1], 16);
            o[0] = dot2_3(e[0], tp[0], dot2_0(e[1], tp[1]));
            o[1] = dot2_3(q0, tp[0], dot2_0(q1, tp[1]));
        }
#pragma unroll
        for (int k = 0; k < NO; k++) o[k] >>= sh1;
    }
}

// Both passes over the loaded rows of one list: emit(y, acc) receives the
// vertical sums (plus init) of output row y, in row order. Rows are paired
// as they come, so that only the last NT - 1 rows' pairs stay live.
template <typename T, int NT, int NO, int R, class Emit>
__device__ __forceinline__ void strip_filter(const uint32_t (*e)[StripRow<T, NO + NT - 1>::NW], const uint32_t* th,
                                             const uint32_t* tp, const uint32_t* tv, int sh1, int init, Emit emit) {
    constexpr int IN = R + NT - 1;
    uint32_t pr[IN - 1][NO];
    int prev[NO];
#pragma unroll
    for (int r = 0; r < IN; r++) {
        int cur[NO];
        strip_hrow<T, NT, NO>(e[r], th, tp, sh1, cur);
        if (r > 0) {
#pragma unroll
            for (int k = 0; k < NO; k++) {
                pr[r - 1][k] = pack16(prev[k], cur[k]);
                // an ordering point (it emits nothing): the row's horizontal
                // pass completes here, ahead of the next row's
                asm volatile("" : "+v"(pr[r - 1][k]));
            }
        }
#pragma unroll
        for (int k = 0; k < NO; k++) prev[k] = cur[k];
        const int y = r - (NT - 1);
        if (y >= 0) {
            int acc[NO];
#pragma unroll
            for (int k = 0; k < NO; k++) {
                acc[k] = init;
#pragma unroll
                for (int i = 0; i < NT / 2; i++) acc[k] = dot2_3(pr[y + 2 * i][k], tv[i], acc[k]);
            }
            emit(y, acc);
        }
    }
}

// A strip's record as separate values (no array a lane-varying index could
// reach: the record stays in registers).
struct StripArgs {
    int x, y, flags;
    int mvx[2], mvy[2], ref[2];
};

__device__ __forceinline__ StripArgs strip_args(const uint4& w) {
    StripArgs a;
    a.x = (int)(w.x & 0xffff);
    a.y = (int)(w.x >> 16);
    a.mvx[0] = (int)(int16_t)(w.y & 0xffff);
    a.mvy[0] = (int)(int16_t)(w.y >> 16);
    a.mvx[1] = (int)(int16_t)(w.z & 0xffff);
    a.mvy[1] = (int)(int16_t)(w.z >> 16);
    a.ref[0] = (int)(int8_t)(w.w & 0xff);
    a.ref[1] = (int)(int8_t)((w.w >> 8) & 0xff);
    a.flags = (int)((w.w >> 16) & 0xff);
    return a;
}

// Geometry of one component of a strip: C = 0 luma (4 columns, 8 taps), else
// chroma (2 columns, 4 taps).
template <typename T, int C>
struct StripPlane {
    static constexpr bool L = C == 0;
    static constexpr int NT = L ? 8 : 4, NO = L ? 4 : 2, NS = NO + NT - 1, SUB = L ? 0 : 1;
    static constexpr int NW = StripRow<T, NS>::NW;
};

// Loads the window rows of component C, list l, of a strip with R output rows.
template <typename T, int C, int R>
__device__ __forceinline__ void strip_fetch(const FrameDev* dp, const StripArgs& st, int l,
                                            uint32_t (*e)[StripPlane<T, C>::NW]) {
    using P = StripPlane<T, C>;
    const FrameConst& fc = dp->fc;
    const int mvx = l ? st.mvx[1] : st.mvx[0], mvy = l ? st.mvy[1] : st.mvy[0];
    const int ri = l ? st.ref[1] : st.ref[0];
    const int xi = P::L ? st.x + (mvx >> 2) - 3 : (st.x >> 1) + (mvx >> 3) - 1;
    const int yi = P::L ? st.y + (mvy >> 2) - 3 : (st.y >> 1) + (mvy >> 3) - 1;
    strip_load<T, P::NS, R + P::NT - 1>(static_cast<const T*>(dp->ref[l][ri][C]), fc.stride[C], xi, yi, e);
}

// Filters and writes component C of a strip from its loaded rows. For a
// bi-predicted strip, pass 0 (list 0) keeps its vertical outputs in `keep`
// and pass 1 (list 1) combines them.
// The filter taps of list l for component class C (luma, or both chroma
// components): horizontal ones as signed bytes (8-bit) or 16-bit pairs, the
// vertical ones as 16-bit pairs. Loaded before the window rows, so that the
// first rows can be filtered while later ones are still on their way.
struct StripTaps {
    uint32_t th[2], tp[4], tv[4];
};

template <typename T, int C>
__device__ __forceinline__ StripTaps strip_taps(const StripArgs& st, int l) {
    constexpr bool L = C == 0;
    const int mvx = l ? st.mvx[1] : st.mvx[0], mvy = l ? st.mvy[1] : st.mvy[0];
    const int fx = L ? mvx & 3 : mvx & 7, fy = L ? mvy & 3 : mvy & 7;
    StripTaps t = {{0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    if constexpr (L) {
        if constexpr (sizeof(T) == 1) {
            t.th[0] = kLumaI8[fx][0];
            t.th[1] = kLumaI8[fx][1];
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) t.tp[i] = kLumaPairs[fx][i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) t.tv[i] = kLumaPairs[fy][i];
    } else {
        if constexpr (sizeof(T) == 1) {
            t.th[0] = kChromaI8[fx];
        } else {
            t.tp[0] = kChromaPairs[fx][0];
            t.tp[1] = kChromaPairs[fx][1];
        }
        t.tv[0] = kChromaPairs[fy][0];
        t.tv[1] = kChromaPairs[fy][1];
    }
    return t;
}

template <typename T, bool BI, int C, int R>
__device__ __forceinline__ void strip_emit(const FrameDev* dp, const StripArgs& st, int pass, const StripTaps& taps,
                                           const uint32_t (*e)[StripPlane<T, C>::NW], uint32_t (*keep)[2]) {
    using P = StripPlane<T, C>;
    constexpr int NT = P::NT, NO = P::NO;
    const FrameConst& fc = dp->fc;
    const int l = BI ? pass : ((st.flags & 1) ? 0 : 1);
    const uint32_t* th = taps.th;
    const uint32_t* tp = taps.tp;
    uint32_t tv[NT / 2];
#pragma unroll
    for (int i = 0; i < NT / 2; i++) tv[i] = taps.tv[i];
    const int sh1 = fc.bd - 8 < 4 ? fc.bd - 8 : 4;
    const int shw = 14 - fc.bd > 2 ? 14 - fc.bd : 2;
    const int vs = 10 - shw;  // vertical taps times 2^vs: the plain result is the sum's high half
#pragma unroll
    for (int i = 0; i < NT / 2; i++) tv[i] = pk_shl16(tv[i], vs);
    // 8-bit: the horizontal sums are H - 8192, and the taps sum to 64
    const int base = sizeof(T) == 1 ? (64 * 8192) << vs : 0;
    const int stride = fc.stride[C];
    const int px = st.x >> P::SUB, py = st.y >> P::SUB;
    if (!BI && !fc.weighted) {
        // one list, no weighting: both roundings at bit 16
        T* dst = static_cast<T*>(dp->rec[C]) + py * stride + px;
        const uint32_t maxv2 = (uint32_t)fc.maxv * 0x10001u;
        strip_filter<T, NT, NO, R>(e, th, tp, tv, sh1, base + (1 << 15), [&](int y, const int* acc) {
            uint32_t h[2];
#pragma unroll
            for (int k = 0; k < NO / 2; k++) h[k] = __builtin_amdgcn_perm((uint32_t)acc[2 * k + 1], (uint32_t)acc[2 * k], 0x07060302u);
            if constexpr (sizeof(T) == 1) {
                if constexpr (NO == 4)
                    *gp(reinterpret_cast<uint32_t*>(dst + y * stride)) =
                        __builtin_amdgcn_perm(sat_pk_u8(h[1]), sat_pk_u8(h[0]), 0x05040100u);
                else
                    *gp(reinterpret_cast<uint16_t*>(dst + y * stride)) = (uint16_t)sat_pk_u8(h[0]);
            } else {
#pragma unroll
                for (int k = 0; k < NO / 2; k++) {
                    const hevcr_short2 v = __builtin_bit_cast(hevcr_short2, h[k]);
                    const hevcr_short2 lo = {0, 0}, hi = __builtin_bit_cast(hevcr_short2, maxv2);
                    h[k] = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_elementwise_max(v, lo), hi));
                }
                if constexpr (NO == 4)
                    *gp(reinterpret_cast<uint2*>(dst + y * stride)) = make_uint2(h[0], h[1]);
                else
                    *gp(reinterpret_cast<uint32_t*>(dst + y * stride)) = h[0];
            }
        });
        return;
    }
    const bool last = !BI || pass == 1;
    const McOut po = mc_out(dp, C, st.ref[0] >= 0 ? st.ref[0] : 0, st.ref[1] >= 0 ? st.ref[1] : 0, st.flags & 3, st.x,
                            st.y);
    const int f0 = st.flags & 1, f1 = (st.flags >> 1) & 1;
    strip_filter<T, NT, NO, R>(e, th, tp, tv, sh1, base, [&](int y, const int* acc) {
        int a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < NO; k++) (l ? b : a)[k] = (int16_t)(acc[k] >> (6 + vs));
        if (!last) {
#pragma unroll
            for (int k = 0; k < NO / 2; k++) keep[y][k] = pack16(a[2 * k], a[2 * k + 1]);
            return;
        }
        if (BI) {
#pragma unroll
            for (int k = 0; k < NO; k++) a[k] = (int16_t)(keep[y][k / 2] >> (16 * (k & 1)));
        }
        mc_weight_row<T>(po, f0, f1, a, b, px, py + y, NO == 4, !P::L);
    });
}

// Luma (LUMA) or both chroma components of one strip of R luma rows, a list
// at a time. The window rows of a list are all loaded before anything is
// written: loads cannot move past the stores, since the compiler cannot tell
// the planes apart.
template <typename T, bool BI, bool LUMA, int R>
__device__ __forceinline__ void strip_body(const FrameDev* dp, const StripArgs& st) {
    constexpr int RC = R / 2;
#pragma unroll
    for (int pass = 0; pass < (BI ? 2 : 1); pass++) {
        const int l = BI ? pass : ((st.flags & 1) ? 0 : 1);
        if constexpr (LUMA) {
            uint32_t kY[BI ? R : 1][2];
            uint32_t eY[R + 7][StripPlane<T, 0>::NW];
            const StripTaps taps = strip_taps<T, 0>(st, l);
            strip_fetch<T, 0, R>(dp, st, l, eY);
            strip_emit<T, BI, 0, R>(dp, st, pass, taps, eY, kY);
        } else {
            uint32_t kB[BI ? RC : 1][2], kR[BI ? RC : 1][2];
            uint32_t eB[RC + 3][StripPlane<T, 1>::NW], eR[RC + 3][StripPlane<T, 2>::NW];
            const StripTaps taps = strip_taps<T, 1>(st, l);
            strip_fetch<T, 1, RC>(dp, st, l, eB);
            strip_fetch<T, 2, RC>(dp, st, l, eR);
            strip_emit<T, BI, 1, RC>(dp, st, pass, taps, eB, kB);
            strip_emit<T, BI, 2, RC>(dp, st, pass, taps, eR, kR);
        }
    }
}

// A wave of 64 strips of one height (at most MAXR rows), their luma (LUMA) or
// their chroma. The lists pad each height class to a multiple of 64 with
// empty strips (no prediction flags). Every strip's window lies inside the
// picture: PUs whose footprint needs the reference's edge clamping go to the
// tile kernels.
template <typename T, bool BI, bool LUMA, int MAXR>
__global__ void __launch_bounds__(64) k_mcs(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ wave_prefix,
                                            int nframes) {
    const uint32_t gi = xcd_remap(blockIdx.x, gridDim.x);
    const int fi = frame_of(wave_prefix, nframes, gi);
    const FrameDev* dp = &fr[fi];
    const McStrip* list = BI ? dp->stripsb : dp->strips;
    const uint4 w = *gp(reinterpret_cast<const uint4*>(list + (size_t)(gi - wave_prefix[fi]) * 64 + threadIdx.x));
    const StripArgs st = strip_args(w);
    const int rows = __builtin_amdgcn_readfirstlane((int)(w.w >> 24));
    if (!(st.flags & 3)) return;
    if constexpr (MAXR > 8) {
        if (rows == 16) return strip_body<T, BI, LUMA, 16>(dp, st);
        if (rows == 12) return strip_body<T, BI, LUMA, 12>(dp, st);
    }
    if (rows == 8) return strip_body<T, BI, LUMA, 8>(dp, st);
    strip_body<T, BI, LUMA, 4>(dp, st);
}

// ---------------------------------------------------------------------------
// Dequantisation + inverse transform of 8x8, 16x16 and 32x32 blocks
// (transform.cpp:45-233, 240-287, 293-368): a lane per column for the
// vertical pass, then a lane per row for the horizontal one, 256 / N blocks
// per workgroup. Each 1-D transform is an even/odd decomposition of the
// matrix product (the reference's partial butterflies: the same integer sums,
// regrouped), with the odd parts as v_dot2 products of packed 16-bit input
// pairs and compile-time coefficient pairs. Dequantisation is exact in 32
// bits: with P = c * m * levelScale (m = 16 flat, else the scaling factor of
// the coefficient, at most 255: |P| < 2^30), s = qP / 6 and the shift b, the
// reference's int64 value ((P << s) + 2^(b-1)) >> b truncated to int is
// P << (s - b) wrapped to 32 bits when s >= b, else (P + 2^(b-s-1)) >> (b - s).
// ---------------------------------------------------------------------------


// N-point inverse transform of one line: y[k] = rnd + sum_m C(m, k) x[m],
// C(m, k) = dct_coef(N, m, k). w holds the inputs as N / 2 words of pairs
// (x[2i], x[2i+1]). Outputs k and N-1-k share the even part (rows 2j of the
// N-point matrix are the N/2-point matrix) and take the odd part with
// opposite signs.
template <int N>
__device__ __forceinline__ void idct_line(const uint32_t* w, int* y, int rnd) {
    if constexpr (N == 2) {
#pragma unroll
        for (int k = 0; k < 2; k++) y[k] = dot2(w[0], pack16c(dct_coef(2, 0, k), dct_coef(2, 1, k)), rnd);
    } else {
        constexpr int H = N / 2, Q = N / 4;
        uint32_t ev[Q > 0 ? Q : 1], od[Q > 0 ? Q : 1];
#pragma unroll
        for (int i = 0; i < Q; i++) {
            ev[i] = __builtin_amdgcn_perm(w[2 * i + 1], w[2 * i], 0x05040100u);  // (x[4i], x[4i+2])
            od[i] = __builtin_amdgcn_perm(w[2 * i + 1], w[2 * i], 0x07060302u);  // (x[4i+1], x[4i+3])
        }
        int e[H];
        idct_line<H>(ev, e, rnd);
#pragma unroll
        for (int k = 0; k < H; k++) {
            int o = 0;
#pragma unroll
            for (int i = 0; i < Q; i++) o = dot2(od[i], pack16c(dct_coef(N, 4 * i + 1, k), dct_coef(N, 4 * i + 3, k)), o);
            y[k] = e[k] + o;
            y[N - 1 - k] = e[k] - o;
        }
    }
}

// NW consecutive 32-bit words at p (16-byte aligned when NW is a multiple of
// 4, else 8-byte aligned; one word: 4-byte aligned), in global memory, or in
// LDS when !G. Stores always go to global memory.
template <typename U, bool G>
__device__ __forceinline__ auto word_ptr(const void* p) {
    if constexpr (G)
        return gp(reinterpret_cast<const U*>(p));
    else
        return reinterpret_cast<const U*>(p);
}
template <int NW, bool G = true>
__device__ __forceinline__ void load_words(const void* p, uint32_t* w) {
    if constexpr (NW == 1) {
        w[0] = *word_ptr<uint32_t, G>(p);
    } else if constexpr (NW % 4 == 0) {
        const auto q = word_ptr<uint4, G>(p);
#pragma unroll
        for (int i = 0; i < NW / 4; i++) {
            const uint4 v = q[i];
            w[4 * i] = v.x;
            w[4 * i + 1] = v.y;
            w[4 * i + 2] = v.z;
            w[4 * i + 3] = v.w;
        }
    } else {
        const auto q = word_ptr<uint2, G>(p);
#pragma unroll
        for (int i = 0; i < NW / 2; i++) {
            const uint2 v = q[i];
            w[2 * i] = v.x;
            w[2 * i + 1] = v.y;
        }
    }
}
template <int NW>
__device__ __forceinline__ void store_words(void* p, const uint32_t* w) {
    if constexpr (NW == 1) {
        *gp(reinterpret_cast<uint32_t*>(p)) = w[0];
    } else if constexpr (NW % 4 == 0) {
        const auto q = gp(reinterpret_cast<uint4*>(p));
#pragma unroll
        for (int i = 0; i < NW / 4; i++) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
    } else {
        const auto q = gp(reinterpret_cast<uint2*>(p));
#pragma unroll
        for (int i = 0; i < NW / 2; i++) q[i] = make_uint2(w[2 * i], w[2 * i + 1]);
    }
}

// A row of N picture samples as N / 2 words of 16-bit pairs.
template <typename T, int N>
__device__ __forceinline__ void load_row(const T* p, uint32_t* w) {
    if constexpr (sizeof(T) == 1) {
        uint32_t raw[N / 4];
        load_words<N / 4>(p, raw);
#pragma unroll
        for (int i = 0; i < N / 4; i++) {
            w[2 * i] = __builtin_amdgcn_perm(0u, raw[i], 0x0c010c00u);
            w[2 * i + 1] = __builtin_amdgcn_perm(0u, raw[i], 0x0c030c02u);
        }
    } else {
        load_words<N / 2>(p, w);
    }
}
template <typename T, int N>
__device__ __forceinline__ void store_row(T* p, const uint32_t* w) {
    if constexpr (sizeof(T) == 1) {
        uint32_t raw[N / 4];
#pragma unroll
        for (int i = 0; i < N / 4; i++) raw[i] = __builtin_amdgcn_perm(w[2 * i + 1], w[2 * i], 0x06040200u);
        store_words<N / 4>(p, raw);
    } else {
        store_words<N / 2>(p, w);
    }
}

template <typename T, int LOG2>
__global__ void __launch_bounds__(256) k_residual_bf(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ wg_prefix,
                                                     int nframes) {
    constexpr int N = 1 << LOG2, BPW = 256 / N;
    constexpr int LD = N == 32 ? 40 : (N == 16 ? 24 : 8);  // LDS row stride: conflict-free 16-byte row reads
    static_assert(BPW == kResPerGroup[LOG2 - 2], "blocks per workgroup");
    __shared__ __align__(16) int16_t tb[BPW * N * LD];
    const int tid = threadIdx.x, bi = tid >> LOG2, j = tid & (N - 1);
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const int fi = frame_of(wg_prefix, nframes, g);
    const FrameDev& d = fr[fi];
    const uint32_t idx = (g - wg_prefix[fi]) * BPW + bi;
    const bool live = idx < d.nblocks[LOG2 - 2];  // the workgroup's first block always is
    const ResBlock rb = *gp(d.blocks[LOG2 - 2] + (live ? idx : (g - wg_prefix[fi]) * BPW));
    int16_t* t = tb + bi * (N * LD);
    const int bd = d.fc.bd;
    int x[N];
    if (d.sig) {  // compact list: the dense block built in LDS (zero, scatter), column j read back
        if (live) {
            uint4* row = reinterpret_cast<uint4*>(t + j * LD);
#pragma unroll
            for (int k = 0; k < N / 8; k++) row[k] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
        if (live) {
            const Glob<const uint32_t>* q = gp(d.sig + rb.coeff_off);
            const uint32_t cnt = q[0];
            for (uint32_t k = (uint32_t)j; k < cnt; k += N) {
                // a position beyond the block wraps into it (only the counts are checked on the host)
                const uint32_t e = q[1 + k], pos = e & 0xffffu;
                t[((pos >> LOG2) & (N - 1)) * LD + (pos & (N - 1))] = (int16_t)(e >> 16);
            }
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int m = 0; m < N; m++) x[m] = t[m * LD + j];
        }
    } else if (live) {
        const Glob<const int16_t>* cp = gp(d.coeffs + rb.coeff_off + j);
#pragma unroll
        for (int m = 0; m < N; m++) x[m] = cp[m * N];
    }
    if (live && (rb.flags & kResRaw)) {  // PCM samples / bypass levels: the block as it is
#pragma unroll
        for (int k = 0; k < N; k++) t[k * LD + j] = (int16_t)x[k];
    } else if (live) {  // column j: dequantise, transform, clip, store to LDS row-major (own column only)
        const int qp = rb.qp, s = qp / 6, b = bd + LOG2 - 5;
        const int ls = kLevelScale[qp - 6 * s];
        const int sl = s > b ? s - b : 0, sr = b > s ? b - s : 0, rnd = sr ? 1 << (sr - 1) : 0;
        // |c * m * levelScale| < 2^30: the product fits before the shifts
        auto dq = [&](int c, int scale) { return clip16(((int)((uint32_t)(c * scale) << sl) + rnd) >> sr); };
        uint32_t w[N / 2];
        if (d.scale_m) {  // scaling lists: factor m(j, row) per coefficient (one frame per workgroup)
            const Glob<const uint8_t>* ms =
                gp(d.scale_m + kScaleOff[LOG2 - 2] + (rb.cidx + ((rb.flags & kResIntra) ? 0 : 3)) * (N * N) + j);
#pragma unroll
            for (int i = 0; i < N / 2; i++)
                w[i] = pack16(dq(x[2 * i], ms[2 * i * N] * ls), dq(x[2 * i + 1], ms[(2 * i + 1) * N] * ls));
        } else {
#pragma unroll
            for (int i = 0; i < N / 2; i++) w[i] = pack16(dq(x[2 * i], 16 * ls), dq(x[2 * i + 1], 16 * ls));
        }
        int y[N];
        idct_line<N>(w, y, 64);
#pragma unroll
        for (int k = 0; k < N; k++) t[k * LD + j] = (int16_t)clip16(y[k] >> 7);
    }
    __syncthreads();
    if (!live) return;
    // row j: transform, clip, then add to the prediction or store the residual
    uint32_t w[N / 2];
    load_words<N / 2, false>(t + j * LD, w);
    if (!(rb.flags & kResRaw)) {
        const int sh2 = 20 - bd;
        int y[N];
        idct_line<N>(w, y, 1 << (sh2 - 1));
#pragma unroll
        for (int i = 0; i < N / 2; i++) w[i] = pack16(clip16(y[2 * i] >> sh2), clip16(y[2 * i + 1] >> sh2));
    }
    const size_t pos = (size_t)(rb.y + j) * d.fc.stride[rb.cidx] + rb.x;
    if (rb.flags & kResPcm) {  // the samples (0..maxv, checked on the host) replace the picture's
        store_row<T, N>(static_cast<T*>(d.rec[rb.cidx]) + pos, w);
    } else if (rb.flags & kResToPlane) {
        T* p = static_cast<T*>(d.rec[rb.cidx]) + pos;
        uint32_t pv[N / 2];
        load_row<T, N>(p, pv);
        const hevcr_short2 zero = {0, 0}, top = {(short)d.fc.maxv, (short)d.fc.maxv};
#pragma unroll
        for (int i = 0; i < N / 2; i++) {
            // saturating: a residual near 32767 plus the prediction clips to maxv, as in int
            hevcr_short2 v = __builtin_elementwise_add_sat(__builtin_bit_cast(hevcr_short2, w[i]),
                                                           __builtin_bit_cast(hevcr_short2, pv[i]));
            v = __builtin_elementwise_min(__builtin_elementwise_max(v, zero), top);
            pv[i] = __builtin_bit_cast(uint32_t, v);
        }
        store_row<T, N>(p, pv);
    } else {
        store_words<N / 2>(d.res[rb.cidx] + pos, w);
    }
}

// Dequantisation + inverse transform of 4x4 blocks (transform.cpp:22-58,
// 240-394, 374-394): DCT, DST-VII (intra luma) or transform skip. A lane per
// column, then per row, 64 blocks per workgroup; each 1-D pass is the 4x4
// matrix product as two v_dot2 per output.
template <typename T>
__global__ void __launch_bounds__(256) k_residual4(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ wg_prefix,
                                                   int nframes) {
    static_assert(kResPerGroup[0] == 64, "blocks per workgroup");
    __shared__ __align__(16) int16_t tb[64 * 16];
    const int tid = threadIdx.x, bi = tid >> 2, j = tid & 3;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const int fi = frame_of(wg_prefix, nframes, g);
    const FrameDev& d = fr[fi];
    const uint32_t first = (g - wg_prefix[fi]) * 64, idx = first + bi;
    const bool live = idx < d.nblocks[0];  // the workgroup's first block always is
    const ResBlock rb = *gp(d.blocks[0] + (live ? idx : first));
    int16_t* t = tb + bi * 16;
    const int bd = d.fc.bd;
    const bool dst = rb.flags & kResDst, tskip = rb.flags & kResTskip;
    uint32_t ma[4], mb[4];  // output k: (M[0][k], M[1][k]) and (M[2][k], M[3][k])
#pragma unroll
    for (int k = 0; k < 4; k++) {
        ma[k] = dst ? pack16c(kDst4.v[0][k], kDst4.v[1][k]) : pack16c(dct_coef(4, 0, k), dct_coef(4, 1, k));
        mb[k] = dst ? pack16c(kDst4.v[2][k], kDst4.v[3][k]) : pack16c(dct_coef(4, 2, k), dct_coef(4, 3, k));
    }
    int x[4];
    if (d.sig) {  // compact list: the dense 4x4 built in LDS (zero, scatter), column j read back
        if (live) *reinterpret_cast<uint2*>(t + 4 * j) = make_uint2(0u, 0u);
        __syncthreads();
        if (live) {
            const Glob<const uint32_t>* q = gp(d.sig + rb.coeff_off);
            const uint32_t cnt = q[0];
            for (uint32_t k = (uint32_t)j; k < cnt; k += 4) {
                const uint32_t e = q[1 + k];
                t[e & 15u] = (int16_t)(e >> 16);
            }
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int m = 0; m < 4; m++) x[m] = t[4 * m + j];
        }
    } else if (live) {
        const Glob<const int16_t>* cp = gp(d.coeffs + rb.coeff_off + j);
#pragma unroll
        for (int m = 0; m < 4; m++) x[m] = cp[4 * m];
    }
    if (live && (rb.flags & kResRaw)) {  // PCM samples / bypass levels: the block as it is
#pragma unroll
        for (int m = 0; m < 4; m++) t[4 * m + j] = (int16_t)x[m];
    } else if (live) {  // column j (own column of the LDS tile only)
        const int qp = rb.qp, s = qp / 6, b = bd - 3;
        const int ls = kLevelScale[qp - 6 * s];
        const int sl = s > b ? s - b : 0, sr = b > s ? b - s : 0, rnd = sr ? 1 << (sr - 1) : 0;
        int scale[4] = {16 * ls, 16 * ls, 16 * ls, 16 * ls};
        if (d.scale_m) {  // scaling lists: factor m(j, row) per coefficient
            const Glob<const uint8_t>* ms = gp(d.scale_m + (rb.cidx + ((rb.flags & kResIntra) ? 0 : 3)) * 16 + j);
#pragma unroll
            for (int m = 0; m < 4; m++) scale[m] = ms[4 * m] * ls;
        }
#pragma unroll
        for (int m = 0; m < 4; m++) x[m] = clip16(((int)((uint32_t)(x[m] * scale[m]) << sl) + rnd) >> sr);
        if (tskip) {  // no transform: the row pass reads the dequantised block
#pragma unroll
            for (int m = 0; m < 4; m++) t[4 * m + j] = (int16_t)x[m];
        } else {
            const uint32_t w0 = pack16(x[0], x[1]), w1 = pack16(x[2], x[3]);
#pragma unroll
            for (int k = 0; k < 4; k++) t[4 * k + j] = (int16_t)clip16(dot2(w1, mb[k], dot2(w0, ma[k], 64)) >> 7);
        }
    }
    __syncthreads();
    if (!live) return;
    // row j
    const uint2 v = *reinterpret_cast<const uint2*>(t + 4 * j);
    int r[4];
    if (rb.flags & kResRaw) {
        r[0] = (int16_t)(v.x & 0xffff);
        r[1] = (int16_t)(v.x >> 16);
        r[2] = (int16_t)(v.y & 0xffff);
        r[3] = (int16_t)(v.y >> 16);
    } else if (tskip) {  // (d + 2^(14-bd)) >> (15-bd), stored as int16 (transform.cpp:380-388)
        const int sh = 15 - bd, rnd = sh > 0 ? 1 << (sh - 1) : 0;
        const int x[4] = {(int16_t)(v.x & 0xffff), (int16_t)(v.x >> 16), (int16_t)(v.y & 0xffff), (int16_t)(v.y >> 16)};
#pragma unroll
        for (int k = 0; k < 4; k++) r[k] = (int16_t)((x[k] + rnd) >> sh);
    } else {
        const int sh2 = 20 - bd;
#pragma unroll
        for (int k = 0; k < 4; k++) r[k] = clip16(dot2(v.y, mb[k], dot2(v.x, ma[k], 1 << (sh2 - 1))) >> sh2);
    }
    uint32_t w[2] = {pack16(r[0], r[1]), pack16(r[2], r[3])};
    const size_t pos = (size_t)(rb.y + j) * d.fc.stride[rb.cidx] + rb.x;
    if (rb.flags & kResPcm) {  // the samples (0..maxv, checked on the host) replace the picture's
        store_row<T, 4>(static_cast<T*>(d.rec[rb.cidx]) + pos, w);
    } else if (rb.flags & kResToPlane) {
        T* p = static_cast<T*>(d.rec[rb.cidx]) + pos;
        uint32_t pv[2];
        load_row<T, 4>(p, pv);
        const hevcr_short2 zero = {0, 0}, top = {(short)d.fc.maxv, (short)d.fc.maxv};
#pragma unroll
        for (int i = 0; i < 2; i++) {
            hevcr_short2 a = __builtin_elementwise_add_sat(__builtin_bit_cast(hevcr_short2, w[i]),
                                                           __builtin_bit_cast(hevcr_short2, pv[i]));
            a = __builtin_elementwise_min(__builtin_elementwise_max(a, zero), top);
            pv[i] = __builtin_bit_cast(uint32_t, a);
        }
        store_row<T, 4>(p, pv);
    } else {
        store_words<2>(d.res[rb.cidx] + pos, w);
    }
}


// ---------------------------------------------------------------------------
// Intra prediction + reconstruction (intra_prediction.cpp:44-481,
// coding_tree.cpp:371-396). One 64-lane workgroup per (CTU, component)
// walks that component's intra blocks of the CTU in decoding order; the CTU,
// its residual and its reconstructed top / left borders live in LDS. CTUs run
// in waves: a CTU comes after every neighbour CTU a block may use (left,
// top-left, top, top-right) that has intra blocks itself (the others are final
// before this stage); in an intra picture the waves are the diagonals x + 2y.
// ---------------------------------------------------------------------------

constexpr int kMaxIntraOps = 256;  // 4x4 blocks of a 64x64 luma CTB

// Orders LDS accesses of the workgroup's single wave without waiting: the LDS
// executes one wave's instructions in issue order, so a read issued after a
// write sees it. Only the compiler must not move accesses across this point.
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// Sum of v over the first K lanes of the wave (the other lanes pass 0): row
// prefix sums by DPP shifts, row totals combined by row broadcasts.
template <int K>
__device__ __forceinline__ int wave_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    if constexpr (K <= 16) return __builtin_amdgcn_readlane(v, 15);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    if constexpr (K <= 32) return __builtin_amdgcn_readlane(v, 31);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

// One intra block of size n = 2^LOG2 at CTU-relative (xt, yt): reference line,
// prediction (intra_prediction.cpp:263-422), residual and the final clip
// (coding_tree.cpp:371-396), written into the CTU array. ext holds the row
// above the CTU (2S + 1 entries, columns cx0 - 1 .. cx0 + 2S - 1), then the
// CTU's S rows of er = S + 1 entries (columns cx0 - 1 .. cx0 + S - 1).
// The reference line lin has 4n + 1 entries: 0 = bottom-left end, 2n = corner,
// 4n = top-right end (top[k] = lin[2n + k], left[k] = lin[2n - k]).
// v * S and v * (S + 1) for the power-of-two CTU size S = 2^sl: shifts and
// adds instead of quarter-rate 32-bit multiplies (two's complement, so also
// for negative v).
__device__ __forceinline__ int mul_pow2(int v, int sl) { return (int)((uint32_t)v << sl); }
__device__ __forceinline__ int mul_pow2p1(int v, int sl) { return (int)(((uint32_t)v << sl) + (uint32_t)v); }

template <int LOG2, typename E>
__device__ __forceinline__ void intra_block(E* ext, const int16_t* rtile, uint16_t* lin, uint16_t* lin2,
                                            int lane, int c, int S, int er, int xt, int yt, int mode, int lo, int hi,
                                            int flags, int ang, int inv, int bd, int maxv) {
    constexpr int n = 1 << LOG2, NS = n * n, PER = NS > 64 ? NS / 64 : 1, tot = 4 * n + 1;
    constexpr int PRE = PER <= 4 ? PER : 1;  // residual samples read ahead (32x32: in the loop, to save registers)
    const int top = 2 * S + 1;
    const int ls2 = 31 - __builtin_clz(S);  // log2 S; rows of ext are er = S + 1 apart
    (void)er;
    E* dst = ext + top + mul_pow2p1(yt, ls2) + xt + 1;
    auto resid = [&](int i) -> int {
        return (flags & kIntraCbf) && (NS >= 64 || i < NS) ? rtile[mul_pow2(yt + (i >> LOG2), ls2) + xt + (i & (n - 1))] : 0;
    };
    int rv[PRE];  // this lane's residual samples, read while the reference line is built
#pragma unroll
    for (int q = 0; q < PRE; q++) rv[q] = resid(lane + 64 * q);
    // entry j: the CTU array sample after substitution (intra_prediction.cpp:177-197): an
    // unavailable entry takes the nearest available one below it (j clamped into [lo, hi),
    // a hole [2n, gend) mapped to 2n - 1), and with none available every entry is
    // 1 << (bd - 1). Left-column entries run up the column x = xt - 1; the others along
    // the row above the block, ext's first row when the block touches the CTU's top.
    {
        const int la = top + mul_pow2p1(yt + 2 * n - 1, ls2) + xt;
        const int lb = (yt == 0 ? 0 : top + mul_pow2p1(yt - 1, ls2)) + xt - 2 * n;
        const bool filt = flags & kIntraFilter;
        const int gend = (flags & kIntraGapTop) ? 3 * n + 1 : ((flags & kIntraGapCorner) ? 2 * n + 1 : 2 * n);
        uint16_t* L = filt ? lin2 : lin;
#pragma unroll
        for (int i = lane; i < tot; i += 64) {
            int v = 1 << (bd - 1);
            if (lo < hi) {
                int jc = i < lo ? lo : (i >= hi ? hi - 1 : i);
                if (jc >= 2 * n && jc < gend) jc = 2 * n - 1;
                v = ext[jc < 2 * n ? la - mul_pow2p1(jc, ls2) : lb + jc];
            }
            L[i] = (uint16_t)v;
        }
        lds_order();
        if (filt) {  // [1,2,1] smoothing or strong bilinear interpolation (intra_prediction.cpp:217-257, 436-471)
            int bil = 0;
            const int tl = lin2[2 * n];
            if (flags & kIntraStrong) {
                const int t = 1 << (bd - 5);
                bil = iabs(tl + lin2[4 * n] - 2 * lin2[3 * n]) < t && iabs(tl + lin2[0] - 2 * lin2[n]) < t;
            }
#pragma unroll
            for (int i = lane; i < tot; i += 64) {
                int v;
                if (i == 0 || i == 4 * n) {
                    v = lin2[i];
                } else if (bil) {
                    if (i == 2 * n)
                        v = tl;
                    else if (i > 2 * n)
                        v = ((2 * n - (i - 2 * n)) * tl + (i - 2 * n) * lin2[4 * n] + n) >> (LOG2 + 1);
                    else
                        v = ((2 * n - (2 * n - i)) * tl + (2 * n - i) * lin2[0] + n) >> (LOG2 + 1);
                } else {
                    v = (lin2[i - 1] + 2 * lin2[i] + lin2[i + 1] + 2) >> 2;
                }
                lin[i] = (uint16_t)v;
            }
            lds_order();
        }
    }
    auto R = [&](int j) -> int { return lin[j]; };
    int dc = 0;
    if (mode == 1) {  // DC: sum of the n top and n left entries
        const int part = lane < 2 * n ? R(lane < n ? 2 * n + 1 + lane : 2 * n - 1 - (lane - n)) : 0;
        dc = (wave_sum<2 * n>(part) + n) >> (LOG2 + 1);
    }
    const bool vert = mode >= 18;
    const bool edge = c == 0 && n < 32;
#pragma unroll PRE
    for (int q = 0; q < PER; q++) {
        const int i = lane + 64 * q;
        if (NS < 64 && i >= NS) break;
        const int y = i >> LOG2, x = i & (n - 1);
        const int r = PER == PRE ? rv[PER == PRE ? q : 0] : resid(i);
        int p;
        if (mode == 0) {
            p = ((n - 1 - x) * R(2 * n - 1 - y) + (x + 1) * R(3 * n + 1) + (n - 1 - y) * R(2 * n + 1 + x) +
                 (y + 1) * R(n - 1) + n) >> (LOG2 + 1);
        } else if (mode == 1) {
            p = dc;
            if (edge) {
                if (x == 0 && y == 0)
                    p = (R(2 * n + 1) + R(2 * n - 1) + 2 * dc + 2) >> 2;
                else if (y == 0)
                    p = (R(2 * n + 1 + x) + 3 * dc + 2) >> 2;
                else if (x == 0)
                    p = (R(2 * n - 1 - y) + 3 * dc + 2) >> 2;
            }
        } else {
            const int pos = vert ? y : x, oth = vert ? x : y;
            const int t = (pos + 1) * ang;
            const int fr5 = t & 31;
            const int k0 = oth + 1 + (t >> 5);
            // main-array entry k, with the negative-index projection of the side
            // array (intra_prediction.cpp:346-368, source index clamped to 2n)
            auto line = [&](int k) -> int {
                int m = k;
                if (ang < 0 && k < 0) {
                    const int si = ((-k) * inv + 128) >> 8;
                    m = -(si > 2 * n ? 2 * n : si);
                }
                return vert ? 2 * n + m : 2 * n - m;
            };
            // (32 - f) * a + f * b with f = 0 is 32 * a: one form for every sample
            p = ((32 - fr5) * R(line(k0)) + fr5 * R(line(k0 + 1)) + 16) >> 5;
            if (edge) {
                if (mode == 26 && x == 0) p = clip3(0, maxv, p + ((R(2 * n - 1 - y) - R(2 * n)) >> 1));
                if (mode == 10 && y == 0) p = clip3(0, maxv, p + ((R(2 * n + 1 + x) - R(2 * n)) >> 1));
            }
        }
        dst[mul_pow2p1(y, ls2) + x] = (E)clip3(0, maxv, (int)(int16_t)p + r);
    }
}

// Up to four independent 4x4 blocks side by side, 16 lanes each (lane group g
// takes block g of the step; groups past `cnt` idle): intra_block<2> with the
// block's parameters per lane (w0..w2: the lane's block record words). Each
// group builds its reference line in its own 17 entries of lin.
template <typename E>
__device__ __forceinline__ void intra_group4(E* ext, const int16_t* rtile, uint16_t* lin, int lane, int c, int S,
                                             int er, int cx0, int cy0, uint32_t w0, uint32_t w1, uint32_t w2, int cnt,
                                             int bd, int maxv) {
    constexpr int n = 4, LOG2 = 2;
    const int g = lane >> 4, sl = lane & 15;
    const bool act = g < cnt;
    const int xt = (int)(w0 & 0xffff) - cx0, yt = (int)(w0 >> 16) - cy0;
    const int mode = (w1 >> 8) & 0xff, lo = (w1 >> 16) & 0xff, hi = w1 >> 24;
    const int flags = w2 & 0xff, ang = (int8_t)((w2 >> 8) & 0xff), inv = (int)(w2 >> 16);
    const int top = 2 * S + 1;
    uint16_t* L = lin + 17 * g;
    const int y = sl >> 2, x = sl & 3;
    const int ls2 = 31 - __builtin_clz(S);  // log2 S; rows of ext are er = S + 1 apart
    (void)er;
    const int rv = act && (flags & kIntraCbf) ? rtile[mul_pow2(yt + y, ls2) + xt + x] : 0;
    {
        const int la = top + mul_pow2p1(yt + 2 * n - 1, ls2) + xt;
        const int lb = (yt == 0 ? 0 : top + mul_pow2p1(yt - 1, ls2)) + xt - 2 * n;
        const int gend = (flags & kIntraGapTop) ? 3 * n + 1 : ((flags & kIntraGapCorner) ? 2 * n + 1 : 2 * n);
#pragma unroll
        for (int i = sl; i < 4 * n + 1; i += 16) {
            int v = 1 << (bd - 1);
            if (lo < hi) {
                int jc = i < lo ? lo : (i >= hi ? hi - 1 : i);
                if (jc >= 2 * n && jc < gend) jc = 2 * n - 1;
                v = ext[jc < 2 * n ? la - mul_pow2p1(jc, ls2) : lb + jc];
            }
            if (act) L[i] = (uint16_t)v;
        }
    }
    lds_order();
    auto R = [&](int j) -> int { return L[j]; };
    // DC sums: lanes 0..7 of each group take the 4 top and 4 left entries; sums
    // over the 16-lane rows by DPP shifts, then from the row's last lane
    int part = sl < 2 * n ? R(sl < n ? 2 * n + 1 + sl : 2 * n - 1 - (sl - n)) : 0;
    part += __builtin_amdgcn_update_dpp(0, part, 0x111, 0xf, 0xf, false);  // row_shr:1
    part += __builtin_amdgcn_update_dpp(0, part, 0x112, 0xf, 0xf, false);  // row_shr:2
    part += __builtin_amdgcn_update_dpp(0, part, 0x114, 0xf, 0xf, false);  // row_shr:4
    part += __builtin_amdgcn_update_dpp(0, part, 0x118, 0xf, 0xf, false);  // row_shr:8
    const int dc = (__shfl(part, (lane & ~15) | 15, 64) + n) >> (LOG2 + 1);
    const bool edge = c == 0;
    int p;
    if (mode == 0) {
        p = ((n - 1 - x) * R(2 * n - 1 - y) + (x + 1) * R(3 * n + 1) + (n - 1 - y) * R(2 * n + 1 + x) +
             (y + 1) * R(n - 1) + n) >> (LOG2 + 1);
    } else if (mode == 1) {
        p = dc;
        if (edge) {
            if (x == 0 && y == 0)
                p = (R(2 * n + 1) + R(2 * n - 1) + 2 * dc + 2) >> 2;
            else if (y == 0)
                p = (R(2 * n + 1 + x) + 3 * dc + 2) >> 2;
            else if (x == 0)
                p = (R(2 * n - 1 - y) + 3 * dc + 2) >> 2;
        }
    } else {
        const bool vert = mode >= 18;
        const int pos = vert ? y : x, oth = vert ? x : y;
        const int t = (pos + 1) * ang;
        const int fr5 = t & 31;
        const int k0 = oth + 1 + (t >> 5);
        auto line = [&](int k) -> int {
            int m = k;
            if (ang < 0 && k < 0) {
                const int si = ((-k) * inv + 128) >> 8;
                m = -(si > 2 * n ? 2 * n : si);
            }
            return vert ? 2 * n + m : 2 * n - m;
        };
        p = ((32 - fr5) * R(line(k0)) + fr5 * R(line(k0 + 1)) + 16) >> 5;
        if (edge) {
            if (mode == 26 && x == 0) p = clip3(0, maxv, p + ((R(2 * n - 1 - y) - R(2 * n)) >> 1));
            if (mode == 10 && y == 0) p = clip3(0, maxv, p + ((R(2 * n + 1 + x) - R(2 * n)) >> 1));
        }
    }
    if (act) ext[top + mul_pow2p1(yt + y, ls2) + xt + x + 1] = (E)clip3(0, maxv, (int)(int16_t)p + rv);
}

// LDS of k_intra, sized for the launch's largest CTB (2^ml luma samples): the
// CTU with its border (ext), its residual (rtile), two reference lines of the
// largest block (32x32, or the CTB; at least four 17-entry lines of 4x4 steps)
// and the words and step spans of up to (2^ml / 4)^2 blocks. Smaller CTBs
// leave more workgroups resident per CU. The CTU is staged in the picture's
// sample type (esz bytes): 64x64 CTBs take 16368 bytes per workgroup at 8
// bits (9 per CU) and 20656 above; 32x32 CTBs 4560 bytes at 8 bits.
struct IntraLds {
    int ext, rtile, lin, lin2, lops, lspan, bytes;  // byte offsets, total
};
constexpr int align16i(int v) { return (v + 15) & ~15; }
__host__ __device__ constexpr IntraLds intra_lds(int ml, int esz) {
    const int sm = 1 << ml, ext_n = 2 * sm + 1 + sm * (sm + 1);
    const int blk = sm < 32 ? sm : 32, lin_n = 4 * blk + 1 > 68 ? 4 * blk + 1 : 68;
    const int ops = (sm / 4) * (sm / 4);
    IntraLds l{};
    l.ext = 0;
    l.rtile = align16i(ext_n * esz);
    l.lin = align16i(l.rtile + sm * sm * 2);
    l.lin2 = align16i(l.lin + lin_n * 2);
    l.lops = align16i(l.lin2 + lin_n * 2);
    l.lspan = l.lops + ops * 12;
    l.bytes = align16i(l.lspan + ops);
    return l;
}
static_assert(intra_lds(6, 2).bytes == 20656 && intra_lds(6, 1).bytes == 16368 && intra_lds(5, 1).bytes == 4560 &&
                  intra_lds(6, 1).lspan - intra_lds(6, 1).lops == kMaxIntraOps * 12,
              "64x64 CTBs: every block of the CTB fits");

// Intra prediction + reconstruction (intra_prediction.cpp:44-481,
// coding_tree.cpp:371-396). One 64-lane workgroup per (CTU, component) walks
// that component's intra blocks of the CTU in steps of dependency levels: a
// step is one block, or up to four independent 4x4 blocks side by side. LDS holds the
// CTU with its reconstructed border as one array `ext`: the row above the CTU
// from x = -1 to 2S - 1 (the top-right reach of the right-most blocks), then
// the CTU's rows from x = -1 (the left CTU's last column) to S - 1. CTUs run in
// dependency waves, so every neighbour CTU a block may use (left, top-left,
// top, top-right) is final when the CTU starts. The workgroup
// is one wave: blocks follow each other with no barrier (lds_order).
// LDS of one intra workgroup (k_intra, k_intra_p).
template <typename E>
struct IntraSmem {
    E* ext;
    uint16_t *lin, *lin2;
    int16_t* rtile;
    uint32_t* lops;
    uint8_t* lspan;
};

#define HEVCR_INTRA_SMEM(T, ML, m)                                                                      \
    IntraSmem<T> m;                                                                                    \
    if constexpr (ML == 6) { /* 64x64 CTBs: arrays of their own (the compiler places them) */          \
        __shared__ T s_ext[2 * 64 + 1 + 64 * 65];                                                      \
        __shared__ int16_t s_rtile[64 * 64];                                                           \
        __shared__ uint16_t s_lin[129];                                                                \
        __shared__ uint16_t s_lin2[129];                                                               \
        __shared__ uint32_t s_lops[kMaxIntraOps * 3];                                                  \
        __shared__ uint8_t s_lspan[kMaxIntraOps];                                                      \
        m = IntraSmem<T>{s_ext, s_lin, s_lin2, s_rtile, s_lops, s_lspan};                             \
    } else { /* smaller CTBs: one launch-sized block (more workgroups per CU) */                        \
        extern __shared__ __align__(16) uint8_t intra_lds_buf[];                                       \
        constexpr IntraLds lay = intra_lds(ML, (int)sizeof(T));                                        \
        m = IntraSmem<T>{reinterpret_cast<T*>(intra_lds_buf + lay.ext),                                \
                         reinterpret_cast<uint16_t*>(intra_lds_buf + lay.lin),                         \
                         reinterpret_cast<uint16_t*>(intra_lds_buf + lay.lin2),                        \
                         reinterpret_cast<int16_t*>(intra_lds_buf + lay.rtile),                        \
                         reinterpret_cast<uint32_t*>(intra_lds_buf + lay.lops), intra_lds_buf + lay.lspan}; \
    }

// Intra prediction + reconstruction of component c of CTU `ctu` of one frame:
// stage the CTU, its residual, its borders and its operations in LDS, run the
// blocks step by step, write the CTU back.
template <typename T, int ML>
__device__ __forceinline__ void intra_ctu(const FrameDev* dp, int c, int ctu, const IntraSmem<T>& m) {
    using E = T;  // the staged CTU in the picture's sample type
    E* ext = m.ext;
    uint16_t *lin = m.lin, *lin2 = m.lin2;
    int16_t* rtile = m.rtile;
    uint32_t* lops = m.lops;
    uint8_t* lspan = m.lspan;
    const int lane = threadIdx.x;
    const int sub = c ? 1 : 0;
    const int ctb_log2 = dp->fc.ctb_log2, wctb = dp->fc.wctb;
    const int S = (1 << ctb_log2) >> sub;
    const int top = 2 * S + 1, er = S + 1;
    const int rx = ctu % wctb, ry = ctu / wctb;
    const int cx0 = rx * S, cy0 = ry * S;
    const int pw = dp->fc.W >> sub, ph = dp->fc.H >> sub;
    const int st = dp->fc.stride[c];
    const int bd = dp->fc.bd, maxv = dp->fc.maxv;
    T* pl = static_cast<T*>(dp->rec[c]);
    const int16_t* rp = dp->res[c];
    const uint32_t o0 = gp(dp->op_begin)[ctu * 3 + c], o1 = gp(dp->op_begin)[ctu * 3 + c + 1];
    const Glob<const uint32_t>* gops = gp(reinterpret_cast<const uint32_t*>(dp->ops));
    const int nops = (int)(o1 - o0);

    // stage the CTU, its residual, its borders and its operations; groups of 4
    // samples, a fixed number per lane. Addresses are clamped into the picture so
    // that every load is valid and all are in flight together; samples outside the
    // picture are never used as references nor written back.
    {
        const int gshift = (ctb_log2 - sub) - 2;  // log2 of 4-sample groups per row
        const int ngroups = (S * S) >> 2;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
            if (64 * k0 >= ngroups) break;
            int v[4][4], rv[4][4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int gi = lane + 64 * (k0 + k);
                const int r = gi >> gshift, gx = (gi & ((1 << gshift) - 1)) << 2;
                const int off = min(cy0 + r, ph - 1) * st + min(cx0 + gx, pw - 4);
                Quad<T>::load(pl + off, v[k]);
                load4_i16(rp + off, rv[k]);  // frames with intra blocks always have residual planes
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int gi = lane + 64 * (k0 + k);
                if (ML < 6 && gi >= ngroups) break;  // smaller CTBs: their LDS holds S rows only
                const int r = gi >> gshift, gx = (gi & ((1 << gshift) - 1)) << 2;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    ext[top + r * er + 1 + gx + q] = (E)v[k][q];
                    rtile[r * S + gx + q] = (int16_t)rv[k][q];
                }
            }
        }
    }
    {
        int tv[3], lv;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int i = lane + 64 * k;
            tv[k] = gp(pl)[max(cy0 - 1, 0) * st + clip3(0, pw - 1, cx0 - 1 + i)];
        }
        lv = gp(pl)[min(cy0 + lane, ph - 1) * st + max(cx0 - 1, 0)];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int i = lane + 64 * k, x = cx0 - 1 + i;
            if (i < top) ext[i] = (cy0 >= 1 && x >= 0 && x < pw) ? (E)tv[k] : (E)0;
        }
        if (lane < S) ext[top + lane * er] = (cx0 >= 1 && cy0 + lane < ph) ? (E)lv : (E)0;
    }
    if (nops > 0) {
        uint32_t ow[12];
#pragma unroll
        for (int k = 0; k < 12; k++) ow[k] = gops[3 * o0 + min(lane + 64 * k, 3 * nops - 1)];
#pragma unroll
        for (int k = 0; k < 12; k++)
            if (lane + 64 * k < 3 * nops) lops[lane + 64 * k] = ow[k];
        uint8_t sw[4];
#pragma unroll
        for (int k = 0; k < 4; k++) sw[k] = gp(dp->op_span)[o0 + min(lane + 64 * k, nops - 1)];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (lane + 64 * k < nops) lspan[lane + 64 * k] = sw[k];
    }
    __syncthreads();

    // the next step's span and block words (lane group g: block g of a 4x4 step) are
    // read while a step runs
    const int lg = lane >> 4;
    const int last = nops > 0 ? nops - 1 : 0;
    int sp = lspan[0];
    uint32_t w0 = lops[3 * min(lg, last)], w1 = lops[3 * min(lg, last) + 1], w2 = lops[3 * min(lg, last) + 2];
    for (int oi = 0; oi < nops;) {
        sp = __builtin_amdgcn_readfirstlane(sp);
        const int cnt = max(sp & 0x7f, 1), nx = oi + cnt;  // every step advances
        const int nk = min(nx + lg, last);
        const int nsp = lspan[min(nx, last)];
        const uint32_t n0 = lops[3 * nk], n1 = lops[3 * nk + 1], n2 = lops[3 * nk + 2];
        if (sp & kIntraGroupStep) {
            intra_group4(ext, rtile, lin, lane, c, S, er, cx0, cy0, w0, w1, w2, cnt, bd, maxv);
        } else {
            const uint32_t u0 = __builtin_amdgcn_readfirstlane(w0), u1 = __builtin_amdgcn_readfirstlane(w1),
                           u2 = __builtin_amdgcn_readfirstlane(w2);
            const int xt = (int)(u0 & 0xffff) - cx0, yt = (int)(u0 >> 16) - cy0;  // CTU-relative position
            const int log2 = u1 & 0xff, mode = (u1 >> 8) & 0xff, lo = (u1 >> 16) & 0xff, hi = u1 >> 24;
            const int flags = u2 & 0xff, ang = (int8_t)((u2 >> 8) & 0xff), inv = (int)(u2 >> 16);
            switch (log2) {
                case 2: intra_block<2>(ext, rtile, lin, lin2, lane, c, S, er, xt, yt, mode, lo, hi, flags, ang, inv, bd, maxv); break;
                case 3: intra_block<3>(ext, rtile, lin, lin2, lane, c, S, er, xt, yt, mode, lo, hi, flags, ang, inv, bd, maxv); break;
                case 4: intra_block<4>(ext, rtile, lin, lin2, lane, c, S, er, xt, yt, mode, lo, hi, flags, ang, inv, bd, maxv); break;
                default: intra_block<5>(ext, rtile, lin, lin2, lane, c, S, er, xt, yt, mode, lo, hi, flags, ang, inv, bd, maxv); break;
            }
        }
        lds_order();
        sp = nsp;
        w0 = n0;
        w1 = n1;
        w2 = n2;
        oi = nx;
    }
    // write the CTU back
    {
        const int gshift = (ctb_log2 - sub) - 2;
        const int ngroups = (S * S) >> 2;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int gi = lane + 64 * k;
            if (gi < ngroups) {
                const int r = gi >> gshift, gx = (gi & ((1 << gshift) - 1)) << 2;
                const int y = cy0 + r, x = cx0 + gx;
                if (x < pw && y < ph) {
                    int v[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) v[q] = ext[top + r * er + 1 + gx + q];
                    Quad<T>::store(pl + y * st + x, v);
                }
            }
        }
    }
}

template <typename T, int ML>
__global__ void __launch_bounds__(64) k_intra(const FrameDev* __restrict__ fr,
                                                  const uint32_t* __restrict__ diag_prefix, int nframes, int diag) {
    HEVCR_INTRA_SMEM(T, ML, m)
    const uint32_t g = blockIdx.x;
    const int fi = frame_of(diag_prefix, nframes, g);
    const FrameDev* dp = &fr[fi];
    const int c = blockIdx.y;
    const int ctu = gp(dp->ictu)[gp(dp->diag_begin)[diag] + (g - diag_prefix[fi])];
    intra_ctu<T, ML>(dp, c, ctu, m);
}

// The intra stage of a whole batch in one launch. Work items are (frame, CTU,
// component) in dependency-wave order (wave w holds items [wave_start[w],
// wave_start[w + 1]); diag_prefix holds one prefix array over the frames per
// wave). Resident workgroups take items from a ticket counter; an item waits
// only for the same component of its left, top-left, top and top-right CTUs,
// which are of earlier waves and so hold smaller tickets, taken by workgroups
// already running: the wait always ends. A finished item makes its samples
// visible device-wide (release) before it raises its flag; a waiting item
// reads its neighbours' samples only after seeing their flags (acquire). A
// wait longer than spin_limit polls gives up, sets err and goes on, so that
// the kernel always exits.
template <typename T, int ML>
__global__ void __launch_bounds__(64) k_intra_p(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ diag_prefix,
                                                    int nframes, const uint32_t* __restrict__ wave_start, int nwaves,
                                                    uint32_t* __restrict__ ticket, uint32_t* __restrict__ err,
                                                    uint32_t spin_limit) {
    HEVCR_INTRA_SMEM(T, ML, m)
    const int lane = threadIdx.x;
    const uint32_t total = wave_start[nwaves];
    // No region of this loop runs on part of the wave only: the compiler may
    // otherwise split the wave at the loop's back edge, and lanes that return
    // to the loop without lane 0 would read a ticket lane 0 never took.
    // Every lane takes part in the ticket atomic (lane 0 adds 1, the others
    // 0), in the flag store and in the error report.
    while (true) {
        const uint32_t t = atomicAdd(ticket, lane == 0 ? 1u : 0u);
        const uint32_t item = __builtin_amdgcn_readfirstlane(t);  // lane 0's ticket
        if (item >= total) break;
        int lo = 0, hi = nwaves - 1;  // its wave: the last w with wave_start[w] <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (wave_start[mid] <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t j = item - wave_start[lo], g = j / 3u;
        const int c = (int)(j - 3u * g);
        const uint32_t* pre = diag_prefix + (size_t)lo * (uint32_t)(nframes + 1);
        const int fi = frame_of(pre, nframes, g);
        const FrameDev* dp = &fr[fi];
        const int ctu = gp(dp->ictu)[gp(dp->diag_begin)[lo] + (g - pre[fi])];
        const int wctb = dp->fc.wctb, rx = ctu % wctb, ry = ctu / wctb;
        uint32_t* flags = dp->iflags;
        // the neighbour this lane watches: lanes 0..3 left, top-left, top,
        // top-right; a lane without one (or outside the picture) watches the
        // item's own flag position, whose value it ignores
        int nb = -1;
        if (lane == 0 && rx > 0) nb = ctu - 1;
        if (lane == 1 && rx > 0 && ry > 0) nb = ctu - wctb - 1;
        if (lane == 2 && ry > 0) nb = ctu - wctb;
        if (lane == 3 && ry > 0 && rx + 1 < wctb) nb = ctu - wctb + 1;
        const uint32_t* watch = flags + 3 * (nb >= 0 ? nb : ctu) + c;
        const uint32_t ctl = gp(dp->ictl)[ctu];
        uint32_t ready = (ctl & kIctlWaits) ? 0u : 1u;  // no intra neighbour above or left: nothing to wait for
        for (uint32_t spins = 0; !ready && spins <= spin_limit; spins++) {
            const uint32_t v = __hip_atomic_load(watch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ready = __builtin_amdgcn_readfirstlane(__all(nb < 0 || v != 0u) ? 1u : 0u);
            if (ready) break;
            __builtin_amdgcn_s_sleep(2);
        }
        if (ctl & kIctlWaits) {
            atomicOr(err, ready ? 0u : 1u);  // the wait gave up: report it (every lane; 0 leaves err unchanged)
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        intra_ctu<T, ML>(dp, c, ctu, m);
        if (ctl & kIctlSignals) {  // an intra neighbour below or right waits for this one
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this wave's stores, device-wide
            __syncthreads();
            __hip_atomic_store(flags + 3 * ctu + c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();  // LDS is restaged by the next item
    }
}

// Before a persistent intra launch: per frame and component, CTUs without
// intra blocks are final already (flag 1), the others not yet (0); the
// ticket counter starts at 0.
__global__ void __launch_bounds__(256) k_intra_init(const FrameDev* __restrict__ fr, uint32_t* __restrict__ ticket) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *ticket = 0u;
    const FrameDev& d = fr[blockIdx.y];
    if (!d.iflags) return;
    const int n = 3 * d.fc.wctb * d.fc.hctb;
    const Glob<const uint32_t>* ob = gp(d.op_begin);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) d.iflags[i] = ob[i + 1] == ob[i] ? 1u : 0u;
}

// ---------------------------------------------------------------------------
// Deblocking (deblocking.cpp:125-614) on tiles of the 64-sample grid, all in
// LDS. Tile (tx, ty) stages luma from (64tx - 4, 64ty - 4) and chroma from
// (32tx - 4, 32ty - 2): NL luma rows and columns, NL / 2 chroma rows. Every
// luma edge x = 64tx + 8m (m < NL / 8) with its 4-line segments, and every
// chroma edge on the 16-sample luma grid, reads and writes inside the tile
// (a luma edge reads x-4..x+3 and writes x-3..x+2, a chroma edge reads
// x-2..x+1 and writes x-1..x). Three passes: a parameter word per segment
// (bS, beta, tc: only side information), the vertical edges, then the
// horizontal edges. A lane filters one line of a segment (luma vertical), two
// lines (luma horizontal: two adjacent columns) or one / two chroma lines;
// the segment's decision takes lines 0 and 3, exchanged between the lanes of
// the segment with DPP quad permutes.
//  - in-place deblocking (NL = 64): tile [64tx-4, 64tx+60) x [64ty-4, 64ty+60);
//    the tiles partition the picture, so they run in place.
//  - deblocking fused with SAO (NL = 72): tile [x0-4, x0+68) for the 64x64
//    area at x0 = 64tx: every deblocked sample of the area and of its
//    one-sample ring, which SAO reads, is computed from the pre-deblock
//    picture (edges x0 and x0 + 64 are filtered by both neighbouring
//    workgroups, identically).
// ---------------------------------------------------------------------------

__device__ __forceinline__ int ref_poc(const FrameConst& fc, const Motion4& m, int l) {
    if (!((m.flags >> l) & 1) || m.ref[l] < 0 || m.ref[l] >= fc.num_ref[l]) return -999999;
    return fc.ref_poc[l][m.ref[l]];
}

__device__ __forceinline__ int mv_far(const int16_t* a, const int16_t* b) {
    return iabs(a[0] - b[0]) >= 4 || iabs(a[1] - b[1]) >= 4;
}

// derive_bs (deblocking.cpp:125-247); ip/iq are the P/Q cells' side
// information words, a/b their motion, qpos the Q-side coordinate across the
// edge.
__device__ int boundary_strength(const FrameConst& fc, uint32_t ip, uint32_t iq, const Motion4& a, const Motion4& b,
                                 int qpos) {
    const int mp = (ip >> info::kModeShift) & 3, mq = (iq >> info::kModeShift) & 3;
    if (mp == HEVCR_MODE_INTRA || mq == HEVCR_MODE_INTRA) return 2;
    const int tp = 1 << (ip & info::kLog2Mask), tq = 1 << (iq & info::kLog2Mask);
    const bool tu_edge = (qpos & (tq - 1)) == 0 || (qpos & (tp - 1)) == 0;  // TU sizes are powers of two
    if (tu_edge && ((ip | iq) & info::kCbf)) return 1;
    const int na = (a.flags & 1) + ((a.flags >> 1) & 1), nb = (b.flags & 1) + ((b.flags >> 1) & 1);
    if (na != nb) return 1;
    if (na == 1) {
        const int la = (a.flags & 1) ? 0 : 1, lb = (b.flags & 1) ? 0 : 1;
        if (ref_poc(fc, a, la) != ref_poc(fc, b, lb)) return 1;
        return mv_far(a.mv[la], b.mv[lb]);
    }
    if (na == 2) {
        const int p0 = ref_poc(fc, a, 0), p1 = ref_poc(fc, a, 1), q0 = ref_poc(fc, b, 0), q1 = ref_poc(fc, b, 1);
        const bool same = p0 == q0 && p1 == q1, swp = p0 == q1 && p1 == q0;
        if (!same && !swp) return 1;
        const int far_same = mv_far(a.mv[0], b.mv[0]) || mv_far(a.mv[1], b.mv[1]);
        const int far_swap = mv_far(a.mv[0], b.mv[1]) || mv_far(a.mv[1], b.mv[0]);
        if (same && !swp) return far_same;
        if (swp && !same) return far_swap;
        return far_same && far_swap;
    }
    return 0;
}

// Parameter word of one edge segment (deblocking.cpp:425-451, 538-560):
// bS (2 bits) | beta' << 2 (7 bits) | tc' << 9 (5 bits) | Cb tc' << 14 |
// Cr tc' << 19 (5 bits each, set on edges of the 16-sample luma grid with
// bS = 2, the only chroma edges) | P side kept << 24 | Q side kept << 25,
// the table values before the bit-depth scaling (beta = beta' << (bd - 8),
// tc likewise). The offsets are the Q side's slice segment's (the slice
// holding q0,0). 0 when the segment is not filtered.
constexpr int kSegTcShift = 9, kSegCbShift = 14, kSegCrShift = 19;
constexpr uint32_t kSegKeepP = 1u << 24, kSegKeepQ = 1u << 25;
__device__ uint32_t segment_word(const FrameConst& fc, uint32_t ip, uint32_t iq, const Motion4& a, const Motion4& b,
                                 int qpos, bool chroma_edge) {
    const int bs = boundary_strength(fc, ip, iq, a, b, qpos);
    if (!bs) return 0;
    const int qpp = (int8_t)(ip >> info::kQpShift), qpq = (int8_t)(iq >> info::kQpShift);
    const int qpl = (qpq + qpp + 1) >> 1;
    const int beta_off = (int)((iq >> info::kBetaShift) & 15) - 6, tc_off = (int)((iq >> info::kTcShift) & 15) - 6;
    const int beta = beta_prime(clip3(0, 51, qpl + 2 * beta_off));
    const int tc = kTcPrime[clip3(0, 53, qpl + 2 * (bs - 1) + 2 * tc_off)];
    uint32_t w = (uint32_t)bs | ((uint32_t)beta << 2) | ((uint32_t)tc << kSegTcShift);
    if (bs == 2 && chroma_edge) {
#pragma unroll
        for (int ci = 1; ci <= 2; ci++) {
            const int off = ci == 1 ? fc.pps_cb_qp_offset : fc.pps_cr_qp_offset;
            const int qpc = qpc_from_qpi(clip3(0, 57, qpl + off));
            const int tcc = kTcPrime[clip3(0, 53, qpc + 2 + 2 * tc_off)];
            w |= (uint32_t)tcc << (ci == 1 ? kSegCbShift : kSegCrShift);
        }
    }
    if (ip & info::kNoFilter) w |= kSegKeepP;
    if (iq & info::kNoFilter) w |= kSegKeepQ;
    return w;
}

// DPP quad permute: lane i of each group of 4 reads lane SEL[i] of its group.
template <int CTRL>
__device__ __forceinline__ int quad_perm(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
constexpr int kQuadAll0 = 0x00, kQuadAll3 = 0xff, kQuadEven = 0xa0, kQuadOdd = 0xf5;  // [0000] [3333] [0022] [1133]

// Decision inputs of one luma line across an edge (P[i] = p_i, Q[i] = q_i),
// packed for the exchange between the lines of a segment: dp | dq << 12 |
// (the line's strong-filter condition, deblocking.cpp:266-273) << 24.
__device__ __forceinline__ int luma_line_key(const int* P, const int* Q, int beta, int tc) {
    const int dp = iabs(P[2] - 2 * P[1] + P[0]), dq = iabs(Q[2] - 2 * Q[1] + Q[0]);
    const bool sc = 2 * (dp + dq) < (beta >> 2) && iabs(P[3] - P[0]) + iabs(Q[0] - Q[3]) < (beta >> 3) &&
                    iabs(P[0] - Q[0]) < ((5 * tc + 1) >> 1);
    return dp | (dq << 12) | ((int)sc << 24);
}

// Segment decision from the keys of its lines 0 and 3 (deblocking.cpp:467-493).
struct LumaDecision {
    bool on, strong, dep, deq;
};
__device__ __forceinline__ LumaDecision luma_decide(int k0, int k3, int beta) {
    const int dp0 = k0 & 4095, dq0 = (k0 >> 12) & 4095, dp3 = k3 & 4095, dq3 = (k3 >> 12) & 4095;
    const int side = (beta + (beta >> 1)) >> 3;
    LumaDecision d;
    d.on = dp0 + dq0 + dp3 + dq3 < beta;
    d.strong = ((k0 & k3) >> 24) & 1;
    d.dep = dp0 + dp3 < side;
    d.deq = dq0 + dq3 < side;
    return d;
}

// Filter of one luma line (deblocking.cpp:279-330); all taps read the
// unfiltered values.
__device__ __forceinline__ void luma_line(int* P, int* Q, const LumaDecision& d, int tc, int maxv) {
    const int P0 = P[0], P1 = P[1], P2 = P[2], P3 = P[3], Q0 = Q[0], Q1 = Q[1], Q2 = Q[2], Q3 = Q[3];
    if (d.strong) {
        P[0] = clip3(P0 - 2 * tc, P0 + 2 * tc, (P2 + 2 * P1 + 2 * P0 + 2 * Q0 + Q1 + 4) >> 3);
        P[1] = clip3(P1 - 2 * tc, P1 + 2 * tc, (P2 + P1 + P0 + Q0 + 2) >> 2);
        P[2] = clip3(P2 - 2 * tc, P2 + 2 * tc, (2 * P3 + 3 * P2 + P1 + P0 + Q0 + 4) >> 3);
        Q[0] = clip3(Q0 - 2 * tc, Q0 + 2 * tc, (P1 + 2 * P0 + 2 * Q0 + 2 * Q1 + Q2 + 4) >> 3);
        Q[1] = clip3(Q1 - 2 * tc, Q1 + 2 * tc, (P0 + Q0 + Q1 + Q2 + 2) >> 2);
        Q[2] = clip3(Q2 - 2 * tc, Q2 + 2 * tc, (P0 + Q0 + Q1 + 3 * Q2 + 2 * Q3 + 4) >> 3);
        return;
    }
    int delta = (9 * (Q0 - P0) - 3 * (Q1 - P1) + 8) >> 4;
    if (iabs(delta) >= tc * 10) return;
    delta = clip3(-tc, tc, delta);
    P[0] = clip3(0, maxv, P0 + delta);
    Q[0] = clip3(0, maxv, Q0 - delta);
    if (d.dep) P[1] = clip3(0, maxv, P1 + clip3(-(tc >> 1), tc >> 1, (((P2 + P0 + 1) >> 1) - P1 + delta) >> 1));
    if (d.deq) Q[1] = clip3(0, maxv, Q1 + clip3(-(tc >> 1), tc >> 1, (((Q2 + Q0 + 1) >> 1) - Q1 - delta) >> 1));
}

// Chroma filter of one line given as words of 16-bit halves: pp = (p1, p0),
// qq = (q0, q1) (deblocking.cpp:338-352).
__device__ __forceinline__ void chroma_line(uint32_t& pp, uint32_t& qq, int tc, int maxv) {
    const int p1 = pp & 0xffff, p0 = pp >> 16, q0 = qq & 0xffff, q1 = qq >> 16;
    const int delta = clip3(-tc, tc, ((q0 - p0) * 4 + p1 - q1 + 4) >> 3);
    pp = (uint32_t)p1 | ((uint32_t)clip3(0, maxv, p0 + delta) << 16);
    qq = (uint32_t)clip3(0, maxv, q0 - delta) | ((uint32_t)q1 << 16);
}

constexpr int kLfL = 72;  // luma LDS row stride: 144-byte rows keep 16-byte alignment
constexpr int kLfC = 40;  // chroma LDS row stride

template <int NL>
struct LfTile {
    static constexpr int NS = NL / 4;         // segments per edge = 4x4 cells per side
    static constexpr int NE = NL / 8;         // luma edges per direction
    static constexpr int NEC = (NE + 1) / 2;  // chroma edges per direction
    static constexpr int NCR = NL / 2;        // chroma rows
    static constexpr int QC = (NL / 2 + 7) / 4;  // chroma 4-sample groups per row
    uint16_t y[NL * kLfL];
    uint16_t c[2][NCR * kLfC];
    uint32_t info[NS * NS];
    Motion4 mot[NS * NS];
    uint32_t sv[NE * NS];  // segment words, vertical edges [m][s]
    uint32_t sh[NE * NS];  // horizontal edges
    // the segments with work, in no particular order: luma vertical, luma
    // horizontal, chroma vertical, chroma horizontal (index m * NS + s)
    uint16_t act[4][NE * NS];
    int nact[4];
    FrameConst fc;
};

// Four consecutive samples as two words of 16-bit halves (the LDS layout).
template <typename T>
struct Wide;
template <>
struct Wide<uint8_t> {
    static __device__ __forceinline__ uint2 load(const uint8_t* p) {
        const uint32_t w = *gp(reinterpret_cast<const uint32_t*>(p));
        return make_uint2(__builtin_amdgcn_perm(0u, w, 0x0c010c00u), __builtin_amdgcn_perm(0u, w, 0x0c030c02u));
    }
    static __device__ __forceinline__ void store(uint8_t* p, uint2 v) {
        *gp(reinterpret_cast<uint32_t*>(p)) = __builtin_amdgcn_perm(v.y, v.x, 0x06040200u);
    }
};
template <>
struct Wide<uint16_t> {
    static __device__ __forceinline__ uint2 load(const uint16_t* p) { return *gp(reinterpret_cast<const uint2*>(p)); }
    static __device__ __forceinline__ void store(uint16_t* p, uint2 v) { *gp(reinterpret_cast<uint2*>(p)) = v; }
};

// Stages the tile's samples from src planes, its 4x4 cells (from
// (16tx - 1, 16ty - 1)) and the frame constants. Every load is issued before
// any LDS store; addresses are clamped into the picture (groups of 4 are
// wholly inside or outside it, since tile origins and plane widths are
// multiples of 4), and what lies outside is never filtered nor written.
template <typename T, int NL>
__device__ __forceinline__ void lf_stage(LfTile<NL>& L, const FrameDev* dp, void* const* src, int tx, int ty,
                                         int tid) {
    using Tile = LfTile<NL>;
    constexpr int QL = NL / 4, NQL = NL * QL, NQC = 2 * Tile::NCR * Tile::QC, NCELL = Tile::NS * Tile::NS;
    constexpr int RL = (NQL + 255) / 256, RC = (NQC + 255) / 256, RS = (NCELL + 255) / 256;
    const int W = dp->fc.W, H = dp->fc.H, cw = W >> 1, ch = H >> 1, gw = dp->fc.gw, gh = dp->fc.gh;
    const int lx0 = 64 * tx - 4, ly0 = 64 * ty - 4, cx0 = 32 * tx - 4, cy0 = 32 * ty - 2;
    const int st0 = dp->fc.stride[0];
    const T* Y = static_cast<const T*>(src[0]);
    uint2 vl[RL], vc[RC];
    uint32_t vi[RS], vm[RS][3];  // a cell's motion as its three words
#pragma unroll
    for (int k = 0; k < RL; k++) {
        const int i = min(tid + 256 * k, NQL - 1), r = i / QL, q = i - r * QL;
        vl[k] = Wide<T>::load(Y + clip3(0, H - 1, ly0 + r) * st0 + clip3(0, W - 4, lx0 + 4 * q));
    }
#pragma unroll
    for (int k = 0; k < RC; k++) {
        const int i = min(tid + 256 * k, NQC - 1), ci = i / (Tile::NCR * Tile::QC);
        const int rem = i - ci * (Tile::NCR * Tile::QC), r = rem / Tile::QC, q = rem - r * Tile::QC;
        const T* C = static_cast<const T*>(src[1 + ci]);
        vc[k] = Wide<T>::load(C + clip3(0, ch - 1, cy0 + r) * dp->fc.stride[1 + ci] + clip3(0, cw - 4, cx0 + 4 * q));
    }
    const bool has_mot = dp->mot != nullptr;
#pragma unroll
    for (int k = 0; k < RS; k++) {
        const int i = min(tid + 256 * k, NCELL - 1), j = i / Tile::NS;
        const int gx = clip3(0, gw - 1, 16 * tx - 1 + (i - j * Tile::NS)), gy = clip3(0, gh - 1, 16 * ty - 1 + j);
        vi[k] = gp(dp->info4)[(size_t)gy * gw + gx];
        if (has_mot) {
            const Glob<const uint32_t>* m = gp(reinterpret_cast<const uint32_t*>(dp->mot + (size_t)gy * gw + gx));
            vm[k][0] = m[0];
            vm[k][1] = m[1];
            vm[k][2] = m[2];
        }
    }
    if (tid < (int)(sizeof(FrameConst) / 4))
        reinterpret_cast<int*>(&L.fc)[tid] = reinterpret_cast<const int*>(&dp->fc)[tid];
    if (tid < 4) L.nact[tid] = 0;
#pragma unroll
    for (int k = 0; k < RL; k++) {
        const int i = tid + 256 * k, r = i / QL, q = i - r * QL;
        if (i < NQL) *reinterpret_cast<uint2*>(&L.y[r * kLfL + 4 * q]) = vl[k];
    }
#pragma unroll
    for (int k = 0; k < RC; k++) {
        const int i = tid + 256 * k, ci = i / (Tile::NCR * Tile::QC);
        const int rem = i - ci * (Tile::NCR * Tile::QC), r = rem / Tile::QC, q = rem - r * Tile::QC;
        if (i < NQC) *reinterpret_cast<uint2*>(&L.c[ci][r * kLfC + 4 * q]) = vc[k];
    }
#pragma unroll
    for (int k = 0; k < RS; k++) {
        const int i = tid + 256 * k;
        if (i < NCELL) {
            L.info[i] = vi[k];
            if (has_mot) {
                uint32_t* m = reinterpret_cast<uint32_t*>(&L.mot[i]);
                m[0] = vm[k][0];
                m[1] = vm[k][1];
                m[2] = vm[k][2];
            }
        }
    }
}

// Segment parameter words of the tile's edges. Vertical edge m, segment s:
// x = 64tx + 8m, rows from 64ty - 4 + 4s; horizontal: y = 64ty + 8m,
// columns from 64tx - 4 + 4s.
template <int NL>
__device__ __forceinline__ void lf_params(LfTile<NL>& L, int tx, int ty, int tid) {
    using Tile = LfTile<NL>;
    constexpr int NS = Tile::NS, NE = Tile::NE;
    const FrameConst& fc = L.fc;
    for (int t = tid; t < 2 * NE * NS; t += 256) {
        const bool vert = t < NE * NS;
        const int u = vert ? t : t - NE * NS, m = u / NS, s = u - m * NS;
        const int x = vert ? 64 * tx + 8 * m : 64 * tx - 4 + 4 * s;
        const int y = vert ? 64 * ty - 4 + 4 * s : 64 * ty + 8 * m;
        uint32_t w = 0;
        if (fc.deblock && x >= 0 && x < fc.W && y >= 0 && y < fc.H && (vert ? x : y) > 0) {
            const int qi = vert ? s * NS + 2 * m + 1 : (2 * m + 1) * NS + s;
            const int pi = vert ? qi - 1 : qi - NS;
            const uint32_t iq = L.info[qi];
            if (iq & (vert ? info::kEdgeV : info::kEdgeH))
                w = segment_word(fc, L.info[pi], iq, L.mot[pi], L.mot[qi], vert ? x : y, (m & 1) == 0);
        }
        (vert ? L.sv : L.sh)[u] = w;
        if (w & 3) L.act[vert ? 0 : 1][atomicAdd(&L.nact[vert ? 0 : 1], 1)] = (uint16_t)u;
        if ((w & 3) == 2 && !(m & 1)) L.act[vert ? 2 : 3][atomicAdd(&L.nact[vert ? 2 : 3], 1)] = (uint16_t)u;
    }
}

// Vertical edges: a lane per luma line (one 16-byte LDS read and write: p3..q3),
// then a lane per chroma line.
template <int NL>
__device__ __forceinline__ void lf_vertical(LfTile<NL>& L, int tid) {
    using Tile = LfTile<NL>;
    constexpr int NS = Tile::NS;
    const int TL = 4 * L.nact[0], TC = 4 * L.nact[2];  // 4 lines per luma segment, 2 rows x 2 components per chroma
    const int maxv = L.fc.maxv, bsh = L.fc.bd - 8;
    for (int t = tid; t < TL + TC; t += 256) {
        if (t < TL) {  // 4 consecutive lanes = the 4 lines of one segment
            const int seg = L.act[0][t >> 2], m = seg / NS, r = 4 * (seg - m * NS) + (t & 3);
            const uint32_t w = L.sv[seg];
            const int beta = ((w >> 2) & 127) << bsh, tc = ((w >> kSegTcShift) & 31) << bsh;
            uint4* p = reinterpret_cast<uint4*>(&L.y[r * kLfL + 8 * m]);
            uint4 v = *p;
            int P[4] = {(int)(v.y >> 16), (int)(v.y & 0xffff), (int)(v.x >> 16), (int)(v.x & 0xffff)};
            int Q[4] = {(int)(v.z & 0xffff), (int)(v.z >> 16), (int)(v.w & 0xffff), (int)(v.w >> 16)};
            const int key = luma_line_key(P, Q, beta, tc);
            const LumaDecision d = luma_decide(quad_perm<kQuadAll0>(key), quad_perm<kQuadAll3>(key), beta);
            if (!d.on) continue;
            const uint4 v0 = v;
            luma_line(P, Q, d, tc, maxv);
            v.x = (uint32_t)P[3] | ((uint32_t)P[2] << 16);
            v.y = (uint32_t)P[1] | ((uint32_t)P[0] << 16);
            v.z = (uint32_t)Q[0] | ((uint32_t)Q[1] << 16);
            v.w = (uint32_t)Q[2] | ((uint32_t)Q[3] << 16);
            if (w & kSegKeepP) {  // nDp = 0 (deblocking.cpp:330-332)
                v.x = v0.x;
                v.y = v0.y;
            }
            if (w & kSegKeepQ) {
                v.z = v0.z;
                v.w = v0.w;
            }
            *p = v;
        } else {
            const int u = t - TL, seg = L.act[2][u >> 2], m = seg / NS, s = seg - m * NS;
            const int ci = (u >> 1) & 1, r = 2 * s + (u & 1), m2 = m >> 1;
            const uint32_t w = L.sv[seg];
            uint32_t* p = reinterpret_cast<uint32_t*>(&L.c[ci][r * kLfC + 2 + 8 * m2]);
            const uint32_t p0 = p[0], q0 = p[1];
            uint32_t pp = p0, qq = q0;
            chroma_line(pp, qq, ((w >> (ci ? kSegCrShift : kSegCbShift)) & 31) << bsh, maxv);
            p[0] = (w & kSegKeepP) ? p0 : pp;
            p[1] = (w & kSegKeepQ) ? q0 : qq;
        }
    }
}

// Horizontal edges: a lane per two adjacent luma columns (8 rows of 4-byte LDS
// words; 2 lanes per segment), then a lane per two chroma columns.
template <int NL>
__device__ __forceinline__ void lf_horizontal(LfTile<NL>& L, int tid) {
    using Tile = LfTile<NL>;
    constexpr int NS = Tile::NS, RW = kLfL / 2, RWC = kLfC / 2;
    const int TL = 2 * L.nact[1], TC = 2 * L.nact[3];  // 2 column pairs per luma segment, 2 components per chroma
    const int maxv = L.fc.maxv, bsh = L.fc.bd - 8;
    for (int t = tid; t < TL + TC; t += 256) {
        if (t < TL) {  // 2 consecutive lanes = the 4 columns of one segment
            const int seg = L.act[1][t >> 1], m = seg / NS, cp = 2 * (seg - m * NS) + (t & 1);
            const uint32_t w = L.sh[seg];
            const int beta = ((w >> 2) & 127) << bsh, tc = ((w >> kSegTcShift) & 31) << bsh;
            uint32_t* p = reinterpret_cast<uint32_t*>(&L.y[(8 * m) * kLfL + 2 * cp]);
            uint32_t rw[8];
#pragma unroll
            for (int i = 0; i < 8; i++) rw[i] = p[i * RW];  // rows p3, p2, p1, p0, q0, q1, q2, q3
            int PA[4], QA[4], PB[4], QB[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                PA[i] = rw[3 - i] & 0xffff;
                PB[i] = rw[3 - i] >> 16;
                QA[i] = rw[4 + i] & 0xffff;
                QB[i] = rw[4 + i] >> 16;
            }
            // segment line 0 = the even lane's first column, line 3 = the odd lane's second
            const int k0 = quad_perm<kQuadEven>(luma_line_key(PA, QA, beta, tc));
            const int k3 = quad_perm<kQuadOdd>(luma_line_key(PB, QB, beta, tc));
            const LumaDecision d = luma_decide(k0, k3, beta);
            if (!d.on) continue;
            luma_line(PA, QA, d, tc, maxv);
            luma_line(PB, QB, d, tc, maxv);
#pragma unroll
            for (int i = 0; i < 3; i++) {  // a kept side (nDp / nDq = 0) is written back unchanged
                if (!(w & kSegKeepP)) p[(2 - i + 1) * RW] = (uint32_t)PA[i] | ((uint32_t)PB[i] << 16);
                if (!(w & kSegKeepQ)) p[(4 + i) * RW] = (uint32_t)QA[i] | ((uint32_t)QB[i] << 16);
            }
        } else {
            const int u = t - TL, seg = L.act[3][u >> 1], m = seg / NS, s = seg - m * NS;
            const int ci = u & 1, m2 = m >> 1;
            const uint32_t w = L.sh[seg];
            const int tc = ((w >> (ci ? kSegCrShift : kSegCbShift)) & 31) << bsh;
            uint32_t* p = reinterpret_cast<uint32_t*>(&L.c[ci][(8 * m2) * kLfC + 2 + 2 * s]);
            const uint32_t r0 = p[0], r1 = p[RWC], r2 = p[2 * RWC], r3 = p[3 * RWC];  // p1, p0, q0, q1
            uint32_t pa = (r0 & 0xffff) | (r1 << 16), qa = (r2 & 0xffff) | (r3 << 16);
            uint32_t pb = (r0 >> 16) | (r1 & 0xffff0000u), qb = (r2 >> 16) | (r3 & 0xffff0000u);
            chroma_line(pa, qa, tc, maxv);
            chroma_line(pb, qb, tc, maxv);
            if (!(w & kSegKeepP)) p[RWC] = (pa >> 16) | (pb & 0xffff0000u);
            if (!(w & kSegKeepQ)) p[2 * RWC] = (qa & 0xffff) | (qb << 16);
        }
    }
}

// In-place deblocking, one 256-lane workgroup per 64x64 tile (NL = 64).
template <typename T>
__global__ void __launch_bounds__(256) k_deblock(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ tile_prefix,
                                                 int nframes) {
    __shared__ __align__(16) LfTile<64> L;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const int f = frame_of(tile_prefix, nframes, g);
    const FrameDev* dp = &fr[f];
    const int tid = threadIdx.x;
    const int W = dp->fc.W, H = dp->fc.H, cw = W >> 1, ch = H >> 1;
    const int t = (int)(g - tile_prefix[f]);
    const int ntx = (W + 4 + 63) >> 6;
    const int tx = t % ntx, ty = t / ntx;
    lf_stage<T, 64>(L, dp, dp->rec, tx, ty, tid);
    __syncthreads();
    lf_params<64>(L, tx, ty, tid);
    __syncthreads();
    lf_vertical<64>(L, tid);
    __syncthreads();
    lf_horizontal<64>(L, tid);
    __syncthreads();
    const int st0 = dp->fc.stride[0];
    T* Yw = static_cast<T*>(dp->rec[0]);
#pragma unroll
    for (int k = 0; k < 4; k++) {  // luma [64tx-4, 64tx+60) x [64ty-4, 64ty+60)
        const int i = tid + 256 * k, r = i >> 4, q = (i & 15) << 2, y = 64 * ty - 4 + r, x = 64 * tx - 4 + q;
        if (x >= 0 && x < W && y >= 0 && y < H)
            Wide<T>::store(Yw + y * st0 + x, *reinterpret_cast<const uint2*>(&L.y[r * kLfL + q]));
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {  // chroma [32tx-2, 32tx+30) x [32ty-2, 32ty+30), in pairs
        const int i = tid + 256 * k, ci = i >> 9, r = (i >> 4) & 31, q = (i & 15) << 1;
        const int y = 32 * ty - 2 + r, x = 32 * tx - 2 + q;
        if (x >= 0 && x < cw && y >= 0 && y < ch)
            Pair<T>::store(static_cast<T*>(dp->rec[1 + ci]) + y * dp->fc.stride[1 + ci] + x,
                           *reinterpret_cast<const uint32_t*>(&L.c[ci][r * kLfC + 2 + q]));
    }
}

// ---------------------------------------------------------------------------
// SAO (sao.cpp:20-243): deblocked picture -> output picture, one 256-lane
// workgroup per CTB (all three components). Band offsets are not wrapped
// modulo 32 (sao.cpp:226-230). A lane handles groups of 4 consecutive
// samples: the CTB with a 1-row / 4-column halo is staged in LDS as 16-bit
// samples, and each group reads its 3 x 12 neighbourhood with 64-bit LDS
// reads. The five offsets of a component are packed as 6-bit fields of one
// word (|offset| <= 31), so a lookup is one bit-field extract.
// ---------------------------------------------------------------------------

// EO on one group of 4: rows[r][j] holds the samples at columns x - 4 + j of
// rows y - 1 + r (x = the group's first column). Edge class offsets
// (sao.cpp:17-18): a = (x + DX, y + DY), b = (x - DX, y - DY).
template <int DX, int DY>
__device__ __forceinline__ void sao_eo4(const int (*rows)[12], uint32_t ofs, int maxv, bool check, int x, int y, int pw,
                                        int ph, int* out) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int v = rows[1][4 + k], av = rows[1 + DY][4 + k + DX], bv = rows[1 - DY][4 + k - DX];
        const int idx = 2 + (v > av) - (v < av) + (v > bv) - (v < bv);
        const int off = __builtin_amdgcn_sbfe(ofs, 6 * idx, 6);
        int o = clip3(0, maxv, v + off);
        if (check) {  // both neighbours inside the picture (sao.cpp:139-140)
            const int xa = x + k + DX, xb = x + k - DX;
            if (xa < 0 || xa >= pw || xb < 0 || xb >= pw || y + DY < 0 || y + DY >= ph || y - DY < 0 || y - DY >= ph) o = v;
        }
        out[k] = o;
    }
}

// Whether an edge-offset neighbour at component position (nx, ny) lies in a
// neighbouring CTB whose samples this CTB may not read (ctl::kSaoMask bits of
// its control word cw; x0, y0, S: the CTB's component origin and size).
__device__ __forceinline__ bool sao_nb_excluded(uint32_t cw, int nx, int ny, int x0, int y0, int S) {
    const int dxc = nx < x0 ? 0 : (nx >= x0 + S ? 2 : 1), dyc = ny < y0 ? 0 : (ny >= y0 + S ? 2 : 1);
    // (dxc, dyc) -> neighbour bit: TL T TR / L - R / BL B BR, 4 bits each (15: the CTB itself)
    constexpr uint64_t kBit = (uint64_t)ctl::kNbTL | (uint64_t)ctl::kNbT << 4 | (uint64_t)ctl::kNbTR << 8 |
                              (uint64_t)ctl::kNbL << 12 | (uint64_t)15 << 16 | (uint64_t)ctl::kNbR << 20 |
                              (uint64_t)ctl::kNbBL << 24 | (uint64_t)ctl::kNbB << 28 | (uint64_t)ctl::kNbBR << 32;
    const int bit = (int)((kBit >> (4 * (3 * dyc + dxc))) & 15);
    return bit != 15 && ((cw >> bit) & 1);
}

template <typename T>
__global__ void __launch_bounds__(256) k_sao(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ ctb_prefix,
                                             int nframes) {
    constexpr int kRS = 64 + 8;  // LDS row: columns x0 - 4 .. x0 + S + 3
    __shared__ __align__(16) uint16_t reg[66 * kRS];
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const int f = frame_of(ctb_prefix, nframes, g);
    const FrameDev* dp = &fr[f];
    const int tid = threadIdx.x;
    const int W = dp->fc.W, H = dp->fc.H, bd = dp->fc.bd, maxv = dp->fc.maxv;
    const int ctb_log2 = dp->fc.ctb_log2, wctb = dp->fc.wctb;
    const int ctb = (int)(g - ctb_prefix[f]);
    const int rx = ctb % wctb, ry = ctb / wctb;
    // slice / tile boundaries and kept (PCM, bypass) samples of this CTB
    const uint32_t cw = dp->ctb_ctl ? gp(dp->ctb_ctl)[ctb] : 0u;
    const bool special = (cw & (ctl::kSaoMask | ctl::kNoFilter)) != 0;
    __shared__ hevcr_sao sp;  // indexed by component below: one copy in LDS, not per lane
    if (tid == 0) sp = gp(dp->sao)[ctb];
    __syncthreads();
    for (int c = 0; c < 3; c++) {
        const int sub = c ? 1 : 0;
        const int slog = ctb_log2 - sub, S = 1 << slog;
        const int pw = W >> sub, ph = H >> sub, st = dp->fc.stride[c];
        const int x0 = rx * S, y0 = ry * S;
        const T* src = static_cast<const T*>(dp->rec[c]);
        T* dst = static_cast<T*>(dp->out[c]);
        const int type = sp.type[c];
        const int gshift = slog - 2;  // log2 of 4-sample groups per row
        if (type == 0) {  // copy
            for (int gi = tid; gi < (S * S) >> 2; gi += 256) {
                const int y = y0 + (gi >> gshift), x = x0 + ((gi & ((1 << gshift) - 1)) << 2);
                if (x < pw && y < ph) {
                    int v[4];
                    Quad<T>::load(src + y * st + x, v);
                    Quad<T>::store(dst + y * st + x, v);
                }
            }
            continue;
        }
        const int e = sp.eo_class[c] & 3;
        const bool need_rows = type == 2 && e != 0;  // vertical and diagonal classes read rows y - 1, y + 1
        __syncthreads();  // the previous component is done with reg
        {   // stage rows y0 - 1 .. y0 + S, columns x0 - 4 .. x0 + S + 3, in groups of 4
            const int qc = (S >> 2) + 2, nq = (S + 2) * qc;
            int r = tid / qc, q = tid - r * qc;
            const int dr = 256 / qc, dq = 256 - dr * qc;
            for (int i = tid; i < nq; i += 256) {
                const int y = y0 - 1 + r, x = x0 - 4 + 4 * q;
                if ((need_rows || (r >= 1 && r <= S)) && y >= 0 && y < ph && x >= 0 && x < pw) {
                    int v[4];
                    Quad<T>::load(src + y * st + x, v);
                    uint2 w;
                    w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
                    w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
                    *reinterpret_cast<uint2*>(&reg[r * kRS + 4 * q]) = w;
                }
                q += dq;
                r += dr;
                if (q >= qc) {
                    q -= qc;
                    r += 1;
                }
            }
        }
        __syncthreads();
        uint32_t ofs = 0;
#pragma unroll
        for (int i = 0; i < 5; i++) ofs |= ((uint32_t)sp.offset[c][i] & 63u) << (6 * i);
        const int shift = bd - 5, pos = sp.band_pos[c];
        const bool border = x0 == 0 || y0 == 0 || x0 + S >= pw || y0 + S >= ph;
        for (int gi = tid; gi < (S * S) >> 2; gi += 256) {
            const int ly = gi >> gshift, lx = (gi & ((1 << gshift) - 1)) << 2;
            const int x = x0 + lx, y = y0 + ly;
            if (x >= pw || y >= ph) continue;
            int rows[3][12];
#pragma unroll
            for (int rr = 0; rr < 3; rr++) {
                if (rr != 1 && !need_rows) continue;
                const uint2* p = reinterpret_cast<const uint2*>(&reg[(ly + rr) * kRS + lx]);
#pragma unroll
                for (int h = 0; h < 3; h++) {
                    const uint2 w = p[h];
                    rows[rr][4 * h + 0] = w.x & 0xffff;
                    rows[rr][4 * h + 1] = w.x >> 16;
                    rows[rr][4 * h + 2] = w.y & 0xffff;
                    rows[rr][4 * h + 3] = w.y >> 16;
                }
            }
            int out[4];
            if (type == 2) {
                switch (e) {
                    case 0: sao_eo4<-1, 0>(rows, ofs, maxv, border, x, y, pw, ph, out); break;
                    case 1: sao_eo4<0, -1>(rows, ofs, maxv, border, x, y, pw, ph, out); break;
                    case 2: sao_eo4<-1, -1>(rows, ofs, maxv, border, x, y, pw, ph, out); break;
                    default: sao_eo4<1, -1>(rows, ofs, maxv, border, x, y, pw, ph, out); break;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int v = rows[1][4 + k];
                    const int bi = (v >> shift) - pos;
                    out[k] = (uint32_t)bi < 4u ? clip3(0, maxv, v + __builtin_amdgcn_sbfe(ofs, 6 * bi, 6)) : v;
                }
            }
            if (special) {  // sao.cpp:122-128, 142-185, 221-225: such samples keep their value
                const int edx = e == 3 ? 1 : (e == 1 ? 0 : -1), edy = e == 0 ? 0 : -1;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    bool keep = false;
                    if (cw & ctl::kNoFilter)
                        keep = gp(dp->info4)[(size_t)((y << sub) >> 2) * dp->fc.gw + (((x + k) << sub) >> 2)] &
                               info::kNoFilter;
                    if (type == 2 && (cw & ctl::kSaoMask))
                        keep = keep || sao_nb_excluded(cw, x + k + edx, y + edy, x0, y0, S) ||
                               sao_nb_excluded(cw, x + k - edx, y - edy, x0, y0, S);
                    if (keep) out[k] = rows[1][4 + k];
                }
            }
            Quad<T>::store(dst + y * st + x, out);
        }
    }
}

// ---------------------------------------------------------------------------
// Deblocking fused with SAO for frames with 64x64 or 32x32 CTBs: one
// 256-lane workgroup per 64x64 area (one CTB, or 2x2 CTBs with their own SAO
// parameters) reads the pre-deblock picture once, deblocks its 72x72 tile
// (the area and 4 samples around it) in LDS and applies SAO to the area from
// the deblocked samples, its one-sample ring included, writing the output
// picture only. The same result as k_deblock then k_sao: SAO reads
// deblocked, pre-SAO samples (sao.cpp:20-243), and every sample of the ring
// is deblocked here from the same inputs as by the workgroup that owns it.
// ---------------------------------------------------------------------------

// EO on one group of 4 samples of an interior CTB (every neighbour inside the
// picture) in packed 16-bit arithmetic: row points at the LDS row of the group
// from column x - 4. A neighbour row shifted by one sample is formed from the
// row's words with alignbit. Returns the 4 outputs as two 16-bit pair words.
template <int DX, int DY>
__device__ __forceinline__ uint2 sao_eo4_packed(const uint16_t* row, int ld, uint32_t ofs, int maxv) {
    auto words = [&](int dy, int dx, uint32_t& w0, uint32_t& w1) {
        const uint16_t* r = row + dy * ld;
        const uint2 m = *reinterpret_cast<const uint2*>(r + 4);  // columns x .. x + 3
        if (dx == 0) {
            w0 = m.x;
            w1 = m.y;
        } else if (dx < 0) {  // x - 1 .. x + 2
            w0 = __builtin_amdgcn_alignbit(m.x, *reinterpret_cast<const uint32_t*>(r + 2), 16);
            w1 = __builtin_amdgcn_alignbit(m.y, m.x, 16);
        } else {  // x + 1 .. x + 4
            w0 = __builtin_amdgcn_alignbit(m.y, m.x, 16);
            w1 = __builtin_amdgcn_alignbit(*reinterpret_cast<const uint32_t*>(r + 8), m.y, 16);
        }
    };
    uint32_t v[2], a[2], b[2];
    words(0, 0, v[0], v[1]);
    words(DY, DX, a[0], a[1]);
    words(-DY, -DX, b[0], b[1]);
    const hevcr_short2 one = {1, 1}, mone = {-1, -1}, two = {2, 2}, zero = {0, 0}, top = {(short)maxv, (short)maxv};
    uint2 out;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const hevcr_short2 vv = __builtin_bit_cast(hevcr_short2, v[k]);
        const hevcr_short2 sa = __builtin_elementwise_max(
            __builtin_elementwise_min(vv - __builtin_bit_cast(hevcr_short2, a[k]), one), mone);
        const hevcr_short2 sb = __builtin_elementwise_max(
            __builtin_elementwise_min(vv - __builtin_bit_cast(hevcr_short2, b[k]), one), mone);
        const hevcr_short2 idx = sa + sb + two;  // edgeIdx 0..4 (sao.cpp:144-151)
        const int o0 = __builtin_amdgcn_sbfe(ofs, 6 * (int)idx.x, 6), o1 = __builtin_amdgcn_sbfe(ofs, 6 * (int)idx.y, 6);
        hevcr_short2 r = vv + __builtin_bit_cast(hevcr_short2, pack16(o0, o1));
        r = __builtin_elementwise_min(__builtin_elementwise_max(r, zero), top);
        (k ? out.y : out.x) = __builtin_bit_cast(uint32_t, r);
    }
    return out;
}

// BO on one group of 4 samples (sao.cpp:220-235) given as two 16-bit pair
// words: band v >> (bd - 5), offsets for bands pos .. pos + 3 (no wrap).
__device__ __forceinline__ uint2 sao_bo4_packed(uint2 w, uint32_t ofs, int shift, int pos, int maxv) {
    const hevcr_short2 sh = {(short)shift, (short)shift}, p = {(short)pos, (short)pos};
    const hevcr_short2 zero = {0, 0}, top = {(short)maxv, (short)maxv};
    uint2 out;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const hevcr_short2 vv = __builtin_bit_cast(hevcr_short2, k ? w.y : w.x);
        const hevcr_short2 bi = (vv >> sh) - p;
        const int b0 = bi.x, b1 = bi.y;
        const int o0 = (uint32_t)b0 < 4u ? __builtin_amdgcn_sbfe(ofs, 6 * b0, 6) : 0;
        const int o1 = (uint32_t)b1 < 4u ? __builtin_amdgcn_sbfe(ofs, 6 * b1, 6) : 0;
        hevcr_short2 r = vv + __builtin_bit_cast(hevcr_short2, pack16(o0, o1));
        r = __builtin_elementwise_min(__builtin_elementwise_max(r, zero), top);
        (k ? out.y : out.x) = __builtin_bit_cast(uint32_t, r);
    }
    return out;
}

// SAO over one 64x64 area of deblocked samples in LDS: PER x PER CTBs, each
// with its own parameters in sp; groups of 4 samples are taken CTB by CTB so
// that a wave (64 groups) stays inside one CTB where a CTB has that many.
template <typename T, int PER>
__device__ __forceinline__ void lf_sao(const LfTile<72>& L, const FrameDev* dp, const hevcr_sao* sp, int tx, int ty,
                                       int tid) {
    constexpr int cl2 = PER == 1 ? 6 : 5;
    constexpr int per = PER;
    const int W = L.fc.W, H = L.fc.H, bd = L.fc.bd, maxv = L.fc.maxv;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int sub = c ? 1 : 0, T64 = 64 >> sub;
        const int sl2 = cl2 - sub, S = 1 << sl2;      // CTB size in this component
        const int gshift = sl2 - 2, gps = 2 * sl2 - 2;  // log2 of 4-sample groups per CTB row / per CTB
        const int pw = W >> sub, ph = H >> sub, st = L.fc.stride[c];
        // the tile's sample (0, 0) sits at LDS column 4, row 4 (luma) or 2 (chroma)
        const int ld = c ? kLfC : kLfL;
        const uint16_t* base = c ? &L.c[c - 1][2 * kLfC] : &L.y[4 * kLfL];
        T* dst = static_cast<T*>(dp->out[c]);
        // a CTB's parameters: type, EO class, packed offsets, band position, whether it touches the picture edge
        struct Prm {
            int type, e, pos;
            uint32_t ofs;
            bool border;
        };
        auto prm_of = [&](int q) {
            const hevcr_sao& r = sp[q];
            const int cx0 = tx * T64 + (q % per) * S, cy0 = ty * T64 + (q / per) * S;
            Prm m;
            m.type = r.type[c];
            m.e = r.eo_class[c] & 3;
            m.pos = r.band_pos[c];
            m.ofs = 0;
#pragma unroll
            for (int i = 0; i < 5; i++) m.ofs |= ((uint32_t)r.offset[c][i] & 63u) << (6 * i);
            m.border = cx0 == 0 || cy0 == 0 || cx0 + S >= pw || cy0 + S >= ph;
            return m;
        };
        const Prm one = prm_of(0);  // the only CTB when PER == 1
        // groups CTB by CTB, so that a wave (64 groups) stays inside one CTB
        for (int gi = tid; gi < (T64 * T64) >> 2; gi += 256) {
            const int q = PER == 1 ? 0 : gi >> gps, local = PER == 1 ? gi : gi & ((1 << gps) - 1);
            const int qx = q % per, qy = q / per;
            const int ly = qy * S + (local >> gshift), lx = qx * S + ((local & ((1 << gshift) - 1)) << 2);
            const int x = tx * T64 + lx, y = ty * T64 + ly;
            if (x >= pw || y >= ph) continue;
            const Prm m = PER == 1 ? one : prm_of(q);
            const int type = m.type, e = m.e;
            const uint32_t ofs = m.ofs;
            const bool border = m.border;
            if (type == 2 && !border) {  // interior CTB, EO: packed arithmetic
                const uint16_t* row = base + ly * ld + lx;
                uint2 o;
                switch (e) {
                    case 0: o = sao_eo4_packed<-1, 0>(row, ld, ofs, maxv); break;
                    case 1: o = sao_eo4_packed<0, -1>(row, ld, ofs, maxv); break;
                    case 2: o = sao_eo4_packed<-1, -1>(row, ld, ofs, maxv); break;
                    default: o = sao_eo4_packed<1, -1>(row, ld, ofs, maxv); break;
                }
                Wide<T>::store(dst + y * st + x, o);
                continue;
            }
            if (type != 2) {  // band offset or none: only the group itself is read
                uint2 o = *reinterpret_cast<const uint2*>(base + ly * ld + lx + 4);
                if (type == 1) o = sao_bo4_packed(o, ofs, bd - 5, m.pos, maxv);
                Wide<T>::store(dst + y * st + x, o);
                continue;
            }
            const bool need_rows = e != 0;  // EO: vertical and diagonal classes read rows y - 1, y + 1
            int rows[3][12];
#pragma unroll
            for (int rr = 0; rr < 3; rr++) {
                if (rr != 1 && !need_rows) continue;
                const uint2* p = reinterpret_cast<const uint2*>(base + (ly - 1 + rr) * ld + lx);
#pragma unroll
                for (int h = 0; h < 3; h++) {
                    const uint2 w = p[h];
                    rows[rr][4 * h + 0] = w.x & 0xffff;
                    rows[rr][4 * h + 1] = w.x >> 16;
                    rows[rr][4 * h + 2] = w.y & 0xffff;
                    rows[rr][4 * h + 3] = w.y >> 16;
                }
            }
            int out[4];  // EO on a CTB at the picture's edge: neighbours outside it leave samples as they are
            switch (e) {
                case 0: sao_eo4<-1, 0>(rows, ofs, maxv, true, x, y, pw, ph, out); break;
                case 1: sao_eo4<0, -1>(rows, ofs, maxv, true, x, y, pw, ph, out); break;
                case 2: sao_eo4<-1, -1>(rows, ofs, maxv, true, x, y, pw, ph, out); break;
                default: sao_eo4<1, -1>(rows, ofs, maxv, true, x, y, pw, ph, out); break;
            }
            Wide<T>::store(dst + y * st + x, make_uint2((uint32_t)out[0] | ((uint32_t)out[1] << 16),
                                                        (uint32_t)out[2] | ((uint32_t)out[3] << 16)));
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_loopfilter(const FrameDev* __restrict__ fr, const uint32_t* __restrict__ tile_prefix,
                                                    int nframes) {
    __shared__ __align__(16) LfTile<72> L;
    __shared__ hevcr_sao sp[4];
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const int f = frame_of(tile_prefix, nframes, g);
    const FrameDev* dp = &fr[f];
    const int tid = threadIdx.x;
    const int tile = (int)(g - tile_prefix[f]);
    const int wtile = (dp->fc.W + 63) >> 6;
    const int tx = tile % wtile, ty = tile / wtile;
    // SAO parameters of the CTBs inside the 64x64 tile: one (CTB 64) or 2x2 (CTB 32)
    const int cl2 = dp->fc.ctb_log2, per = 64 >> cl2, wctb = dp->fc.wctb, hctb = dp->fc.hctb;
    if (tid < per * per) {
        const int cx = tx * per + (tid % per), cy = ty * per + (tid / per);
        sp[tid] = gp(dp->sao)[(size_t)min(cy, hctb - 1) * wctb + min(cx, wctb - 1)];
    }
    lf_stage<T, 72>(L, dp, dp->rec, tx, ty, tid);
    __syncthreads();
    lf_params<72>(L, tx, ty, tid);
    __syncthreads();
    lf_vertical<72>(L, tid);
    __syncthreads();
    lf_horizontal<72>(L, tid);
    __syncthreads();
    if (per == 1)
        lf_sao<T, 1>(L, dp, sp, tx, ty, tid);
    else
        lf_sao<T, 2>(L, dp, sp, tx, ty, tid);
}

// Position-keyed 64-bit digest of one plane: sum over samples of
// (v + 1) * mix(index, plane key) modulo 2^64, with a SplitMix64 finaliser as
// the mix. The sum is order-free, so partial sums from any grid combine
// exactly (64-bit atomics); a changed sample changes it unless the weights
// collide (probability about 2^-64 per change pattern).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void __launch_bounds__(256) k_digest(const PlaneRef* __restrict__ planes, uint64_t* __restrict__ out) {
    const PlaneRef pr = planes[blockIdx.y];
    const T* src = static_cast<const T*>(pr.p);
    const uint64_t n = (uint64_t)pr.w * pr.h;
    uint64_t acc = 0;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull) {
        const uint32_t y = (uint32_t)(i / (uint32_t)pr.w), x = (uint32_t)(i - (uint64_t)y * pr.w);
        acc += ((uint64_t)src[(size_t)y * pr.stride + x] + 1) * mix64(i ^ pr.key);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&out[blockIdx.y]), (unsigned long long)acc);
}

// ---------------------------------------------------------------------------
// SHA-256 (FIPS 180-4) of pictures' planes, serialised as write_yuv writes
// them (picture.cpp:42-99), so that a plane's digest equals the hash of the
// bytes the reference writes for it. A message is hashed sequentially, so the
// parallelism is across planes: one lane per plane, the host sorting planes by
// length so that a wavefront's lanes run the same number of blocks. A lane
// keeps the chaining value and a 16-word schedule window in registers and
// fetches the next block's 16 words before compressing the current one.
// ---------------------------------------------------------------------------
__constant__ uint32_t kSha256K[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return __builtin_amdgcn_alignbit(x, x, n); }

// One compression (FIPS 180-4 6.2.2) of the 16 words w[] into st[].
__device__ __forceinline__ void sha256_block(uint32_t st[8], uint32_t w[16]) {
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6], h = st[7];
#pragma unroll
    for (int t = 0; t < 64; t++) {
        if (t >= 16) {
            const uint32_t w15 = w[(t - 15) & 15], w2 = w[(t - 2) & 15];
            const uint32_t s0 = rotr(w15, 7) ^ rotr(w15, 18) ^ (w15 >> 3);
            const uint32_t s1 = rotr(w2, 17) ^ rotr(w2, 19) ^ (w2 >> 10);
            w[t & 15] += s0 + w[(t - 7) & 15] + s1;
        }
        const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
        const uint32_t ch = (e & f) ^ (~e & g);
        const uint32_t t1 = h + S1 + ch + kSha256K[t] + w[t & 15];
        const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
        const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + S0 + maj;
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
    st[4] += e; st[5] += f; st[6] += g; st[7] += h;
}

// Reads the message of one plane 4 bytes at a time: row y, byte column x.
struct ShaCursor {
    const uint8_t* row;
    uint32_t x, rb, y, h;
    size_t rstride;  // bytes between rows in memory
    __device__ __forceinline__ uint32_t byte() {
        const uint32_t v = row[x];
        if (++x == rb) {
            x = 0;
            row += rstride;
            y++;
        }
        return v;
    }
    // The next 4 message bytes as a big-endian word; the caller guarantees
    // that they exist.
    __device__ __forceinline__ uint32_t word() {
        const uint8_t* p = row + x;
        if (x + 4 <= rb && (reinterpret_cast<uintptr_t>(p) & 3) == 0) {
            const uint32_t v = *gp(reinterpret_cast<const uint32_t*>(p));
            x += 4;
            if (x == rb) {
                x = 0;
                row += rstride;
                y++;
            }
            return __builtin_bswap32(v);
        }
        uint32_t v = byte() << 24;
        v |= byte() << 16;
        v |= byte() << 8;
        return v | byte();
    }
};

template <typename T>
__global__ void __launch_bounds__(64) k_sha256(const ShaPlane* __restrict__ planes, int nplanes, uint32_t* __restrict__ out) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= nplanes) return;
    const ShaPlane pl = planes[i];
    ShaCursor cur;
    cur.row = static_cast<const uint8_t*>(pl.src);
    cur.x = 0;
    cur.rb = (uint32_t)pl.w * (uint32_t)sizeof(T);
    cur.y = 0;
    cur.h = (uint32_t)pl.h;
    cur.rstride = (size_t)pl.stride * sizeof(T);
    const uint64_t len = (uint64_t)cur.rb * (uint64_t)pl.h;
    const uint64_t full = len / 64;
    uint32_t st[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                      0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    uint32_t nxt[16];
    if (full) {
#pragma unroll
        for (int k = 0; k < 16; k++) nxt[k] = cur.word();
    }
    for (uint64_t blk = 0; blk < full; blk++) {
        uint32_t w[16];
#pragma unroll
        for (int k = 0; k < 16; k++) w[k] = nxt[k];
        if (blk + 1 < full) {
#pragma unroll
            for (int k = 0; k < 16; k++) nxt[k] = cur.word();
        }
        sha256_block(st, w);
    }
    // padding (5.1.1): the rem remaining bytes, 0x80, zeros, the bit length
    const uint32_t rem = (uint32_t)(len - full * 64);
    uint32_t w[16];
    for (int k = 0; k < 16; k++) {
        uint32_t v = 0;
        for (int j = 0; j < 4; j++) {
            const uint32_t at = (uint32_t)(4 * k + j);
            const uint32_t b = at < rem ? cur.byte() : (at == rem ? 0x80u : 0u);
            v |= b << (24 - 8 * j);
        }
        w[k] = v;
    }
    const uint64_t bits = len * 8;
    if (rem < 56) {
        w[14] = (uint32_t)(bits >> 32);
        w[15] = (uint32_t)bits;
        sha256_block(st, w);
    } else {
        sha256_block(st, w);
        for (int k = 0; k < 14; k++) w[k] = 0;
        w[14] = (uint32_t)(bits >> 32);
        w[15] = (uint32_t)bits;
        sha256_block(st, w);
    }
    uint32_t* o = out + 8 * (size_t)pl.out;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = st[k];
}

// Plain 16-byte vector copy (counter calibration).
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Launchers (called by the host runtime; all on the caller's stream)
// ---------------------------------------------------------------------------

void launch_expand(hipStream_t s, const FrameDev* fr, const uint32_t* ctb_prefix, int nframes, uint32_t total_ctbs,
                   int max_ctb_log2) {
    // a lane per 4x4 cell of the largest CTB of the launch: 256, 64 or 16 lanes
    const unsigned lanes = 1u << (2 * (max_ctb_log2 - 2));
    if (total_ctbs) hipLaunchKernelGGL(k_expand, dim3(total_ctbs), dim3(lanes), 0, s, fr, ctb_prefix, nframes);
}

template <typename T>
void launch_mc(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total) {
    if (total) hipLaunchKernelGGL((k_mc<T, MC_UNI>), dim3(total), dim3(64), 0, s, fr, tile_prefix, nframes);
}

template <typename T>
void launch_mcb(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total) {
    if (total) hipLaunchKernelGGL((k_mc<T, MC_BI>), dim3(total), dim3(64), 0, s, fr, tile_prefix, nframes);
}

template <typename T>
void launch_mc8(hipStream_t s, const FrameDev* fr, const uint32_t* wg_prefix, int nframes, uint32_t total) {
    if (total) hipLaunchKernelGGL(k_mc8<T>, dim3(total), dim3(64), 0, s, fr, wg_prefix, nframes);
}

template <typename T>
void launch_mcs(hipStream_t s, const FrameDev* fr, const uint32_t* wave_prefix, int nframes, uint32_t total, bool bi,
                int max_rows) {
    if (!total) return;
    const dim3 g(total), b(64);
    if (bi) {
        hipLaunchKernelGGL((k_mcs<T, true, true, 8>), g, b, 0, s, fr, wave_prefix, nframes);
        hipLaunchKernelGGL((k_mcs<T, true, false, 8>), g, b, 0, s, fr, wave_prefix, nframes);
    } else if (max_rows > 8) {
        hipLaunchKernelGGL((k_mcs<T, false, true, 16>), g, b, 0, s, fr, wave_prefix, nframes);
        hipLaunchKernelGGL((k_mcs<T, false, false, 16>), g, b, 0, s, fr, wave_prefix, nframes);
    } else {
        hipLaunchKernelGGL((k_mcs<T, false, true, 8>), g, b, 0, s, fr, wave_prefix, nframes);
        hipLaunchKernelGGL((k_mcs<T, false, false, 8>), g, b, 0, s, fr, wave_prefix, nframes);
    }
}

template <typename T>
void launch_residual(hipStream_t s, const FrameDev* fr, const uint32_t* const* wg_prefix, int nframes, const uint32_t* total) {
    if (total[0]) hipLaunchKernelGGL(k_residual4<T>, dim3(total[0]), dim3(256), 0, s, fr, wg_prefix[0], nframes);
    if (total[1]) hipLaunchKernelGGL((k_residual_bf<T, 3>), dim3(total[1]), dim3(256), 0, s, fr, wg_prefix[1], nframes);
    if (total[2]) hipLaunchKernelGGL((k_residual_bf<T, 4>), dim3(total[2]), dim3(256), 0, s, fr, wg_prefix[2], nframes);
    if (total[3]) hipLaunchKernelGGL((k_residual_bf<T, 5>), dim3(total[3]), dim3(256), 0, s, fr, wg_prefix[3], nframes);
}

template <typename T>
void launch_intra(hipStream_t s, const FrameDev* fr, const uint32_t* diag_prefix, int nframes, int diag, uint32_t total,
                  int max_ctb_log2) {
    if (!total) return;
    const dim3 grid(total, 3);
    if (max_ctb_log2 <= 4)
        hipLaunchKernelGGL((k_intra<T, 4>), grid, dim3(64), (size_t)intra_lds(4, (int)sizeof(T)).bytes, s, fr, diag_prefix, nframes, diag);
    else if (max_ctb_log2 == 5)
        hipLaunchKernelGGL((k_intra<T, 5>), grid, dim3(64), (size_t)intra_lds(5, (int)sizeof(T)).bytes, s, fr, diag_prefix, nframes, diag);
    else
        hipLaunchKernelGGL((k_intra<T, 6>), grid, dim3(64), 0, s, fr, diag_prefix, nframes, diag);
}

template <typename T>
void launch_intra_persistent(hipStream_t s, const FrameDev* fr, int nframes, int max_nctb, const uint32_t* diag_prefix,
                             const uint32_t* wave_start, int nwaves, uint32_t total_items, uint32_t* ticket,
                             uint32_t* err, uint32_t spin_limit, int max_ctb_log2) {
    if (!total_items) return;
    const int per = (3 * max_nctb + 255) / 256;
    hipLaunchKernelGGL(k_intra_init, dim3((unsigned)(per < 64 ? per : 64), (unsigned)nframes), dim3(256), 0, s, fr, ticket);
    const dim3 grid(total_items < 4096 ? total_items : 4096);
    if (max_ctb_log2 <= 4)
        hipLaunchKernelGGL((k_intra_p<T, 4>), grid, dim3(64), (size_t)intra_lds(4, (int)sizeof(T)).bytes, s, fr, diag_prefix,
                           nframes, wave_start, nwaves, ticket, err, spin_limit);
    else if (max_ctb_log2 == 5)
        hipLaunchKernelGGL((k_intra_p<T, 5>), grid, dim3(64), (size_t)intra_lds(5, (int)sizeof(T)).bytes, s, fr, diag_prefix,
                           nframes, wave_start, nwaves, ticket, err, spin_limit);
    else
        hipLaunchKernelGGL((k_intra_p<T, 6>), grid, dim3(64), 0, s, fr, diag_prefix, nframes, wave_start, nwaves, ticket,
                           err, spin_limit);
}

template <typename T>
void launch_deblock(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total_tiles) {
    if (total_tiles) hipLaunchKernelGGL(k_deblock<T>, dim3(total_tiles), dim3(256), 0, s, fr, tile_prefix, nframes);
}

template <typename T>
void launch_sao(hipStream_t s, const FrameDev* fr, const uint32_t* ctb_prefix, int nframes, uint32_t total_ctbs) {
    if (total_ctbs) hipLaunchKernelGGL(k_sao<T>, dim3(total_ctbs), dim3(256), 0, s, fr, ctb_prefix, nframes);
}

template <typename T>
void launch_loopfilter(hipStream_t s, const FrameDev* fr, const uint32_t* tile_prefix, int nframes, uint32_t total_tiles) {
    if (total_tiles) hipLaunchKernelGGL(k_loopfilter<T>, dim3(total_tiles), dim3(256), 0, s, fr, tile_prefix, nframes);
}

// ---------------------------------------------------------------------------
// Egress: pictures serialised as the reference writes them (picture.cpp:42-99
// write_yuv: the conformance window of each plane, rows without padding,
// samples as bytes at 8 bits and as 16-bit little-endian words above) or as
// hevc_api.h's uint16 frame planes (hevc_api.cpp:49-77). A plane's output is
// one flat array of w * h samples; each lane produces four consecutive ones
// (one 4- or 8-byte store where the plane's destination is aligned to it)
// and finds their source rows itself, so crops and odd widths need no special
// case. The destination may be mapped pinned host memory: stores go over the
// bus directly, and the same generic stores serve both.
// ---------------------------------------------------------------------------
template <typename T, typename U>
__global__ void __launch_bounds__(256) k_egress(const EgressPlane* __restrict__ planes) {
    const EgressPlane pl = planes[blockIdx.y];
    const uint64_t n = (uint64_t)pl.w * (uint64_t)pl.h;
    const Glob<const T>* src = gp(static_cast<const T*>(pl.src));
    U* dst = static_cast<U*>(pl.dst);
    const bool aligned = (reinterpret_cast<uintptr_t>(dst) & (4 * sizeof(U) - 1)) == 0;
    const uint64_t step = (uint64_t)gridDim.x * 1024ull;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * 256ull + threadIdx.x) * 4ull; i0 < n; i0 += step) {
        uint32_t y = (uint32_t)(i0 / (uint64_t)pl.w), x = (uint32_t)(i0 - (uint64_t)y * (uint64_t)pl.w);
        uint32_t v[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (i0 + k < n) v[k] = src[(size_t)y * (size_t)pl.stride + x];
            if (++x == (uint32_t)pl.w) {
                x = 0;
                y++;
            }
        }
        if (aligned && i0 + 4 <= n) {
            if (sizeof(U) == 1)
                *reinterpret_cast<uint32_t*>(dst + i0) = v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24);
            else
                *reinterpret_cast<uint2*>(dst + i0) = make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16));
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (i0 + k < n) dst[i0 + k] = (U)v[k];
        }
    }
}

void launch_egress(hipStream_t s, const EgressPlane* planes, int nplanes, int in_bytes, int out_bytes, uint64_t max_samples) {
    if (nplanes <= 0 || max_samples == 0) return;
    const uint64_t blocks = (max_samples + 1023) / 1024;
    const dim3 grid((unsigned)(blocks < 512 ? blocks : 512), (unsigned)nplanes);
    if (in_bytes == 1 && out_bytes == 1)
        hipLaunchKernelGGL((k_egress<uint8_t, uint8_t>), grid, dim3(256), 0, s, planes);
    else if (in_bytes == 1)
        hipLaunchKernelGGL((k_egress<uint8_t, uint16_t>), grid, dim3(256), 0, s, planes);
    else
        hipLaunchKernelGGL((k_egress<uint16_t, uint16_t>), grid, dim3(256), 0, s, planes);
}

void launch_digest(hipStream_t s, const PlaneRef* planes, int nplanes, int bpp, uint64_t* out) {
    if (nplanes <= 0) return;
    if (bpp == 1)
        hipLaunchKernelGGL(k_digest<uint8_t>, dim3(64, nplanes), dim3(256), 0, s, planes, out);
    else
        hipLaunchKernelGGL(k_digest<uint16_t>, dim3(64, nplanes), dim3(256), 0, s, planes, out);
}

void launch_sha256(hipStream_t s, const ShaPlane* planes, int nplanes, int bpp, uint32_t* out) {
    if (nplanes <= 0) return;
    const dim3 grid((unsigned)((nplanes + 63) / 64));
    if (bpp == 1)
        hipLaunchKernelGGL(k_sha256<uint8_t>, grid, dim3(64), 0, s, planes, nplanes, out);
    else
        hipLaunchKernelGGL(k_sha256<uint16_t>, grid, dim3(64), 0, s, planes, nplanes, out);
}

// The motion-compensation strips of a frame from its PU bands: a lane per
// band writes the band's strips (16 bytes each, consecutive).
__global__ void __launch_bounds__(256) k_strip_expand(const McBand* __restrict__ bands, uint32_t n,
                                                     McStrip* __restrict__ strips) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const McBand b = *gp(bands + i);
    uint4 w;
    __builtin_memcpy(&w, &b.s, sizeof(w));
    for (uint32_t k = 0; k < b.n; k++) {
        uint4 v = w;
        v.x = (w.x & 0xffff0000u) | ((w.x + 4 * k) & 0xffffu);  // x: the low half of the first dword
        *gp(reinterpret_cast<uint4*>(strips + b.dst + k)) = v;
    }
}

void launch_strip_expand(hipStream_t s, const McBand* bands, uint32_t n, McStrip* strips) {
    if (n) hipLaunchKernelGGL(k_strip_expand, dim3((n + 255) / 256), dim3(256), 0, s, bands, n, strips);
}

void launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
}

#define HEVCR_INSTANTIATE(T)                                                                                  \
    template void launch_mc<T>(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);                 \
    template void launch_mc8<T>(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);                \
    template void launch_mcs<T>(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t, bool, int);     \
    template void launch_mcb<T>(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);                \
    template void launch_residual<T>(hipStream_t, const FrameDev*, const uint32_t* const*, int, const uint32_t*); \
    template void launch_intra<T>(hipStream_t, const FrameDev*, const uint32_t*, int, int, uint32_t, int);    \
    template void launch_intra_persistent<T>(hipStream_t, const FrameDev*, int, int, const uint32_t*, const uint32_t*, \
                                             int, uint32_t, uint32_t*, uint32_t*, uint32_t, int);                 \
    template void launch_deblock<T>(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);           \
    template void launch_sao<T>(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);            \
    template void launch_loopfilter<T>(hipStream_t, const FrameDev*, const uint32_t*, int, uint32_t);

HEVCR_INSTANTIATE(uint8_t)
HEVCR_INSTANTIATE(uint16_t)

}  // namespace hevcr
