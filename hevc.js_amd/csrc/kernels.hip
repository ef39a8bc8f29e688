// HIP kernels of the reconstruction engine (gfx950, 64-lane wavefronts).
//
// Stage order for one picture (reference frame driver: decoder.cpp:52-304,
// with the bit-exact re-ordering justified by the dependency structure):
//   expand  : CU / TU / PU records -> per-4x4 side information and motion
//   mc      : inter prediction of every PU into the picture
//   residual: dequantisation + inverse transform of every coded block; inter
//             blocks are added to the prediction in place, intra residuals
//             are kept for the intra pass
//   intra   : CTU-diagonal wavefront of intra prediction + reconstruction
//   deblock : vertical edges, then horizontal edges, in place
//   sao     : deblocked picture -> output picture
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"
#include "tables.hpp"

namespace hevcr {

__device__ __forceinline__ int clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int clip16(int v) { return clip3(-32768, 32767, v); }
__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

// ---------------------------------------------------------------------------
// Syntax expansion: the reference's grid writes done while parsing
// (coding_tree.cpp:956-1054 for CUs, 1233-1247 for TUs,
// inter_prediction.cpp:17-27 for motion).
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(64) k_expand_cu(const hevcr_cu* __restrict__ cus, uint32_t n, FrameConst fc,
                                                  uint32_t* __restrict__ info4) {
    const uint32_t ci = blockIdx.x;
    if (ci >= n) return;
    const hevcr_cu cu = cus[ci];
    const int cb = 1 << cu.log2_size, cells = cb >> 2;
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
    const bool nxn = cu.part_mode == HEVCR_PART_NxN;
    const uint32_t base = (uint32_t)cu.log2_size | ((uint32_t)cu.pred_mode << info::kModeShift) |
                          ((uint32_t)(uint8_t)cu.qp_y << info::kQpShift) |
                          ((uint32_t)cu.intra_chroma << info::kChromaModeShift);
    for (int i = threadIdx.x; i < cells * cells; i += blockDim.x) {
        const int dx = (i % cells) * 4, dy = (i / cells) * 4;
        uint32_t w = base;
        const int pu = nxn ? ((dy >= cb / 2) * 2 + (dx >= cb / 2)) : 0;
        w |= (uint32_t)cu.intra_luma[pu] << info::kLumaModeShift;
        if (dx == 0 || dx == vx) w |= info::kEdgeV;
        if (dy == 0 || dy == hy) w |= info::kEdgeH;
        info4[(size_t)((cu.y + dy) >> 2) * fc.gw + ((cu.x + dx) >> 2)] = w;
    }
}

__global__ void __launch_bounds__(256) k_expand_tu(const hevcr_tu* __restrict__ tus, uint32_t n, FrameConst fc,
                                                   uint32_t* __restrict__ info4) {
    const uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= n) return;
    const hevcr_tu t = tus[ti];
    const int cells = 1 << (t.log2_size - 2);
    const uint32_t set = (uint32_t)t.log2_size | ((t.cbf & HEVCR_CBF_Y) ? info::kCbf : 0u);
    for (int j = 0; j < cells; j++)
        for (int i = 0; i < cells; i++) {
            uint32_t* p = &info4[(size_t)((t.y >> 2) + j) * fc.gw + (t.x >> 2) + i];
            uint32_t w = (*p & ~(info::kLog2Mask | info::kCbf)) | set;
            if (i == 0) w |= info::kEdgeV;
            if (j == 0) w |= info::kEdgeH;
            *p = w;
        }
}

__global__ void __launch_bounds__(64) k_expand_pu(const hevcr_pu* __restrict__ pus, uint32_t n, FrameConst fc,
                                                  Motion4* __restrict__ mot) {
    const uint32_t pi = blockIdx.x;
    if (pi >= n) return;
    const hevcr_pu pu = pus[pi];
    Motion4 m;
    m.mv[0][0] = pu.mv[0][0];
    m.mv[0][1] = pu.mv[0][1];
    m.mv[1][0] = pu.mv[1][0];
    m.mv[1][1] = pu.mv[1][1];
    m.ref[0] = pu.ref_idx[0];
    m.ref[1] = pu.ref_idx[1];
    m.flags = pu.pred_flags;
    m.pad = 0;
    const int cw = pu.w >> 2, ch = pu.h >> 2;
    for (int i = threadIdx.x; i < cw * ch; i += blockDim.x)
        mot[(size_t)((pu.y >> 2) + i / cw) * fc.gw + (pu.x >> 2) + i % cw] = m;
}

// ---------------------------------------------------------------------------
// Motion compensation (interpolation.cpp:45-411): one 64-lane workgroup per
// tile of at most 16x16 luma (+ two 8x8 chroma). Reference samples of the
// tile's footprint are staged in LDS with clamp-on-load (the reference's
// edge clamping), filtered horizontally into int16, then vertically into
// int16 (the reference stores both intermediates as int16), then weighted.
// ---------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(64) k_mc(const McTile* __restrict__ tiles, uint32_t ntiles,
                                           const hevcr_pu* __restrict__ pus, FrameConst fc, RefPlanes refs,
                                           T* __restrict__ y_pl, T* __restrict__ cb_pl, T* __restrict__ cr_pl) {
    __shared__ int16_t fp[23 * 23];
    __shared__ int16_t tmp[23 * 16];
    __shared__ int16_t pred[2][16 * 16];
    const uint32_t tix = blockIdx.x;
    if (tix >= ntiles) return;
    const McTile tile = tiles[tix];
    const hevcr_pu pu = pus[tile.pu];
    const int lane = threadIdx.x;
    const int sh1 = fc.bd - 8 < 4 ? fc.bd - 8 : 4;
    const int shw = 14 - fc.bd > 2 ? 14 - fc.bd : 2;  // shift1 of the weighting
    T* planes[3] = {y_pl, cb_pl, cr_pl};

    for (int c = 0; c < 3; c++) {
        const int sub = c ? 1 : 0;
        const int bw = tile.w >> sub, bh = tile.h >> sub;
        const int ntap = c ? 4 : 8, half = c ? 1 : 3;
        const int pw = c ? fc.W >> 1 : fc.W, ph = c ? fc.H >> 1 : fc.H;
        const int fw = bw + ntap - 1, fh = bh + ntap - 1;
        const int tx = tile.x >> sub, ty = tile.y >> sub;
        for (int l = 0; l < 2; l++) {
            if (!((pu.pred_flags >> l) & 1)) continue;
            const T* ref = static_cast<const T*>(refs.p[l][pu.ref_idx[l]][c]);
            const int mvx = pu.mv[l][0], mvy = pu.mv[l][1];
            const int xi = tx + (c ? (mvx >> 3) : (mvx >> 2)), yi = ty + (c ? (mvy >> 3) : (mvy >> 2));
            const int fx = c ? (mvx & 7) : (mvx & 3), fy = c ? (mvy & 7) : (mvy & 3);
            const int st = fc.stride[c];
            for (int i = lane; i < fw * fh; i += 64) {
                const int yy = clip3(0, ph - 1, yi - half + i / fw), xx = clip3(0, pw - 1, xi - half + i % fw);
                fp[i] = (int16_t)ref[(size_t)yy * st + xx];
            }
            __syncthreads();
            const int* hx = c ? kChromaTaps[fx] : kLumaTaps[fx];
            const int* vy = c ? kChromaTaps[fy] : kLumaTaps[fy];
            for (int i = lane; i < fh * bw; i += 64) {
                const int r = i / bw, x = i % bw;
                int s = 0;
                for (int k = 0; k < ntap; k++) s += hx[k] * fp[r * fw + x + k];
                tmp[i] = (int16_t)(s >> sh1);
            }
            __syncthreads();
            for (int i = lane; i < bh * bw; i += 64) {
                const int r = i / bw, x = i % bw;
                int s = 0;
                for (int k = 0; k < ntap; k++) s += vy[k] * tmp[(r + k) * bw + x];
                pred[l][i] = (int16_t)(s >> 6);
            }
            __syncthreads();
        }
        // Weighted sample prediction (interpolation.cpp:215-300, 392-410)
        const int f0 = pu.pred_flags & 1, f1 = (pu.pred_flags >> 1) & 1;
        int lwd = 0, w0 = 0, w1 = 0, o0 = 0, o1 = 0;
        if (fc.weighted) {
            lwd = (c ? fc.chroma_log2_wd : fc.luma_log2_wd) + shw;
            const int r0 = pu.ref_idx[0] >= 0 ? pu.ref_idx[0] : 0, r1 = pu.ref_idx[1] >= 0 ? pu.ref_idx[1] : 0;
            w0 = fc.wt[0][r0][c];
            w1 = fc.wt[1][r1][c];
            o0 = fc.off[0][r0][c] * (1 << (fc.bd - 8));
            o1 = fc.off[1][r1][c] * (1 << (fc.bd - 8));
        }
        T* dst = planes[c];
        const int st = fc.stride[c];
        for (int i = lane; i < bh * bw; i += 64) {
            const int a = pred[0][i], b = pred[1][i];
            int v;
            if (!fc.weighted) {
                v = (f0 && f1) ? (a + b + (1 << shw)) >> (shw + 1) : ((f0 ? a : b) + (1 << (shw - 1))) >> shw;
            } else if (f0 && f1) {
                v = (a * w0 + b * w1 + (o0 + o1 + 1) * (1 << lwd)) >> (lwd + 1);
            } else {
                v = f0 ? ((a * w0 + (1 << (lwd - 1))) >> lwd) + o0 : ((b * w1 + (1 << (lwd - 1))) >> lwd) + o1;
            }
            dst[(size_t)(ty + i / bw) * st + tx + i % bw] = (T)clip3(0, fc.maxv, v);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Dequantisation + inverse transform (transform.cpp:240-394). One 256-lane
// workgroup handles 1024 coefficients: one 32x32 block, or 4 16x16, 16 8x8,
// 64 4x4 blocks. Both 1-D passes are integer matrix products over the
// 128-entry cosine table (identical to the reference's partial butterflies)
// with the reference's int16 clip after each pass.
// ---------------------------------------------------------------------------

template <typename T, int LOG2>
__global__ void __launch_bounds__(256) k_residual(const ResBlock* __restrict__ blocks, uint32_t nblocks,
                                                  const int16_t* __restrict__ coeffs, int16_t* __restrict__ resid,
                                                  FrameConst fc, T* __restrict__ y_pl, T* __restrict__ cb_pl,
                                                  T* __restrict__ cr_pl) {
    constexpr int N = 1 << LOG2, NN = N * N, BPW = 1024 / NN, STEP = 32 >> LOG2;
    __shared__ int16_t buf0[1024];
    __shared__ int16_t buf1[1024];
    __shared__ ResBlock blk[BPW];
    __shared__ signed char cosT[128];
    const int tid = threadIdx.x;
    const uint32_t first = blockIdx.x * BPW;
    if (tid < 128) cosT[tid] = kCos128[tid];
    if (tid < BPW) {
        if (first + tid < nblocks) {
            blk[tid] = blocks[first + tid];
        } else {
            blk[tid].log2 = 0;  // empty slot
        }
    }
    __syncthreads();
    const int bd = fc.bd;
    // load + dequantise (transform.cpp:293-368, flat scaling)
    for (int i = tid; i < 1024; i += 256) {
        const int b = i / NN, e = i % NN;
        const ResBlock& rb = blk[b];
        int16_t d = 0;
        if (rb.log2) {
            const int c = coeffs[(size_t)rb.coeff_off + e];
            if (c) {
                const int bdshift = bd + LOG2 - 5;
                const int64_t add = (int64_t)1 << (bdshift - 1);
                int64_t v = (int64_t)c * 16 * kLevelScale[rb.qp % 6];
                v = (v * ((int64_t)1 << (rb.qp / 6)) + add) >> bdshift;
                d = (int16_t)clip16((int)v);
            }
        }
        buf0[i] = d;
    }
    __syncthreads();
    // column pass: t[k][j] = clip16((sum_m M(m,k) d[m][j] + 64) >> 7)
    for (int i = tid; i < 1024; i += 256) {
        const int b = i / NN, e = i % NN, k = e / N, j = e % N;
        const ResBlock& rb = blk[b];
        if (!rb.log2) continue;
        const int16_t* d = &buf0[b * NN];
        int s = 0;
        if (LOG2 == 2 && (rb.flags & kResDst)) {
            for (int m = 0; m < 4; m++) s += kDst4[m][k] * d[m * 4 + j];
        } else {
            for (int m = 0; m < N; m++) s += cosT[((2 * k + 1) * m * STEP) & 127] * d[m * N + j];
        }
        buf1[i] = (int16_t)clip16((s + 64) >> 7);
    }
    __syncthreads();
    // row pass: r[y][x] = clip16((sum_m M(m,x) t[y][m] + rnd) >> (20 - bd)), then output
    const int sh2 = 20 - bd;
    for (int i = tid; i < 1024; i += 256) {
        const int b = i / NN, e = i % NN, y = e / N, x = e % N;
        const ResBlock& rb = blk[b];
        if (!rb.log2) continue;
        const int16_t* t = &buf1[b * NN];
        int s = 0;
        int r;
        if (LOG2 == 2 && (rb.flags & kResTskip)) {
            // transform skip as the reference computes it: (d + 2^(14-bd)) >> (15-bd)
            // (transform.cpp:380-388), no clip
            const int sh = 15 - bd;
            r = (int16_t)((buf0[i] + (sh > 0 ? 1 << (sh - 1) : 0)) >> sh);
        } else {
        if (LOG2 == 2 && (rb.flags & kResDst)) {
            for (int m = 0; m < 4; m++) s += kDst4[m][x] * t[y * 4 + m];
        } else {
            for (int m = 0; m < N; m++) s += cosT[((2 * x + 1) * m * STEP) & 127] * t[y * N + m];
        }
        r = clip16((s + (1 << (sh2 - 1))) >> sh2);
        }
        if (rb.flags & kResToPlane) {
            T* pl = rb.cidx == 0 ? y_pl : (rb.cidx == 1 ? cb_pl : cr_pl);
            T* p = &pl[(size_t)(rb.y + y) * fc.stride[rb.cidx] + rb.x + x];
            *p = (T)clip3(0, fc.maxv, (int)*p + r);
        } else {
            resid[(size_t)rb.coeff_off + e] = (int16_t)r;
        }
    }
}

// ---------------------------------------------------------------------------
// Intra prediction + reconstruction (intra_prediction.cpp:44-481,
// coding_tree.cpp:371-396). One 64-lane workgroup per (CTU, component)
// walks that component's intra blocks of the CTU in decoding order; the CTU
// and its reconstructed top / left borders live in LDS. CTUs run in
// diagonal waves (x + 2y constant), so every neighbour CTU a block may use
// (left, top-left, top, top-right) finished in an earlier wave.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t zscan4(int bx, int by) {
    uint32_t z = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) z |= (((uint32_t)bx >> i) & 1u) << (2 * i) | (((uint32_t)by >> i) & 1u) << (2 * i + 1);
    return z;
}

template <typename T>
__global__ void __launch_bounds__(64) k_intra(const uint32_t* __restrict__ ctus, uint32_t nctus,
                                              const uint32_t* __restrict__ op_begin,
                                              const IntraOp* __restrict__ ops, const int16_t* __restrict__ resid,
                                              FrameConst fc, T* __restrict__ y_pl, T* __restrict__ cb_pl,
                                              T* __restrict__ cr_pl) {
    __shared__ uint16_t tile[64 * 64];
    __shared__ uint16_t top[129];
    __shared__ uint16_t left[64];
    __shared__ int lin[129];
    __shared__ int lin2[129];
    __shared__ int reduce_tmp[2];
    const uint32_t bi = blockIdx.x;
    if (bi >= nctus) return;
    const int c = blockIdx.y;
    const uint32_t ctu = ctus[bi];
    const int lane = threadIdx.x;
    const int sub = c ? 1 : 0;
    const int S = (1 << fc.ctb_log2) >> sub, slog = fc.ctb_log2 - sub;
    const int rx = (int)(ctu % (uint32_t)fc.wctb), ry = (int)(ctu / (uint32_t)fc.wctb);
    const int cx0 = rx * S, cy0 = ry * S;
    const int pw = c ? fc.W >> 1 : fc.W, ph = c ? fc.H >> 1 : fc.H;
    const int st = fc.stride[c];
    T* pl = c == 0 ? y_pl : (c == 1 ? cb_pl : cr_pl);
    const int bd = fc.bd, maxv = fc.maxv;

    // stage the CTU and its borders
    for (int i = lane; i < S * S; i += 64) {
        const int y = cy0 + i / S, x = cx0 + i % S;
        tile[i] = (x < pw && y < ph) ? (uint16_t)pl[(size_t)y * st + x] : 0;
    }
    for (int i = lane; i <= 2 * S; i += 64) {
        const int x = cx0 - 1 + i, y = cy0 - 1;
        top[i] = (y >= 0 && x >= 0 && x < pw) ? (uint16_t)pl[(size_t)y * st + x] : 0;
    }
    for (int i = lane; i < S; i += 64) {
        const int x = cx0 - 1, y = cy0 + i;
        left[i] = (x >= 0 && y < ph) ? (uint16_t)pl[(size_t)y * st + x] : 0;
    }
    __syncthreads();

    const uint32_t o0 = op_begin[ctu * 3 + c], o1 = op_begin[ctu * 3 + c + 1];
    for (uint32_t oi = o0; oi < o1; oi++) {
        const IntraOp op = ops[oi];
        const int n = 1 << op.log2, log2 = op.log2, mode = op.mode;
        const int xt = op.x, yt = op.y;  // component coordinates
        const uint32_t zcur = zscan4(((xt - cx0) << sub) >> 2, ((yt - cy0) << sub) >> 2);
        const int tot = 4 * n + 1;

        // 1. gather with availability (intra_prediction.cpp:102-175)
        //    linear index: 0..2n-1 left column bottom-up, 2n corner, 2n+1.. top row
        int myv[3], mya[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int i = lane * 3 + q;
            int sx, sy;
            if (i < 2 * n) {
                sx = xt - 1;
                sy = yt + 2 * n - 1 - i;
            } else if (i == 2 * n) {
                sx = xt - 1;
                sy = yt - 1;
            } else {
                sx = xt + (i - 2 * n - 1);
                sy = yt - 1;
            }
            int a = 0, v = 0;
            if (i < tot && sx >= 0 && sy >= 0 && sx < pw && sy < ph) {
                const int nx = sx >> slog, ny = sy >> slog;
                if (nx != rx || ny != ry) {
                    a = ny * fc.wctb + nx < (int)ctu;
                } else {
                    a = zscan4(((sx - cx0) << sub) >> 2, ((sy - cy0) << sub) >> 2) < zcur;
                }
                if (a) {
                    if (sy < cy0)
                        v = top[sx - (cx0 - 1)];
                    else if (sx < cx0)
                        v = left[sy - cy0];
                    else
                        v = tile[(sy - cy0) * S + (sx - cx0)];
                }
            }
            myv[q] = v;
            mya[q] = a;
            if (i < tot) lin[i] = v;
        }
        // 2. substitution (intra_prediction.cpp:177-197): nearest available
        //    index at or below i, else the first available index
        int last_local = -1, agg = -1;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (mya[q] && lane * 3 + q < tot) agg = lane * 3 + q;
        int incl = agg;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(incl, off, 64);
            if (lane >= off) incl = o > incl ? o : incl;
        }
        int excl = __shfl_up(incl, 1, 64);
        if (lane == 0) excl = -1;
        int first = 1 << 20;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (mya[q] && lane * 3 + q < tot && lane * 3 + q < first) first = lane * 3 + q;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const int o = __shfl_xor(first, off, 64);
            first = o < first ? o : first;
        }
        __syncthreads();
        last_local = excl;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int i = lane * 3 + q;
            if (i >= tot) break;
            if (mya[q]) last_local = i;
            int v;
            if (first >= (1 << 20))
                v = 1 << (bd - 1);
            else if (mya[q])
                v = myv[q];
            else
                v = lin[last_local >= 0 ? last_local : first];
            lin2[i] = v;
        }
        __syncthreads();

        // 3. reference smoothing (luma only in 4:2:0; intra_prediction.cpp:217-257, 436-471)
        int* ref = lin2;
        if (c == 0 && !fc.intra_smoothing_disabled && mode != 1 && log2 > 2) {
            const int dv = iabs(mode - 26) < iabs(mode - 10) ? iabs(mode - 26) : iabs(mode - 10);
            const int thr = log2 == 3 ? 7 : (log2 == 4 ? 1 : 0);
            if (dv > thr) {
                int bil = 0;
                if (fc.strong_intra_smoothing && n == 32) {
                    const int tl = lin2[2 * n], t = 1 << (bd - 5);
                    bil = iabs(tl + lin2[4 * n] - 2 * lin2[3 * n]) < t && iabs(tl + lin2[0] - 2 * lin2[n]) < t;
                }
                for (int i = lane; i < tot; i += 64) {
                    int v;
                    if (i == 0 || i == 4 * n) {
                        v = lin2[i];
                    } else if (bil) {
                        const int tl = lin2[2 * n];
                        if (i == 2 * n)
                            v = tl;
                        else if (i > 2 * n)
                            v = ((2 * n - (i - 2 * n)) * tl + (i - 2 * n) * lin2[4 * n] + n) >> (log2 + 1);
                        else
                            v = ((2 * n - (2 * n - i)) * tl + (2 * n - i) * lin2[0] + n) >> (log2 + 1);
                    } else {
                        v = (lin2[i - 1] + 2 * lin2[i] + lin2[i + 1] + 2) >> 2;
                    }
                    lin[i] = v;
                }
                __syncthreads();
                ref = lin;
            }
        }
        // top[j] = ref[2n + j], left[j] = ref[2n - j]
#define TOP(j) ref[2 * n + (j)]
#define LEFT(j) ref[2 * n - (j)]
        int dc = 0;
        if (mode == 1) {
            int part = 0;
            if (lane < n) part = TOP(lane + 1) + LEFT(lane + 1);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off, 64);
            dc = (part + n) >> (log2 + 1);
        }
        const int ang = kIntraAngle[mode], vert = mode >= 18;
        const int inv = kInvAngle[mode];
        for (int i = lane; i < n * n; i += 64) {
            const int y = i >> log2, x = i & (n - 1);
            int p;
            if (mode == 0) {
                p = ((n - 1 - x) * LEFT(y + 1) + (x + 1) * TOP(n + 1) + (n - 1 - y) * TOP(x + 1) + (y + 1) * LEFT(n + 1) + n) >>
                    (log2 + 1);
            } else if (mode == 1) {
                p = dc;
                if (c == 0 && n < 32) {
                    if (x == 0 && y == 0)
                        p = (TOP(1) + LEFT(1) + 2 * dc + 2) >> 2;
                    else if (y == 0)
                        p = (TOP(x + 1) + 3 * dc + 2) >> 2;
                    else if (x == 0)
                        p = (LEFT(y + 1) + 3 * dc + 2) >> 2;
                }
            } else {
                const int pos = vert ? y : x, oth = vert ? x : y;
                const int idx = ((pos + 1) * ang) >> 5, fr = ((pos + 1) * ang) & 31;
                const int r = oth + 1 + idx;
                // main array with the negative-index projection of the side array
                // (intra_prediction.cpp:346-368: source index clamped to 2n)
                auto mainv = [&](int k) -> int {
                    if (k >= 0) return vert ? TOP(k) : LEFT(k);
                    int si = ((-k) * inv + 128) >> 8;
                    si = si > 2 * n ? 2 * n : si;
                    return vert ? LEFT(si) : TOP(si);
                };
                p = fr ? ((32 - fr) * mainv(r) + fr * mainv(r + 1) + 16) >> 5 : mainv(r);
                if (c == 0 && n < 32) {
                    if (mode == 26 && x == 0) p = clip3(0, maxv, p + ((LEFT(y + 1) - LEFT(0)) >> 1));
                    if (mode == 10 && y == 0) p = clip3(0, maxv, p + ((TOP(x + 1) - TOP(0)) >> 1));
                }
            }
            const int rv = op.cbf ? resid[(size_t)op.resid_off + i] : 0;
            tile[(yt - cy0 + y) * S + (xt - cx0 + x)] = (uint16_t)clip3(0, maxv, (int)(int16_t)p + rv);
        }
#undef TOP
#undef LEFT
        __syncthreads();
    }
    (void)reduce_tmp;
    // write the CTU back
    for (int i = lane; i < S * S; i += 64) {
        const int y = cy0 + i / S, x = cx0 + i % S;
        if (x < pw && y < ph) pl[(size_t)y * st + x] = (T)tile[i];
    }
}

// ---------------------------------------------------------------------------
// Deblocking (deblocking.cpp:125-614). One lane per 4-sample edge segment.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int ref_poc(const FrameConst& fc, const Motion4& m, int l) {
    if (!((m.flags >> l) & 1) || m.ref[l] < 0 || m.ref[l] >= fc.num_ref[l]) return -999999;
    return fc.ref_poc[l][m.ref[l]];
}

__device__ __forceinline__ int mv_far(const int16_t* a, const int16_t* b) {
    return iabs(a[0] - b[0]) >= 4 || iabs(a[1] - b[1]) >= 4;
}

// derive_bs (deblocking.cpp:125-247); gp/gq are 4x4-cell indices, vert the
// edge direction, qpos the Q-side coordinate across the edge.
__device__ int boundary_strength(const FrameConst& fc, const uint32_t* info4, const Motion4* mot, size_t gp, size_t gq,
                                 int qpos) {
    const uint32_t ip = info4[gp], iq = info4[gq];
    const int mp = (ip >> info::kModeShift) & 3, mq = (iq >> info::kModeShift) & 3;
    if (mp == HEVCR_MODE_INTRA || mq == HEVCR_MODE_INTRA) return 2;
    const int tp = 1 << (ip & info::kLog2Mask), tq = 1 << (iq & info::kLog2Mask);
    const bool tu_edge = (qpos % tq == 0) || (qpos % tp == 0);
    if (tu_edge && ((ip | iq) & info::kCbf)) return 1;
    const Motion4 a = mot[gp], b = mot[gq];
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

// Luma decision + filter of one 4-line segment; samples addressed as
// base[k*lstep + i*istep] with i = 0..3 on the Q side and -1..-4 on P.
template <typename T>
__device__ void luma_segment(T* base, ptrdiff_t lstep, ptrdiff_t istep, int beta, int tc, int maxv) {
#define SP(k, i) base[(k) * lstep - ((i) + 1) * istep]
#define SQ(k, i) base[(k) * lstep + (i) * istep]
    const int dp0 = iabs(SP(0, 2) - 2 * SP(0, 1) + SP(0, 0)), dp3 = iabs(SP(3, 2) - 2 * SP(3, 1) + SP(3, 0));
    const int dq0 = iabs(SQ(0, 2) - 2 * SQ(0, 1) + SQ(0, 0)), dq3 = iabs(SQ(3, 2) - 2 * SQ(3, 1) + SQ(3, 0));
    if (dp0 + dq0 + dp3 + dq3 >= beta) return;
    bool strong = true;
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
        const int k = kk * 3, dpq = kk ? dp3 + dq3 : dp0 + dq0;
        const int p0 = SP(k, 0), p3 = SP(k, 3), q0 = SQ(k, 0), q3 = SQ(k, 3);
        if (!(2 * dpq < (beta >> 2) && iabs(p3 - p0) + iabs(q0 - q3) < (beta >> 3) && iabs(p0 - q0) < ((5 * tc + 1) >> 1)))
            strong = false;
    }
    const int side = (beta + (beta >> 1)) >> 3;
    const bool dep = dp0 + dp3 < side, deq = dq0 + dq3 < side;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int P0 = SP(k, 0), P1 = SP(k, 1), P2 = SP(k, 2), P3 = SP(k, 3);
        const int Q0 = SQ(k, 0), Q1 = SQ(k, 1), Q2 = SQ(k, 2), Q3 = SQ(k, 3);
        if (strong) {
            SP(k, 0) = (T)clip3(P0 - 2 * tc, P0 + 2 * tc, (P2 + 2 * P1 + 2 * P0 + 2 * Q0 + Q1 + 4) >> 3);
            SP(k, 1) = (T)clip3(P1 - 2 * tc, P1 + 2 * tc, (P2 + P1 + P0 + Q0 + 2) >> 2);
            SP(k, 2) = (T)clip3(P2 - 2 * tc, P2 + 2 * tc, (2 * P3 + 3 * P2 + P1 + P0 + Q0 + 4) >> 3);
            SQ(k, 0) = (T)clip3(Q0 - 2 * tc, Q0 + 2 * tc, (P1 + 2 * P0 + 2 * Q0 + 2 * Q1 + Q2 + 4) >> 3);
            SQ(k, 1) = (T)clip3(Q1 - 2 * tc, Q1 + 2 * tc, (P0 + Q0 + Q1 + Q2 + 2) >> 2);
            SQ(k, 2) = (T)clip3(Q2 - 2 * tc, Q2 + 2 * tc, (P0 + Q0 + Q1 + 3 * Q2 + 2 * Q3 + 4) >> 3);
        } else {
            int delta = (9 * (Q0 - P0) - 3 * (Q1 - P1) + 8) >> 4;
            if (iabs(delta) < tc * 10) {
                delta = clip3(-tc, tc, delta);
                SP(k, 0) = (T)clip3(0, maxv, P0 + delta);
                SQ(k, 0) = (T)clip3(0, maxv, Q0 - delta);
                if (dep) SP(k, 1) = (T)clip3(0, maxv, P1 + clip3(-(tc >> 1), tc >> 1, (((P2 + P0 + 1) >> 1) - P1 + delta) >> 1));
                if (deq) SQ(k, 1) = (T)clip3(0, maxv, Q1 + clip3(-(tc >> 1), tc >> 1, (((Q2 + Q0 + 1) >> 1) - Q1 - delta) >> 1));
            }
        }
    }
#undef SP
#undef SQ
}

// Chroma filter of 2 lines (deblocking.cpp:338-352, 538-608).
template <typename T>
__device__ void chroma_segment(T* base, ptrdiff_t lstep, ptrdiff_t istep, int tc, int maxv) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
        T* q = base + k * lstep;
        const int q0 = q[0], q1 = q[istep], p0 = q[-istep], p1 = q[-2 * istep];
        const int delta = clip3(-tc, tc, ((q0 - p0) * 4 + p1 - q1 + 4) >> 3);
        q[-istep] = (T)clip3(0, maxv, p0 + delta);
        q[0] = (T)clip3(0, maxv, q0 - delta);
    }
}

template <typename T, bool VERT>
__global__ void __launch_bounds__(256) k_deblock(FrameConst fc, const uint32_t* __restrict__ info4,
                                                 const Motion4* __restrict__ mot, T* __restrict__ y_pl,
                                                 T* __restrict__ cb_pl, T* __restrict__ cr_pl) {
    // VERT: lanes over (4-row segment, edge column x = 8e); else (edge row y = 8e, 4-col segment)
    const int ne = VERT ? (fc.W + 7) >> 3 : fc.W >> 2;  // lanes per row of work
    const int nr = VERT ? fc.H >> 2 : (fc.H + 7) >> 3;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ne * nr) return;
    const int a = id % ne, b = id / ne;
    const int x = VERT ? a * 8 : a * 4, y = VERT ? b * 4 : b * 8;
    if (x >= fc.W || y >= fc.H) return;
    if (VERT ? x == 0 : y == 0) return;  // picture edge
    const size_t gq = (size_t)(y >> 2) * fc.gw + (x >> 2);
    const uint32_t iq = info4[gq];
    if (!(iq & (VERT ? info::kEdgeV : info::kEdgeH))) return;
    const size_t gp = VERT ? gq - 1 : gq - fc.gw;
    const int bs = boundary_strength(fc, info4, mot, gp, gq, VERT ? x : y);
    if (!bs) return;
    const uint32_t ip = info4[gp];
    const int qpp = (int8_t)(ip >> info::kQpShift), qpq = (int8_t)(iq >> info::kQpShift);
    const int qpl = (qpq + qpp + 1) >> 1, bdm = 1 << (fc.bd - 8);
    const int beta = beta_prime(clip3(0, 51, qpl + 2 * fc.beta_offset_div2)) * bdm;
    const int tc = kTcPrime[clip3(0, 53, qpl + 2 * (bs - 1) + 2 * fc.tc_offset_div2)] * bdm;
    {
        T* base = y_pl + (size_t)y * fc.stride[0] + x;
        const ptrdiff_t st = fc.stride[0];
        if (VERT)
            luma_segment<T>(base, st, 1, beta, tc, fc.maxv);
        else
            luma_segment<T>(base, 1, st, beta, tc, fc.maxv);
    }
    if (bs == 2 && ((((VERT ? x : y) >> 1) & 7) == 0)) {
        for (int ci = 1; ci <= 2; ci++) {
            const int off = ci == 1 ? fc.pps_cb_qp_offset : fc.pps_cr_qp_offset;
            const int qpc = qpc_from_qpi(clip3(0, 57, ((qpq + qpp + 1) >> 1) + off));
            const int tcc = kTcPrime[clip3(0, 53, qpc + 2 + 2 * fc.tc_offset_div2)] * bdm;
            T* pl = ci == 1 ? cb_pl : cr_pl;
            const ptrdiff_t st = fc.stride[ci];
            T* base = pl + (size_t)(y >> 1) * st + (x >> 1);
            if (VERT)
                chroma_segment<T>(base, st, 1, tcc, fc.maxv);
            else
                chroma_segment<T>(base, 1, st, tcc, fc.maxv);
        }
    }
}

// ---------------------------------------------------------------------------
// SAO (sao.cpp:20-243): deblocked picture -> output picture, one lane per
// sample. Band offsets are not wrapped modulo 32 (sao.cpp:226-230).
// ---------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) k_sao(FrameConst fc, const hevcr_sao* __restrict__ sao, int c,
                                             const T* __restrict__ src, T* __restrict__ dst) {
    const int pw = c ? fc.W >> 1 : fc.W, ph = c ? fc.H >> 1 : fc.H;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= pw * ph) return;
    const int x = id % pw, y = id / pw;
    const int st = fc.stride[c];
    const int slog = fc.ctb_log2 - (c ? 1 : 0);
    const hevcr_sao& sp = sao[(y >> slog) * fc.wctb + (x >> slog)];
    const int v = src[(size_t)y * st + x];
    int out = v;
    const int type = sp.type[c];
    if (type == 2) {
        const int e = sp.eo_class[c] & 3;
        // EO neighbour offsets (sao.cpp:17-18)
        const int dxa = e == 1 ? 0 : (e == 3 ? 1 : -1), dya = e == 0 ? 0 : -1;
        const int dxb = -dxa, dyb = -dya;
        const int ax = x + dxa, ay = y + dya, bx = x + dxb, by = y + dyb;
        if (ax >= 0 && ax < pw && ay >= 0 && ay < ph && bx >= 0 && bx < pw && by >= 0 && by < ph) {
            const int av = src[(size_t)ay * st + ax], bv = src[(size_t)by * st + bx];
            const int idx = 2 + (v > av) - (v < av) + (v > bv) - (v < bv);
            out = clip3(0, fc.maxv, v + sp.offset[c][idx]);
        }
    } else if (type == 1) {
        const int bi = (v >> (fc.bd - 5)) - sp.band_pos[c];
        if (bi >= 0 && bi < 4) out = clip3(0, fc.maxv, v + sp.offset[c][bi]);
    }
    dst[(size_t)y * st + x] = (T)out;
}

// Plain 16-byte vector copy (counter calibration).
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Launchers (called by the host runtime; all on the caller's stream)
// ---------------------------------------------------------------------------

static inline unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

void launch_expand(hipStream_t s, const FrameConst& fc, const hevcr_cu* cus, uint32_t ncu, const hevcr_tu* tus,
                   uint32_t ntu, const hevcr_pu* pus, uint32_t npu, uint32_t* info4, Motion4* mot) {
    if (ncu) hipLaunchKernelGGL(k_expand_cu, dim3(ncu), dim3(64), 0, s, cus, ncu, fc, info4);
    if (ntu) hipLaunchKernelGGL(k_expand_tu, dim3(blocks_for(ntu, 256)), dim3(256), 0, s, tus, ntu, fc, info4);
    if (npu) hipLaunchKernelGGL(k_expand_pu, dim3(npu), dim3(64), 0, s, pus, npu, fc, mot);
}

template <typename T>
void launch_mc(hipStream_t s, const FrameConst& fc, const RefPlanes& refs, const McTile* tiles, uint32_t ntiles,
               const hevcr_pu* pus, void* const* planes) {
    if (!ntiles) return;
    hipLaunchKernelGGL(k_mc<T>, dim3(ntiles), dim3(64), 0, s, tiles, ntiles, pus, fc, refs, (T*)planes[0],
                       (T*)planes[1], (T*)planes[2]);
}

template <typename T>
void launch_residual(hipStream_t s, const FrameConst& fc, const ResBlock* const* lists, const uint32_t* counts,
                     const int16_t* coeffs, int16_t* resid, void* const* planes) {
    T* p0 = (T*)planes[0];
    T* p1 = (T*)planes[1];
    T* p2 = (T*)planes[2];
    if (counts[0]) hipLaunchKernelGGL((k_residual<T, 2>), dim3(blocks_for(counts[0], 64)), dim3(256), 0, s, lists[0], counts[0], coeffs, resid, fc, p0, p1, p2);
    if (counts[1]) hipLaunchKernelGGL((k_residual<T, 3>), dim3(blocks_for(counts[1], 16)), dim3(256), 0, s, lists[1], counts[1], coeffs, resid, fc, p0, p1, p2);
    if (counts[2]) hipLaunchKernelGGL((k_residual<T, 4>), dim3(blocks_for(counts[2], 4)), dim3(256), 0, s, lists[2], counts[2], coeffs, resid, fc, p0, p1, p2);
    if (counts[3]) hipLaunchKernelGGL((k_residual<T, 5>), dim3(counts[3]), dim3(256), 0, s, lists[3], counts[3], coeffs, resid, fc, p0, p1, p2);
}

template <typename T>
void launch_intra(hipStream_t s, const FrameConst& fc, const uint32_t* ctus, uint32_t nctus, const uint32_t* op_begin,
                  const IntraOp* ops, const int16_t* resid, void* const* planes) {
    if (!nctus) return;
    hipLaunchKernelGGL(k_intra<T>, dim3(nctus, 3), dim3(64), 0, s, ctus, nctus, op_begin, ops, resid, fc, (T*)planes[0],
                       (T*)planes[1], (T*)planes[2]);
}

template <typename T>
void launch_deblock(hipStream_t s, const FrameConst& fc, const uint32_t* info4, const Motion4* mot, void* const* planes) {
    const size_t nv = (size_t)((fc.W + 7) >> 3) * (fc.H >> 2), nh = (size_t)(fc.W >> 2) * ((fc.H + 7) >> 3);
    hipLaunchKernelGGL((k_deblock<T, true>), dim3(blocks_for(nv, 256)), dim3(256), 0, s, fc, info4, mot, (T*)planes[0], (T*)planes[1], (T*)planes[2]);
    hipLaunchKernelGGL((k_deblock<T, false>), dim3(blocks_for(nh, 256)), dim3(256), 0, s, fc, info4, mot, (T*)planes[0], (T*)planes[1], (T*)planes[2]);
}

template <typename T>
void launch_sao(hipStream_t s, const FrameConst& fc, const hevcr_sao* sao, void* const* src, void* const* dst) {
    for (int c = 0; c < 3; c++) {
        const size_t n = c ? (size_t)(fc.W >> 1) * (fc.H >> 1) : (size_t)fc.W * fc.H;
        hipLaunchKernelGGL(k_sao<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, fc, sao, c, (const T*)src[c], (T*)dst[c]);
    }
}

void launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes) {
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
}

#define HEVCR_INSTANTIATE(T)                                                                                          \
    template void launch_mc<T>(hipStream_t, const FrameConst&, const RefPlanes&, const McTile*, uint32_t,          \
                               const hevcr_pu*, void* const*);                                                     \
    template void launch_residual<T>(hipStream_t, const FrameConst&, const ResBlock* const*, const uint32_t*,      \
                                     const int16_t*, int16_t*, void* const*);                                      \
    template void launch_intra<T>(hipStream_t, const FrameConst&, const uint32_t*, uint32_t, const uint32_t*,      \
                                  const IntraOp*, const int16_t*, void* const*);                                   \
    template void launch_deblock<T>(hipStream_t, const FrameConst&, const uint32_t*, const Motion4*, void* const*); \
    template void launch_sao<T>(hipStream_t, const FrameConst&, const hevcr_sao*, void* const*, void* const*);

HEVCR_INSTANTIATE(uint8_t)
HEVCR_INSTANTIATE(uint16_t)

}  // namespace hevcr
