// Constant tables of H.265 section 8 used by the device kernels. Each table is
// cited where the reference decoder keeps its copy; they are rebuilt from the
// standard (closed forms where the standard gives one).
#pragma once

#include <hip/hip_runtime.h>

namespace hevcr {

// levelScale, H.265 eq. 8-309 (reference: decoding/cabac_tables.h:285)
__constant__ static const int kLevelScale[6] = {40, 45, 51, 57, 64, 72};

// 128-entry cosine table: C[j] = the H.265 transform-matrix entry for angle
// j*pi/64 (eq. 8-318), C[0] = 64 for the DC row. The NxN matrix entry of
// row m, column k is C[((2k+1) * m * (32/N)) & 127]
// (reference: decoding/transform.cpp:151-184 holds the expanded 32x32 matrix).
// A constant expression: the transforms, unrolled at compile time, take their
// coefficients as immediates.
struct CosTable {
    signed char v[128];
};
constexpr CosTable kCosC = {{
    64,  90,  90,  90,  89,  88,  87,  85,  83,  82,  80,  78,  75,  73,  70,  67,
    64,  61,  57,  54,  50,  46,  43,  38,  36,  31,  25,  22,  18,  13,  9,   4,
    0,   -4,  -9,  -13, -18, -22, -25, -31, -36, -38, -43, -46, -50, -54, -57, -61,
    -64, -67, -70, -73, -75, -78, -80, -82, -83, -85, -87, -88, -89, -90, -90, -90,
    -64, -90, -90, -90, -89, -88, -87, -85, -83, -82, -80, -78, -75, -73, -70, -67,
    -64, -61, -57, -54, -50, -46, -43, -38, -36, -31, -25, -22, -18, -13, -9,  -4,
    0,   4,   9,   13,  18,  22,  25,  31,  36,  38,  43,  46,  50,  54,  57,  61,
    64,  67,  70,  73,  75,  78,  80,  82,  83,  85,  87,  88,  89,  90,  90,  90}};

// Entry (row m, column k) of the N-point inverse-transform matrix.
__host__ __device__ constexpr int dct_coef(int n, int m, int k) { return kCosC.v[((2 * k + 1) * m * (32 / n)) & 127]; }

// DST-VII 4x4, eq. 8-317 (reference: decoding/transform.cpp:22-43); row m, column k.
struct Dst4Table {
    int v[4][4];
};
constexpr Dst4Table kDst4 = {{{29, 55, 74, 84}, {74, 74, 0, -74}, {84, -29, -74, 55}, {55, -84, 74, -29}}};

// Interpolation taps, Tables 8-1/8-2 (reference: interpolation.cpp:17-38)
__constant__ static const int kLumaTaps[4][8] = {{0, 0, 0, 64, 0, 0, 0, 0},
                                                 {-1, 4, -10, 58, 17, -5, 1, 0},
                                                 {-1, 4, -11, 40, 40, -11, 4, -1},
                                                 {0, 1, -5, 17, 58, -10, 4, -1}};
__constant__ static const int kChromaTaps[8][4] = {{0, 64, 0, 0},   {-2, 58, 10, -2}, {-4, 54, 16, -2},
                                                   {-6, 46, 28, -4}, {-4, 36, 36, -4}, {-4, 28, 46, -6},
                                                   {-2, 16, 54, -4}, {-2, 10, 58, -2}};

// The same taps as signed 16-bit pairs (low half = even tap) for v_dot2_i32_i16.
#define HEVCR_PK(a, b) ((uint32_t)(uint16_t)(int16_t)(a) | ((uint32_t)(uint16_t)(int16_t)(b) << 16))
__constant__ static const uint32_t kLumaPairs[4][4] = {
    {HEVCR_PK(0, 0), HEVCR_PK(0, 64), HEVCR_PK(0, 0), HEVCR_PK(0, 0)},
    {HEVCR_PK(-1, 4), HEVCR_PK(-10, 58), HEVCR_PK(17, -5), HEVCR_PK(1, 0)},
    {HEVCR_PK(-1, 4), HEVCR_PK(-11, 40), HEVCR_PK(40, -11), HEVCR_PK(4, -1)},
    {HEVCR_PK(0, 1), HEVCR_PK(-5, 17), HEVCR_PK(58, -10), HEVCR_PK(4, -1)}};
__constant__ static const uint32_t kChromaPairs[8][2] = {
    {HEVCR_PK(0, 64), HEVCR_PK(0, 0)},   {HEVCR_PK(-2, 58), HEVCR_PK(10, -2)}, {HEVCR_PK(-4, 54), HEVCR_PK(16, -2)},
    {HEVCR_PK(-6, 46), HEVCR_PK(28, -4)}, {HEVCR_PK(-4, 36), HEVCR_PK(36, -4)}, {HEVCR_PK(-4, 28), HEVCR_PK(46, -6)},
    {HEVCR_PK(-2, 16), HEVCR_PK(54, -4)}, {HEVCR_PK(-2, 10), HEVCR_PK(58, -2)}};
#undef HEVCR_PK

// The same taps as signed bytes (byte i = tap i) for v_dot4_i32_i8: luma taps
// 0..3 and 4..7, chroma taps 0..3.
#define HEVCR_B4(a, b, c, d) \
    ((uint32_t)(uint8_t)(int8_t)(a) | ((uint32_t)(uint8_t)(int8_t)(b) << 8) | ((uint32_t)(uint8_t)(int8_t)(c) << 16) | \
     ((uint32_t)(uint8_t)(int8_t)(d) << 24))
__constant__ static const uint32_t kLumaI8[4][2] = {{HEVCR_B4(0, 0, 0, 64), HEVCR_B4(0, 0, 0, 0)},
                                                    {HEVCR_B4(-1, 4, -10, 58), HEVCR_B4(17, -5, 1, 0)},
                                                    {HEVCR_B4(-1, 4, -11, 40), HEVCR_B4(40, -11, 4, -1)},
                                                    {HEVCR_B4(0, 1, -5, 17), HEVCR_B4(58, -10, 4, -1)}};
__constant__ static const uint32_t kChromaI8[8] = {HEVCR_B4(0, 64, 0, 0),   HEVCR_B4(-2, 58, 10, -2), HEVCR_B4(-4, 54, 16, -2),
                                                   HEVCR_B4(-6, 46, 28, -4), HEVCR_B4(-4, 36, 36, -4), HEVCR_B4(-4, 28, 46, -6),
                                                   HEVCR_B4(-2, 16, 54, -4), HEVCR_B4(-2, 10, 58, -2)};
#undef HEVCR_B4

// tC', Table 8-12 (reference: filters/deblocking.cpp:26-33)
__constant__ static const int kTcPrime[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,
                                              1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3,  3,  3,  4,
                                              4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};

// beta', Table 8-12: 0 below 16, Q-10 up to 28, 2Q-38 above
// (reference: filters/deblocking.cpp:16-23)
__device__ __forceinline__ int beta_prime(int q) { return q < 16 ? 0 : (q <= 28 ? q - 10 : 2 * q - 38); }

// QpC from qPi, Table 8-10 (reference: cabac_tables.h:275-282, deblocking.cpp:35-42)
__device__ __forceinline__ int qpc_from_qpi(int qpi) {
    if (qpi < 30) return qpi;
    if (qpi > 42) return qpi - 6;
    const int mid = qpi - 30;  // 29,30,31,32,33,33,34,34,35,35,36,36,37
    return mid < 4 ? 29 + mid : 33 + ((mid - 4) >> 1);
}

}  // namespace hevcr
