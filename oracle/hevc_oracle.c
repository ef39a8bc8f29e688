/*
 * hevc_oracle.c — CPU restatement of the reference decoder's reconstruction
 * path (privaloops/hevc.js, src/decoding + src/filters).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the HIP engine is compared
 * against (and the CPU baseline leg of the benchmark). The product never
 * links, loads or calls it.
 *
 * Written from the behaviour of the reference functions (cited per function)
 * and H.265 section 8, not from their text. Pinned by the golden vectors that
 * the reference's own code produces through oracle/ref_harness.cpp, by the
 * reference's fixture MD5s (through that harness), and by closed-form checks
 * in tests/test_oracle_*.py.
 *
 * Scope matches include/hevcr_syntax.h: 4:2:0, one slice, no tiles, flat
 * scaling lists, no PCM / transquant bypass.
 */
#include "hevc_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * Constant tables, rebuilt from H.265 (each cited where the reference keeps it)
 * ------------------------------------------------------------------------- */

/* levelScale, H.265 eq. 8-309; reference: decoding/cabac_tables.h:285 */
static const int kLevelScale[6] = {40, 45, 51, 57, 64, 72};

/* QpC as a function of qPi (H.265 Table 8-10) for qPi in 30..42; below 30
 * QpC == qPi, above 42 QpC == qPi - 6. Reference: cabac_tables.h:275-282,
 * filters/deblocking.cpp:35-42. */
static int qpc_from_qpi(int qpi) {
    static const int mid[13] = {29, 30, 31, 32, 33, 33, 34, 34, 35, 35, 36, 36, 37};
    if (qpi < 30) return qpi;
    if (qpi <= 42) return mid[qpi - 30];
    return qpi - 6;
}

/* cos(j*pi/64) scaled (H.265 eq. 8-318 matrix entries), j = 0..31, with the
 * DC row value 64 at j = 0. The 32x32 matrix is T[k][n] = a((2n+1)k mod 128)
 * with the cosine's quadrant symmetry. Reference: decoding/transform.cpp:151-184. */
static const int kCosA[32] = {64, 90, 90, 90, 89, 88, 87, 85, 83, 82, 80, 78, 75, 73, 70, 67,
                              64, 61, 57, 54, 50, 46, 43, 38, 36, 31, 25, 22, 18, 13, 9,  4};

static int g_tm[32][32];
static int g_tm_ready;

static int cos_entry(int j) {
    j &= 127;
    if (j == 0) return 64;
    if (j < 32) return kCosA[j];
    if (j == 32) return 0;
    if (j < 64) return -kCosA[64 - j];
    if (j == 64) return -64; /* only reached by k == 0 rows, never used (k=0 -> j=0) */
    if (j < 96) return -kCosA[j - 64];
    if (j == 96) return 0;
    return kCosA[128 - j];
}

static void init_tm(void) {
    if (g_tm_ready) return;
    for (int k = 0; k < 32; k++)
        for (int n = 0; n < 32; n++) g_tm[k][n] = k == 0 ? 64 : cos_entry((2 * n + 1) * k);
    g_tm_ready = 1;
}

/* DST-VII 4x4 (H.265 eq. 8-317); reference: transform.cpp:22-43 */
static const int kDst[4][4] = {{29, 55, 74, 84}, {74, 74, 0, -74}, {84, -29, -74, 55}, {55, -84, 74, -29}};

/* intraPredAngle / invAngle (H.265 Tables 8-4, 8-5); reference: intra_prediction.cpp:13-29 */
static const int kAngle[35] = {0,   0,   32,  26,  21,  17,  13,  9,  5,  2,  0,  -2,
                               -5,  -9,  -13, -17, -21, -26, -32, -26, -21, -17, -13, -9,
                               -5,  -2,  0,   2,   5,   9,   13,  17,  21,  26,  32};
static const int kInvAngle[35] = {0,    0,    256,  315,  390,  482,  630, 910, 1638, 4096, 0,   4096,
                                  1638, 910,  630,  482,  390,  315,  256, 315, 390,  482,  630, 910,
                                  1638, 4096, 0,    4096, 1638, 910,  630, 482, 390,  315,  256};

/* Interpolation taps (H.265 Tables 8-1/8-2); reference: interpolation.cpp:17-38 */
static const int kLumaTap[4][8] = {{0, 0, 0, 64, 0, 0, 0, 0},
                                   {-1, 4, -10, 58, 17, -5, 1, 0},
                                   {-1, 4, -11, 40, 40, -11, 4, -1},
                                   {0, 1, -5, 17, 58, -10, 4, -1}};
static const int kChromaTap[8][4] = {{0, 64, 0, 0},    {-2, 58, 10, -2}, {-4, 54, 16, -2}, {-6, 46, 28, -4},
                                     {-4, 36, 36, -4}, {-4, 28, 46, -6}, {-2, 16, 54, -4}, {-2, 10, 58, -2}};

/* beta' / tC' (H.265 Table 8-12); reference: deblocking.cpp:16-33.
 * beta': 0 below Q=16, Q-10 for 16..28, 2Q-38 for 29..51. */
static int beta_prime(int q) {
    if (q < 16) return 0;
    if (q <= 28) return q - 10;
    return 2 * q - 38;
}
static int tc_prime(int q) {
    static const int t[54] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1,  1,  1,  1,  1,  1,  1,  1, 1,
                              2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7, 8, 9, 10, 11, 13, 14, 16, 18, 20, 22, 24};
    return t[q];
}

static inline int clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int iabs(int v) { return v < 0 ? -v : v; }

/* ---------------------------------------------------------------------------
 * Dequantisation + inverse transform
 * ------------------------------------------------------------------------- */

/* Scaling factor m of coefficient (x, y) of an NxN block, transform.cpp:308-356:
 * 16 without scaling lists; else the 4x4 list, or the 8x8 base list replicated
 * over r x r squares (r = N / 8) with the DC factor at (0, 0) for N >= 16
 * (a stored 0 reads as 16 there). matrixId = cIdx + 3 for non-intra CUs. */
static int scale_m(const hevcr_scaling *sl, int log2, int cidx, int intra, int x, int y) {
    if (!sl) return 16;
    const int size_id = log2 - 2, matrix = cidx + (intra ? 0 : 3);
    if (size_id == 0) return sl->list[0][matrix][y * 4 + x];
    const int r = (1 << log2) / 8;
    if (size_id >= 2 && x == 0 && y == 0) {
        const int dc = sl->dc[size_id - 2][matrix];
        return dc ? dc : 16;
    }
    return sl->list[size_id][matrix][(y / r) * 8 + x / r];
}

/* perform_dequant, transform.cpp:293-368. The int64 scaled value is narrowed
 * to int before the int16 clip (transform.cpp:362-365). */
static void dequant(const int16_t *c, int16_t *d, int log2, int qp, int bd, const hevcr_scaling *sl, int cidx,
                    int intra) {
    const int n = 1 << log2, bdshift = bd + log2 - 5;
    const int64_t add = bdshift > 0 ? (int64_t)1 << (bdshift - 1) : 0;
    const int64_t scale = kLevelScale[qp % 6];
    for (int i = 0; i < n * n; i++) {
        if (!c[i]) {
            d[i] = 0;
            continue;
        }
        int64_t v = (int64_t)c[i] * scale_m(sl, log2, cidx, intra, i % n, i / n) * scale;
        v = (v * ((int64_t)1 << (qp / 6)) + add) >> bdshift;
        d[i] = (int16_t)clip3(-32768, 32767, (int)v);
    }
}

/* One 1-D inverse pass over the columns of an NxN block: out[k][j] =
 * clip16((sum_m M[m][k] * in[m][j] + rnd) >> shift). Integer matrix product,
 * identical to the reference's partial butterflies (transform.cpp:22-233). */
static void inv_pass(const int16_t *in, int16_t *out, int log2, int dst, int shift) {
    const int n = 1 << log2, step = 32 >> log2, add = 1 << (shift - 1);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
            int s = 0;
            for (int m = 0; m < n; m++) s += (dst ? kDst[m][k] : g_tm[m * step][k]) * in[m * n + j];
            out[k * n + j] = (int16_t)clip3(-32768, 32767, (s + add) >> shift);
        }
}

/* perform_transform_inverse, transform.cpp:374-394 (+ inverse_transform_2d 240-287):
 * column pass (shift 7), then row pass (shift 20 - bd); transform skip uses the
 * reference's (d + 2^(14-bd)) >> (15-bd). */
static void inv_transform(const int16_t *d, int16_t *r, int log2, int cidx, int intra, int tskip, int bd) {
    const int n = 1 << log2;
    if (tskip) {
        const int sh = 15 - bd > 0 ? 15 - bd : 0, add = sh > 0 ? 1 << (sh - 1) : 0;
        for (int i = 0; i < n * n; i++) r[i] = (int16_t)((d[i] + add) >> sh);
        return;
    }
    init_tm();
    const int dst = intra && cidx == 0 && log2 == 2;
    int16_t t1[32 * 32], t2[32 * 32] = {0};
    inv_pass(d, t1, log2, dst, 7);
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) t2[y * n + x] = t1[x * n + y];
    inv_pass(t2, t1, log2, dst, 20 - bd);
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) r[y * n + x] = t1[x * n + y];
}

int oracle_residual_block(int log2, int cidx, int is_intra, int qp_prime, int bit_depth, int tskip,
                          const int16_t *coeffs, int16_t *out) {
    if (log2 < 2 || log2 > 5 || !coeffs || !out || qp_prime < 0) return -1;
    int16_t d[32 * 32];
    dequant(coeffs, d, log2, qp_prime, bit_depth, NULL, cidx, is_intra);
    inv_transform(d, out, log2, cidx, is_intra, tskip, bit_depth);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Frame state
 * ------------------------------------------------------------------------- */

typedef struct {
    int16_t mv[2][2];
    int8_t ref[2];
    uint8_t flags;
} Motion;

/* Slice-segment parameters the operators read (slice_header.h:16-108), one
 * set per segment; a picture without segment records is one segment, p's. */
typedef struct {
    int type, cb_off, cr_off, dbk_off, beta, tc, lf_across, wp, lwd, cwd;
    const hevcr_weight (*w)[HEVCR_MAX_REFS];
} SegP;

typedef struct {
    const hevcr_frame_syntax *s;
    const hevcr_frame_params *p;
    SegP *seg;
    const uint8_t *ctb_seg; /* segment per CTB (raster), NULL: all 0 */
    int *tile;              /* tile per CTB (raster), NULL: one tile */
    int lf_tiles;           /* loop_filter_across_tiles_enabled_flag */
    uint8_t *cu_flags;      /* per 8x8: HEVCR_CU_PCM (as the reference's is_pcm), HEVCR_CU_BYPASS */
    int W, H, bd, maxv, ctb_log2, ctb, wctb, hctb;
    uint16_t *pl[3];
    int pw[3], ph[3];
    const uint16_t *const *refs;
    int nrefs;
    int gw, gh; /* 4x4 grid */
    uint8_t *cbf, *log2tu, *ev, *eh, *mode_y, *mode_c;
    uint8_t *cu_pred; /* per 8x8 */
    int8_t *cu_qp;
    Motion *mot;
} Ctx;

static int ctx_init(Ctx *c, const hevcr_frame_syntax *s) {
    memset(c, 0, sizeof(*c));
    c->s = s;
    c->p = &s->p;
    const hevcr_frame_params *p = &s->p;
    if (p->width < 8 || p->height < 8 || (p->width & 7) || (p->height & 7)) return -10;
    if (p->bit_depth < 8 || p->bit_depth > 12) return -11;
    if (p->ctb_log2 < 4 || p->ctb_log2 > 6 || p->min_cb_log2 < 3 || p->min_cb_log2 > p->ctb_log2) return -12;
    c->W = p->width;
    c->H = p->height;
    c->bd = p->bit_depth;
    c->maxv = (1 << c->bd) - 1;
    c->ctb_log2 = p->ctb_log2;
    c->ctb = 1 << p->ctb_log2;
    c->wctb = (c->W + c->ctb - 1) / c->ctb;
    c->hctb = (c->H + c->ctb - 1) / c->ctb;
    c->pw[0] = c->W;
    c->ph[0] = c->H;
    c->pw[1] = c->pw[2] = c->W / 2;
    c->ph[1] = c->ph[2] = c->H / 2;
    c->gw = c->W / 4;
    c->gh = c->H / 4;
    const size_t g = (size_t)c->gw * c->gh, g8 = (size_t)(c->W / 8) * (c->H / 8);
    c->cbf = calloc(g, 1);
    c->log2tu = malloc(g);
    c->ev = calloc(g, 1);
    c->eh = calloc(g, 1);
    c->mode_y = malloc(g);
    c->mode_c = calloc(g, 1);
    c->cu_pred = malloc(g8);
    c->cu_qp = malloc(g8);
    c->mot = calloc(g, sizeof(Motion));
    for (int i = 0; i < 3; i++) c->pl[i] = calloc((size_t)c->pw[i] * c->ph[i], 2);
    if (!c->cbf || !c->log2tu || !c->ev || !c->eh || !c->mode_y || !c->mode_c || !c->cu_pred || !c->cu_qp ||
        !c->mot || !c->pl[0] || !c->pl[1] || !c->pl[2])
        return -13;
    /* Per-picture grid defaults, decoder.cpp:117-149 */
    memset(c->log2tu, c->ctb_log2, g);
    memset(c->mode_y, 1, g);
    memset(c->cu_pred, HEVCR_MODE_INTRA, g8);
    memset(c->cu_qp, 26, g8);
    for (size_t i = 0; i < g; i++) c->mot[i].ref[0] = c->mot[i].ref[1] = -1;
    c->cu_flags = calloc(g8, 1);
    const int nseg = s->slices ? (int)s->num_slices : 1;
    c->seg = calloc((size_t)nseg, sizeof(SegP));
    if (!c->cu_flags || !c->seg) return -13;
    if (s->slices && (!s->ctb_slice || s->num_slices < 1 || s->num_slices > HEVCR_MAX_SLICES)) return -14;
    for (int k = 0; k < nseg; k++) {
        SegP *q = &c->seg[k];
        if (s->slices) {
            const hevcr_slice *r = &s->slices[k];
            q->type = r->slice_type;
            q->cb_off = r->slice_cb_qp_offset;
            q->cr_off = r->slice_cr_qp_offset;
            q->dbk_off = r->deblocking_disabled;
            q->beta = r->beta_offset_div2;
            q->tc = r->tc_offset_div2;
            q->lf_across = r->loop_filter_across_slices;
            q->wp = r->weighted_pred;
            q->lwd = r->luma_log2_wd;
            q->cwd = r->chroma_log2_wd;
            q->w = r->weights;
        } else {
            q->type = p->slice_type;
            q->cb_off = p->slice_cb_qp_offset;
            q->cr_off = p->slice_cr_qp_offset;
            q->dbk_off = p->deblocking_disabled;
            q->beta = p->beta_offset_div2;
            q->tc = p->tc_offset_div2;
            q->lf_across = 1;
            q->wp = p->weighted_pred;
            q->lwd = p->luma_log2_wd;
            q->cwd = p->chroma_log2_wd;
            q->w = p->weights;
        }
    }
    c->ctb_seg = s->slices ? s->ctb_slice : NULL;
    for (int a = 0; a < c->wctb * c->hctb && c->ctb_seg; a++)
        if (c->ctb_seg[a] >= nseg) return -15;
    c->lf_tiles = 1;
    if (s->tiles) { /* TileId per CTB from the column / row boundaries (pps.cpp derive_tile_scan) */
        const hevcr_tiles *t = s->tiles;
        if (t->cols < 1 || t->cols > HEVCR_MAX_TILE_COLS || t->rows < 1 || t->rows > HEVCR_MAX_TILE_ROWS) return -16;
        c->tile = malloc((size_t)c->wctb * c->hctb * sizeof(int));
        if (!c->tile) return -13;
        int y0 = 0;
        for (int r = 0; r < t->rows; r++) {
            int x0 = 0;
            for (int q = 0; q < t->cols; q++) {
                for (int y = y0; y < y0 + t->row_height[r] && y < c->hctb; y++)
                    for (int x = x0; x < x0 + t->col_width[q] && x < c->wctb; x++) c->tile[y * c->wctb + x] = r * t->cols + q;
                x0 += t->col_width[q];
            }
            if (x0 != c->wctb) return -16;
            y0 += t->row_height[r];
        }
        if (y0 != c->hctb) return -16;
        c->lf_tiles = t->loop_filter_across_tiles;
    }
    return 0;
}

/* CTB raster address of a luma position, its slice segment and tile. */
static int ctb_at(const Ctx *c, int x, int y) { return (y >> c->ctb_log2) * c->wctb + (x >> c->ctb_log2); }
static int seg_of(const Ctx *c, int a) { return c->ctb_seg ? c->ctb_seg[a] : 0; }
static int tile_of(const Ctx *c, int a) { return c->tile ? c->tile[a] : 0; }

static void ctx_free(Ctx *c) {
    free(c->cbf);
    free(c->log2tu);
    free(c->ev);
    free(c->eh);
    free(c->mode_y);
    free(c->mode_c);
    free(c->cu_pred);
    free(c->cu_qp);
    free(c->mot);
    free(c->cu_flags);
    free(c->seg);
    free(c->tile);
    for (int i = 0; i < 3; i++) free(c->pl[i]);
}

#define G4(c, x, y) ((size_t)((y) >> 2) * (c)->gw + ((x) >> 2))
#define G8(c, x, y) ((size_t)((y) >> 3) * ((c)->W >> 3) + ((x) >> 3))

/* ---------------------------------------------------------------------------
 * Intra prediction (intra_prediction.cpp:44-481)
 * ------------------------------------------------------------------------- */

static uint32_t zscan(int bx, int by) {
    uint32_t z = 0;
    for (int i = 0; i < 8; i++) z |= (uint32_t)(((bx >> i) & 1) << (2 * i)) | (uint32_t)(((by >> i) & 1) << (2 * i + 1));
    return z;
}

/* Availability of a neighbouring sample (component coordinates) for the block
 * whose top-left luma position is (x0,y0): inside the picture and either in an
 * earlier CTB in raster order of the same slice segment and tile, or in the
 * same CTB at a smaller z-scan address of 4x4 luma blocks
 * (intra_prediction.cpp:102-133). */
static int avail(const Ctx *c, int cidx, int x0, int y0, int rx, int ry) {
    if (rx < 0 || ry < 0 || rx >= c->pw[cidx] || ry >= c->ph[cidx]) return 0;
    const int sh = cidx ? 1 : 0;
    const int lx = rx << sh, ly = ry << sh;
    const int cx = x0 >> c->ctb_log2, cy = y0 >> c->ctb_log2;
    const int nx = lx >> c->ctb_log2, ny = ly >> c->ctb_log2;
    if (nx != cx || ny != cy) {
        const int ra = ny * c->wctb + nx, ca = cy * c->wctb + cx;
        return ra < ca && seg_of(c, ra) == seg_of(c, ca) && tile_of(c, ra) == tile_of(c, ca);
    }
    const int ox = cx << c->ctb_log2, oy = cy << c->ctb_log2;
    return zscan((lx - ox) >> 2, (ly - oy) >> 2) < zscan((x0 - ox) >> 2, (y0 - oy) >> 2);
}

/* Builds top[0..2N] (top[0] = corner, top[1+i] = p[i][-1]) and left[0..2N]
 * (left[1+i] = p[-1][i]) with substitution (intra_prediction.cpp:44-211). */
static void intra_refs(const Ctx *c, int cidx, int x0, int y0, int n, int *top, int *left) {
    const int xc = cidx ? x0 >> 1 : x0, yc = cidx ? y0 >> 1 : y0;
    const uint16_t *pl = c->pl[cidx];
    const int st = c->pw[cidx];
    /* linear order: left column bottom-up (2N), corner, top row left-right (2N) */
    int v[4 * 64 + 1], a[4 * 64 + 1];
    const int tot = 4 * n + 1;
    for (int k = 0; k < 2 * n; k++) {
        const int ry = yc + 2 * n - 1 - k, rx = xc - 1;
        a[k] = avail(c, cidx, x0, y0, rx, ry);
        v[k] = a[k] ? pl[ry * st + rx] : 0;
    }
    a[2 * n] = avail(c, cidx, x0, y0, xc - 1, yc - 1);
    v[2 * n] = a[2 * n] ? pl[(yc - 1) * st + xc - 1] : 0;
    for (int k = 0; k < 2 * n; k++) {
        const int rx = xc + k, ry = yc - 1;
        a[2 * n + 1 + k] = avail(c, cidx, x0, y0, rx, ry);
        v[2 * n + 1 + k] = a[2 * n + 1 + k] ? pl[ry * st + rx] : 0;
    }
    int first = -1;
    for (int i = 0; i < tot; i++)
        if (a[i]) {
            first = i;
            break;
        }
    if (first < 0) {
        for (int i = 0; i < tot; i++) v[i] = 1 << (c->bd - 1);
    } else {
        for (int i = 0; i < first; i++) v[i] = v[first];
        for (int i = first + 1; i < tot; i++)
            if (!a[i]) v[i] = v[i - 1];
    }
    top[0] = left[0] = v[2 * n];
    for (int i = 0; i < 2 * n; i++) {
        top[1 + i] = v[2 * n + 1 + i];
        left[1 + i] = v[2 * n - 1 - i];
    }
}

/* [1 2 1] smoothing or the 32x32 bilinear (strong) smoothing of one array,
 * corner excluded (intra_prediction.cpp:232-257). */
static void smooth(int *r, int n, int bilinear) {
    int f[2 * 64 + 1];
    f[0] = r[0];
    f[2 * n] = r[2 * n];
    for (int i = 1; i < 2 * n; i++)
        f[i] = bilinear ? ((2 * n - i) * r[0] + i * r[2 * n] + n) / (2 * n) : (r[i - 1] + 2 * r[i] + r[i + 1] + 2) >> 2;
    memcpy(r, f, sizeof(int) * (size_t)(2 * n + 1));
}

static void intra_predict(const Ctx *c, int cidx, int x0, int y0, int log2, int mode, int16_t *pred) {
    const int n = 1 << log2;
    int top[2 * 64 + 1], left[2 * 64 + 1];
    intra_refs(c, cidx, x0, y0, n, top, left);

    /* Filtering decision (intra_prediction.cpp:217-230, 436-471): luma only in
     * 4:2:0, never for DC or 4x4, threshold on min(|m-26|,|m-10|). */
    int filt = 0;
    if (!c->p->intra_smoothing_disabled && cidx == 0 && mode != 1 && log2 > 2) {
        const int d = iabs(mode - 26) < iabs(mode - 10) ? iabs(mode - 26) : iabs(mode - 10);
        const int thr = log2 == 3 ? 7 : (log2 == 4 ? 1 : 0);
        filt = d > thr;
    }
    if (filt) {
        int bil = 0;
        if (c->p->strong_intra_smoothing && n == 32) {
            const int t = 1 << (c->bd - 5);
            bil = iabs(top[0] + top[2 * n] - 2 * top[n]) < t && iabs(left[0] + left[2 * n] - 2 * left[n]) < t;
        }
        const int corner = (left[1] + 2 * top[0] + top[1] + 2) >> 2;
        smooth(top, n, bil);
        smooth(left, n, bil);
        if (!bil) top[0] = left[0] = corner;
    }

    if (mode == 0) { /* planar, intra_prediction.cpp:263-289 */
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++)
                pred[y * n + x] = (int16_t)(((n - 1 - x) * left[y + 1] + (x + 1) * top[n + 1] + (n - 1 - y) * top[x + 1] +
                                             (y + 1) * left[n + 1] + n) >> (log2 + 1));
        return;
    }
    if (mode == 1) { /* DC, intra_prediction.cpp:295-317 */
        int sum = n;
        for (int i = 1; i <= n; i++) sum += top[i] + left[i];
        const int dc = sum >> (log2 + 1);
        for (int i = 0; i < n * n; i++) pred[i] = (int16_t)dc;
        if (cidx == 0 && n < 32) {
            pred[0] = (int16_t)((top[1] + left[1] + 2 * dc + 2) >> 2);
            for (int x = 1; x < n; x++) pred[x] = (int16_t)((top[x + 1] + 3 * dc + 2) >> 2);
            for (int y = 1; y < n; y++) pred[y * n] = (int16_t)((left[y + 1] + 3 * dc + 2) >> 2);
        }
        return;
    }
    /* Angular, intra_prediction.cpp:323-422 */
    const int ang = kAngle[mode], vert = mode >= 18;
    const int *mainr = vert ? top : left, *side = vert ? left : top;
    int ext[3 * 64 + 1];
    int *ref = ext + 64;
    for (int i = 0; i <= 2 * n; i++) ref[i] = mainr[i];
    if (ang < 0) {
        /* Projection of the side array; the reference always projects at least
         * index -1 and clamps the source index to 2N (intra_prediction.cpp:346-368). */
        const int num = (-(n * ang) + 31) >> 5;
        for (int i = 1; i <= num; i++) {
            int si = (i * kInvAngle[mode] + 128) >> 8;
            if (si > 2 * n) si = 2 * n;
            ref[-i] = side[si];
        }
    }
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
            const int pos = vert ? y : x, oth = vert ? x : y;
            const int idx = ((pos + 1) * ang) >> 5, fr = ((pos + 1) * ang) & 31;
            const int r = oth + 1 + idx;
            pred[y * n + x] = (int16_t)(fr ? ((32 - fr) * ref[r] + fr * ref[r + 1] + 16) >> 5 : ref[r]);
        }
    if (cidx == 0 && n < 32) {
        if (mode == 26)
            for (int y = 0; y < n; y++) pred[y * n] = (int16_t)clip3(0, c->maxv, pred[y * n] + ((left[y + 1] - left[0]) >> 1));
        else if (mode == 10)
            for (int x = 0; x < n; x++) pred[x] = (int16_t)clip3(0, c->maxv, pred[x] + ((top[x + 1] - top[0]) >> 1));
    }
}

/* ---------------------------------------------------------------------------
 * Inter prediction (interpolation.cpp:45-411)
 * ------------------------------------------------------------------------- */

static inline int ref_at(const uint16_t *pl, int w, int h, int x, int y) {
    x = clip3(0, w - 1, x);
    y = clip3(0, h - 1, y);
    return pl[(size_t)y * w + x];
}

/* Separable interpolation of one list: luma 8-tap quarter-pel, chroma 4-tap
 * eighth-pel; intermediates stored as int16 (interpolation.cpp:79-115, 162-198);
 * reference coordinates clamped into the picture (edge replication). */
static void interp(const Ctx *c, int cidx, const uint16_t *pl, int xint, int yint, int fx, int fy, int w, int h,
                   int16_t *out) {
    const int pw = c->pw[cidx], ph = c->ph[cidx];
    const int ntap = cidx ? 4 : 8, half = cidx ? 1 : 3;
    const int *tx = cidx ? kChromaTap[fx] : kLumaTap[fx];
    const int *ty = cidx ? kChromaTap[fy] : kLumaTap[fy];
    const int sh1 = c->bd - 8 < 4 ? c->bd - 8 : 4, sh3 = 14 - c->bd > 2 ? 14 - c->bd : 2;
    if (!fx && !fy) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) out[y * w + x] = (int16_t)(ref_at(pl, pw, ph, xint + x, yint + y) << sh3);
        return;
    }
    if (!fy) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int s = 0;
                for (int k = 0; k < ntap; k++) s += tx[k] * ref_at(pl, pw, ph, xint + x + k - half, yint + y);
                out[y * w + x] = (int16_t)(s >> sh1);
            }
        return;
    }
    if (!fx) {
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int s = 0;
                for (int k = 0; k < ntap; k++) s += ty[k] * ref_at(pl, pw, ph, xint + x, yint + y + k - half);
                out[y * w + x] = (int16_t)(s >> sh1);
            }
        return;
    }
    int16_t tmp[64 * 71];
    for (int y = 0; y < h + ntap - 1; y++)
        for (int x = 0; x < w; x++) {
            int s = 0;
            for (int k = 0; k < ntap; k++) s += tx[k] * ref_at(pl, pw, ph, xint + x + k - half, yint + y - half);
            tmp[y * w + x] = (int16_t)(s >> sh1);
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int s = 0;
            for (int k = 0; k < ntap; k++) s += ty[k] * tmp[(y + k) * w + x];
            out[y * w + x] = (int16_t)(s >> 6);
        }
}

/* perform_inter_prediction for one PU and component, with default
 * (interpolation.cpp:215-240) or explicit (246-300) weighting; the
 * prediction is written into the current picture (coding_tree.cpp:888-922). */
static int mc_component(Ctx *c, const hevcr_pu *pu, int cidx) {
    const hevcr_frame_params *p = c->p;
    const SegP *sg = &c->seg[seg_of(c, ctb_at(c, pu->x, pu->y))];
    const int w = cidx ? pu->w >> 1 : pu->w, h = cidx ? pu->h >> 1 : pu->h;
    int16_t pr[2][64 * 64];
    for (int l = 0; l < 2; l++) {
        if (!((pu->pred_flags >> l) & 1)) continue;
        const int ri = pu->ref_idx[l];
        if (ri < 0 || ri >= p->num_ref[l]) return -20;
        const int slot = p->ref_slot[l][ri];
        if (slot < 0 || slot >= c->nrefs) return -21;
        const uint16_t *pl = c->refs[slot * 3 + cidx];
        const int mvx = pu->mv[l][0], mvy = pu->mv[l][1];
        if (cidx == 0)
            interp(c, 0, pl, pu->x + (mvx >> 2), pu->y + (mvy >> 2), mvx & 3, mvy & 3, w, h, pr[l]);
        else
            interp(c, cidx, pl, (pu->x >> 1) + (mvx >> 3), (pu->y >> 1) + (mvy >> 3), mvx & 7, mvy & 7, w, h, pr[l]);
    }
    const int f0 = pu->pred_flags & 1, f1 = (pu->pred_flags >> 1) & 1;
    if (!f0 && !f1) return -22;
    const int sh1 = 14 - c->bd > 2 ? 14 - c->bd : 2;
    uint16_t *dst = c->pl[cidx];
    const int st = c->pw[cidx], xo = cidx ? pu->x >> 1 : pu->x, yo = cidx ? pu->y >> 1 : pu->y;
    const int wp = sg->wp && (sg->type == HEVCR_SLICE_P || sg->type == HEVCR_SLICE_B);
    int lwd = 0, w0 = 0, w1 = 0, o0 = 0, o1 = 0;
    if (wp) {
        lwd = (cidx ? sg->cwd : sg->lwd) + sh1;
        const int r0 = pu->ref_idx[0] >= 0 ? pu->ref_idx[0] : 0, r1 = pu->ref_idx[1] >= 0 ? pu->ref_idx[1] : 0;
        w0 = sg->w[0][r0].w[cidx];
        w1 = sg->w[1][r1].w[cidx];
        o0 = sg->w[0][r0].o[cidx] * (1 << (c->bd - 8));
        o1 = sg->w[1][r1].o[cidx] * (1 << (c->bd - 8));
    }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int a = pr[0][y * w + x], b = pr[1][y * w + x];
            int v;
            if (!wp) {
                if (f0 && f1)
                    v = (a + b + (1 << (sh1 + 1 - 1))) >> (sh1 + 1);
                else
                    v = ((f0 ? a : b) + (1 << (sh1 - 1))) >> sh1;
            } else if (f0 && f1) {
                v = (a * w0 + b * w1 + (o0 + o1 + 1) * (1 << lwd)) >> (lwd + 1);
            } else {
                v = f0 ? ((a * w0 + (1 << (lwd - 1))) >> lwd) + o0 : ((b * w1 + (1 << (lwd - 1))) >> lwd) + o1;
            }
            dst[(size_t)(yo + y) * st + xo + x] = (uint16_t)(int16_t)clip3(0, c->maxv, v);
        }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Reconstruction in decoding order (coding_tree.cpp:749-1419)
 * ------------------------------------------------------------------------- */

static void put_block(Ctx *c, int cidx, int xl, int yl, int log2, const int16_t *pred, const int16_t *res) {
    const int n = 1 << log2, xc = cidx ? xl >> 1 : xl, yc = cidx ? yl >> 1 : yl;
    uint16_t *pl = c->pl[cidx];
    const int st = c->pw[cidx];
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            if (xc + i >= c->pw[cidx] || yc + j >= c->ph[cidx]) continue;
            pl[(size_t)(yc + j) * st + xc + i] = (uint16_t)clip3(0, c->maxv, pred[j * n + i] + (res ? res[j * n + i] : 0));
        }
}

static void read_block(const Ctx *c, int cidx, int xl, int yl, int log2, int16_t *pred) {
    const int n = 1 << log2, xc = cidx ? xl >> 1 : xl, yc = cidx ? yl >> 1 : yl, st = c->pw[cidx];
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) pred[j * n + i] = (int16_t)c->pl[cidx][(size_t)(yc + j) * st + xc + i];
}

/* Chroma qP' from the CU's QpY and its slice segment's offsets
 * (coding_tree.cpp:1366-1377). */
static int chroma_qp(const Ctx *c, int qpy, int cidx, const SegP *sg) {
    const hevcr_frame_params *p = c->p;
    const int qbd = 6 * (c->bd - 8);
    const int off = cidx == 1 ? p->pps_cb_qp_offset + sg->cb_off : p->pps_cr_qp_offset + sg->cr_off;
    const int qpi = clip3(-qbd, 57, qpy + off);
    return (qpi < 0 ? qpi : qpc_from_qpi(qpi)) + qbd;
}

static int do_tu(Ctx *c, const hevcr_cu *cu, const hevcr_tu *t) {
    const hevcr_frame_syntax *s = c->s;
    const int intra = cu->pred_mode == HEVCR_MODE_INTRA, n = 1 << t->log2_size;
    const int bypass = (cu->flags & HEVCR_CU_BYPASS) != 0; /* residual = levels (coding_tree.cpp:1292-1305) */
    const SegP *sg = &c->seg[seg_of(c, ctb_at(c, cu->x, cu->y))];
    int16_t d[32 * 32], res[32 * 32], pred[32 * 32];
    const int qbd = 6 * (c->bd - 8);
    if (t->log2_size < 2 || t->log2_size > 5) return -30;
    if ((t->cbf & HEVCR_CBF_Y) || intra) {
        if (intra)
            intra_predict(c, 0, t->x, t->y, t->log2_size, c->mode_y[G4(c, t->x, t->y)], pred);
        else
            read_block(c, 0, t->x, t->y, t->log2_size, pred);
        if (t->cbf & HEVCR_CBF_Y) {
            if ((uint64_t)t->coeff_off[0] + (uint64_t)n * n > s->num_coeffs) return -31;
            if (bypass) {
                memcpy(res, s->coeffs + t->coeff_off[0], sizeof(int16_t) * n * n);
            } else {
                dequant(s->coeffs + t->coeff_off[0], d, t->log2_size, cu->qp_y + qbd, c->bd, s->scaling, 0, intra);
                inv_transform(d, res, t->log2_size, 0, intra, t->tskip & 1, c->bd);
            }
            put_block(c, 0, t->x, t->y, t->log2_size, pred, res);
        } else {
            put_block(c, 0, t->x, t->y, t->log2_size, pred, NULL);
        }
    }
    if (t->log2_size > 2 || t->blk_idx == 3) {
        const int l2c = t->log2_size > 2 ? t->log2_size - 1 : 2, nc = 1 << l2c;
        const int xb = t->log2_size == 2 ? t->xbase : t->x, yb = t->log2_size == 2 ? t->ybase : t->y;
        for (int ci = 1; ci <= 2; ci++) {
            const int coded = (t->cbf >> ci) & 1;
            if (!coded && !intra) continue;
            if (intra)
                intra_predict(c, ci, xb, yb, l2c, c->mode_c[G4(c, xb, yb)], pred);
            else
                read_block(c, ci, xb, yb, l2c, pred);
            if (coded) {
                if ((uint64_t)t->coeff_off[ci] + (uint64_t)nc * nc > s->num_coeffs) return -32;
                if (bypass) {
                    memcpy(res, s->coeffs + t->coeff_off[ci], sizeof(int16_t) * nc * nc);
                } else {
                    dequant(s->coeffs + t->coeff_off[ci], d, l2c, chroma_qp(c, cu->qp_y, ci, sg), c->bd, s->scaling,
                            ci, intra);
                    inv_transform(d, res, l2c, ci, intra, (t->tskip >> ci) & 1, c->bd);
                }
                put_block(c, ci, xb, yb, l2c, pred, res);
            } else {
                put_block(c, ci, xb, yb, l2c, pred, NULL);
            }
        }
    }
    /* deblocking side info of the TU (coding_tree.cpp:1233-1247) */
    for (int dy = 0; dy < n; dy += 4)
        for (int dx = 0; dx < n; dx += 4) {
            c->cbf[G4(c, t->x + dx, t->y + dy)] = (t->cbf & HEVCR_CBF_Y) ? 1 : 0;
            c->log2tu[G4(c, t->x + dx, t->y + dy)] = t->log2_size;
        }
    for (int k = 0; k < n; k += 4) {
        c->ev[G4(c, t->x, t->y + k)] = 1;
        c->eh[G4(c, t->x + k, t->y)] = 1;
    }
    return 0;
}

/* CU-level side info (coding_tree.cpp:956-1021): CU grid, TU-size grid reset,
 * CU edges and internal PU edges. */
static void cu_side_info(Ctx *c, const hevcr_cu *cu) {
    const int x0 = cu->x, y0 = cu->y, cb = 1 << cu->log2_size;
    for (int j = 0; j < cb; j += 8)
        for (int i = 0; i < cb; i += 8) {
            c->cu_pred[G8(c, x0 + i, y0 + j)] = cu->pred_mode;
            c->cu_flags[G8(c, x0 + i, y0 + j)] = (uint8_t)(cu->flags & HEVCR_CU_BYPASS);
        }
    for (int j = 0; j < cb; j += 4)
        for (int i = 0; i < cb; i += 4) {
            c->log2tu[G4(c, x0 + i, y0 + j)] = cu->log2_size;
            c->cbf[G4(c, x0 + i, y0 + j)] = 0;
        }
    int vx[2] = {-1, -1}, hy[2] = {-1, -1};
    switch (cu->part_mode) {
        case HEVCR_PART_Nx2N: vx[0] = cb / 2; break;
        case HEVCR_PART_NxN: vx[0] = cb / 2; hy[0] = cb / 2; break;
        case HEVCR_PART_nLx2N: vx[0] = cb / 4; break;
        case HEVCR_PART_nRx2N: vx[0] = 3 * cb / 4; break;
        case HEVCR_PART_2NxN: hy[0] = cb / 2; break;
        case HEVCR_PART_2NxnU: hy[0] = cb / 4; break;
        case HEVCR_PART_2NxnD: hy[0] = 3 * cb / 4; break;
        default: break;
    }
    for (int k = 0; k < cb; k += 4) {
        c->ev[G4(c, x0, y0 + k)] = 1;
        c->eh[G4(c, x0 + k, y0)] = 1;
        if (vx[0] >= 0) c->ev[G4(c, x0 + vx[0], y0 + k)] = 1;
        if (hy[0] >= 0) c->eh[G4(c, x0 + k, y0 + hy[0])] = 1;
    }
}

/* PCM CU (coding_tree.cpp:852-875, 1425-1460): the samples are written as
 * read (no shift to the bit depth), the CU grid says intra, PCM and the QP
 * carried by the record, and none of the filter grids is written, so they
 * keep their per-picture defaults (decoder.cpp:117-149). */
static void pcm_flags(Ctx *c, const hevcr_cu *cu) {
    const int cb = 1 << cu->log2_size;
    for (int j = 0; j < cb; j += 8)
        for (int i = 0; i < cb; i += 8) {
            c->cu_pred[G8(c, cu->x + i, cu->y + j)] = HEVCR_MODE_INTRA;
            c->cu_flags[G8(c, cu->x + i, cu->y + j)] = HEVCR_CU_PCM;
            c->cu_qp[G8(c, cu->x + i, cu->y + j)] = cu->qp_y;
        }
}

static int pcm_cu(Ctx *c, const hevcr_cu *cu, const hevcr_tu *t) {
    const int cb = 1 << cu->log2_size;
    if (!t || cu->part_mode != HEVCR_PART_2Nx2N || cu->pred_mode != HEVCR_MODE_INTRA || cu->log2_size > 5) return -43;
    for (int ci = 0; ci < 3; ci++) {
        const int n = ci ? cb >> 1 : cb, xc = ci ? cu->x >> 1 : cu->x, yc = ci ? cu->y >> 1 : cu->y;
        if ((uint64_t)t->coeff_off[ci] + (uint64_t)n * n > c->s->num_coeffs) return -44;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                c->pl[ci][(size_t)(yc + j) * c->pw[ci] + xc + i] = (uint16_t)c->s->coeffs[t->coeff_off[ci] + j * n + i];
    }
    pcm_flags(c, cu);
    return 0;
}

static int reconstruct(Ctx *c) {
    const hevcr_frame_syntax *s = c->s;
    uint32_t pi = 0, ti = 0;
    for (uint32_t ci = 0; ci < s->num_cus; ci++) {
        const hevcr_cu *cu = &s->cus[ci];
        const int cb = 1 << cu->log2_size;
        if (cu->log2_size < 3 || cu->log2_size > c->ctb_log2 || cu->x + cb > c->W || cu->y + cb > c->H) return -40;
        if (cu->flags & HEVCR_CU_PCM) {
            const int rc = pcm_cu(c, cu, ti < s->num_tus && s->tus[ti].cu_idx == ci ? &s->tus[ti] : NULL);
            if (rc) return rc;
            ti++;
            continue;
        }
        cu_side_info(c, cu);
        if (cu->pred_mode == HEVCR_MODE_INTRA) {
            /* luma mode per 4x4 per PU, chroma mode per CU (coding_tree.cpp:1124, 1134) */
            const int nxn = cu->part_mode == HEVCR_PART_NxN, h = nxn ? cb / 2 : cb;
            for (int j = 0; j < cb; j += 4)
                for (int i = 0; i < cb; i += 4) {
                    const int k = nxn ? (j >= h) * 2 + (i >= h) : 0;
                    c->mode_y[G4(c, cu->x + i, cu->y + j)] = cu->intra_luma[k];
                    c->mode_c[G4(c, cu->x + i, cu->y + j)] = cu->intra_chroma;
                }
        } else {
            for (int k = 0; k < cu->num_pu; k++, pi++) {
                if (pi >= s->num_pus) return -41;
                const hevcr_pu *pu = &s->pus[pi];
                Motion m;
                memset(&m, 0, sizeof(m));
                for (int l = 0; l < 2; l++) {
                    m.mv[l][0] = pu->mv[l][0];
                    m.mv[l][1] = pu->mv[l][1];
                    m.ref[l] = pu->ref_idx[l];
                }
                m.flags = pu->pred_flags;
                for (int j = 0; j < pu->h; j += 4)
                    for (int i = 0; i < pu->w; i += 4) c->mot[G4(c, pu->x + i, pu->y + j)] = m;
                for (int cc = 0; cc < 3; cc++) {
                    const int rc = mc_component(c, pu, cc);
                    if (rc) return rc;
                }
            }
        }
        for (; ti < s->num_tus && s->tus[ti].cu_idx == ci; ti++) {
            const int rc = do_tu(c, cu, &s->tus[ti]);
            if (rc) return rc;
        }
        for (int j = 0; j < cb; j += 8)
            for (int i = 0; i < cb; i += 8) c->cu_qp[G8(c, cu->x + i, cu->y + j)] = cu->qp_y;
    }
    if (pi != s->num_pus || ti != s->num_tus) return -42;
    return 0;
}

/* ---------------------------------------------------------------------------
 * Deblocking (deblocking.cpp:125-614)
 * ------------------------------------------------------------------------- */

static int ref_poc_of(const Ctx *c, const Motion *m, int l) {
    if (!((m->flags >> l) & 1) || m->ref[l] < 0 || m->ref[l] >= c->p->num_ref[l]) return -999999;
    return c->p->ref_poc[l][m->ref[l]];
}

static int mv_far(const int16_t *a, const int16_t *b) { return iabs(a[0] - b[0]) >= 4 || iabs(a[1] - b[1]) >= 4; }

/* Boundary strength between P (xp,yp) and Q (xq,yq), derive_bs
 * (deblocking.cpp:125-247): intra -> 2; transform edge (TU-size heuristic)
 * with coded luma on either side -> 1; else motion comparison by reference POC. */
static int boundary_strength(const Ctx *c, int xp, int yp, int xq, int yq) {
    if (c->cu_pred[G8(c, xp, yp)] == HEVCR_MODE_INTRA || c->cu_pred[G8(c, xq, yq)] == HEVCR_MODE_INTRA) return 2;
    const size_t gp = G4(c, xp, yp), gq = G4(c, xq, yq);
    const int tp = 1 << c->log2tu[gp], tq = 1 << c->log2tu[gq];
    int tu_edge;
    if (xp != xq)
        tu_edge = (xq % tq == 0) || ((xp + 1) % tp == 0 && xp + 1 == xq);
    else
        tu_edge = (yq % tq == 0) || ((yp + 1) % tp == 0 && yp + 1 == yq);
    if (tu_edge && (c->cbf[gp] || c->cbf[gq])) return 1;
    const Motion *mp = &c->mot[gp], *mq = &c->mot[gq];
    const int np = (mp->flags & 1) + ((mp->flags >> 1) & 1), nq = (mq->flags & 1) + ((mq->flags >> 1) & 1);
    if (np != nq) return 1;
    if (np == 1) {
        const int lp = (mp->flags & 1) ? 0 : 1, lq = (mq->flags & 1) ? 0 : 1;
        if (ref_poc_of(c, mp, lp) != ref_poc_of(c, mq, lq)) return 1;
        return mv_far(mp->mv[lp], mq->mv[lq]);
    }
    if (np == 2) {
        const int p0 = ref_poc_of(c, mp, 0), p1 = ref_poc_of(c, mp, 1);
        const int q0 = ref_poc_of(c, mq, 0), q1 = ref_poc_of(c, mq, 1);
        const int same = p0 == q0 && p1 == q1, swap = p0 == q1 && p1 == q0;
        if (!same && !swap) return 1;
        const int far_same = mv_far(mp->mv[0], mq->mv[0]) || mv_far(mp->mv[1], mq->mv[1]);
        const int far_swap = mv_far(mp->mv[0], mq->mv[1]) || mv_far(mp->mv[1], mq->mv[0]);
        if (same && !swap) return far_same;
        if (swap && !same) return far_swap;
        return far_same && far_swap;
    }
    return 0;
}

/* Filters one 4-sample luma segment and, for bS 2 on the chroma 8x8 grid, the
 * matching chroma segments (deblocking.cpp:419-608). vert: edge at x. */
/* Samples the filters leave alone: PCM CUs when pcm_loop_filter_disabled_flag
 * is set, and transquant-bypass CUs (deblocking.cpp:330-351, sao.cpp:85-128). */
static int no_filter(const Ctx *c, int x, int y) {
    const int f = c->cu_flags[G8(c, x, y)];
    return ((f & HEVCR_CU_PCM) && c->s->pcm_loop_filter_disabled) || (f & HEVCR_CU_BYPASS);
}

/* is_boundary_excluded (deblocking.cpp:50-120) past the picture edge: the
 * Q-side segment's slice_deblocking_filter_disabled_flag, a tile boundary
 * without loop_filter_across_tiles_enabled_flag, a slice-segment boundary
 * without the Q-side segment's slice_loop_filter_across_slices_enabled_flag. */
static int edge_excluded(const Ctx *c, int x, int y, int vert) {
    const int aq = ctb_at(c, x, y), ap = vert ? ctb_at(c, x - 1, y) : ctb_at(c, x, y - 1);
    const SegP *sq = &c->seg[seg_of(c, aq)];
    if (sq->dbk_off) return 1;
    if (!c->lf_tiles && tile_of(c, aq) != tile_of(c, ap)) return 1;
    if (seg_of(c, aq) != seg_of(c, ap) && !sq->lf_across) return 1;
    return 0;
}

static void deblock_segment(Ctx *c, int x, int y, int vert) {
    const int xp = vert ? x - 1 : x, yp = vert ? y : y - 1;
    const int bs = boundary_strength(c, xp, yp, x, y);
    if (!bs) return;
    const hevcr_frame_params *p = c->p;
    const SegP *sq = &c->seg[seg_of(c, ctb_at(c, x, y))]; /* the slice holding q0,0 (deblocking.cpp:426-429) */
    const int qpp = c->cu_qp[G8(c, xp, yp)], qpq = c->cu_qp[G8(c, x, y)];
    const int qpl = (qpq + qpp + 1) >> 1, bdm = 1 << (c->bd - 8);
    const int beta = beta_prime(clip3(0, 51, qpl + 2 * sq->beta)) * bdm;
    const int tc = tc_prime(clip3(0, 53, qpl + 2 * (bs - 1) + 2 * sq->tc)) * bdm;
    const int keep_p = no_filter(c, xp, yp), keep_q = no_filter(c, x, y);
    uint16_t *Y = c->pl[0];
    const int W = c->W;
    /* sample i (0..3) on P side / Q side of line k */
#define PS(k, i) (vert ? &Y[(size_t)(y + (k)) * W + x - 1 - (i)] : &Y[(size_t)(y - 1 - (i)) * W + x + (k)])
#define QS(k, i) (vert ? &Y[(size_t)(y + (k)) * W + x + (i)] : &Y[(size_t)(y + (i)) * W + x + (k)])
    const int dp0 = iabs(*PS(0, 2) - 2 * *PS(0, 1) + *PS(0, 0)), dp3 = iabs(*PS(3, 2) - 2 * *PS(3, 1) + *PS(3, 0));
    const int dq0 = iabs(*QS(0, 2) - 2 * *QS(0, 1) + *QS(0, 0)), dq3 = iabs(*QS(3, 2) - 2 * *QS(3, 1) + *QS(3, 0));
    const int d = dp0 + dq0 + dp3 + dq3;
    if (d < beta) {
        int strong = 1;
        for (int kk = 0; kk < 2; kk++) {
            const int k = kk * 3, dpq = kk ? dp3 + dq3 : dp0 + dq0;
            const int p0 = *PS(k, 0), p3 = *PS(k, 3), q0 = *QS(k, 0), q3 = *QS(k, 3);
            if (!(2 * dpq < (beta >> 2) && iabs(p3 - p0) + iabs(q0 - q3) < (beta >> 3) && iabs(p0 - q0) < ((5 * tc + 1) >> 1)))
                strong = 0;
        }
        const int side_thr = (beta + (beta >> 1)) >> 3;
        const int dep = dp0 + dp3 < side_thr, deq = dq0 + dq3 < side_thr;
        for (int k = 0; k < 4; k++) {
            int P[4], Q[4];
            for (int i = 0; i < 4; i++) {
                P[i] = *PS(k, i);
                Q[i] = *QS(k, i);
            }
            if (strong) {
                *PS(k, 0) = (uint16_t)clip3(P[0] - 2 * tc, P[0] + 2 * tc, (P[2] + 2 * P[1] + 2 * P[0] + 2 * Q[0] + Q[1] + 4) >> 3);
                *PS(k, 1) = (uint16_t)clip3(P[1] - 2 * tc, P[1] + 2 * tc, (P[2] + P[1] + P[0] + Q[0] + 2) >> 2);
                *PS(k, 2) = (uint16_t)clip3(P[2] - 2 * tc, P[2] + 2 * tc, (2 * P[3] + 3 * P[2] + P[1] + P[0] + Q[0] + 4) >> 3);
                *QS(k, 0) = (uint16_t)clip3(Q[0] - 2 * tc, Q[0] + 2 * tc, (P[1] + 2 * P[0] + 2 * Q[0] + 2 * Q[1] + Q[2] + 4) >> 3);
                *QS(k, 1) = (uint16_t)clip3(Q[1] - 2 * tc, Q[1] + 2 * tc, (P[0] + Q[0] + Q[1] + Q[2] + 2) >> 2);
                *QS(k, 2) = (uint16_t)clip3(Q[2] - 2 * tc, Q[2] + 2 * tc, (P[0] + Q[0] + Q[1] + 3 * Q[2] + 2 * Q[3] + 4) >> 3);
            } else {
                int delta = (9 * (Q[0] - P[0]) - 3 * (Q[1] - P[1]) + 8) >> 4;
                if (iabs(delta) < tc * 10) {
                    delta = clip3(-tc, tc, delta);
                    *PS(k, 0) = (uint16_t)clip3(0, c->maxv, P[0] + delta);
                    *QS(k, 0) = (uint16_t)clip3(0, c->maxv, Q[0] - delta);
                    if (dep)
                        *PS(k, 1) = (uint16_t)clip3(0, c->maxv, P[1] + clip3(-(tc >> 1), tc >> 1, (((P[2] + P[0] + 1) >> 1) - P[1] + delta) >> 1));
                    if (deq)
                        *QS(k, 1) = (uint16_t)clip3(0, c->maxv, Q[1] + clip3(-(tc >> 1), tc >> 1, (((Q[2] + Q[0] + 1) >> 1) - Q[1] - delta) >> 1));
                }
            }
            for (int i = 0; i < 3; i++) { /* nDp / nDq = 0 for a side that is not filtered */
                if (keep_p) *PS(k, i) = (uint16_t)P[i];
                if (keep_q) *QS(k, i) = (uint16_t)Q[i];
            }
        }
    }
#undef PS
#undef QS
    /* chroma: bS 2 and the chroma edge on the 8-sample chroma grid */
    if (bs == 2 && (((vert ? x : y) >> 1) & 7) == 0) {
        for (int ci = 1; ci <= 2; ci++) {
            const int off = ci == 1 ? p->pps_cb_qp_offset : p->pps_cr_qp_offset;
            const int qpc = qpc_from_qpi(clip3(0, 57, ((qpq + qpp + 1) >> 1) + off));
            const int tcc = tc_prime(clip3(0, 53, qpc + 2 + 2 * sq->tc)) * bdm;
            uint16_t *Cp = c->pl[ci];
            const int cw = c->pw[ci];
            for (int k = 0; k < 2; k++) {
                uint16_t *q0, *q1, *p0, *p1;
                if (vert) {
                    uint16_t *row = Cp + (size_t)((y >> 1) + k) * cw + (x >> 1);
                    q0 = row;
                    q1 = row + 1;
                    p0 = row - 1;
                    p1 = row - 2;
                } else {
                    uint16_t *col = Cp + (size_t)(y >> 1) * cw + (x >> 1) + k;
                    q0 = col;
                    q1 = col + cw;
                    p0 = col - cw;
                    p1 = col - 2 * cw;
                }
                const int delta = clip3(-tcc, tcc, ((((int)*q0 - *p0) * 4) + *p1 - *q1 + 4) >> 3);
                const int np0 = clip3(0, c->maxv, *p0 + delta), nq0 = clip3(0, c->maxv, *q0 - delta);
                if (!keep_p) *p0 = (uint16_t)np0;
                if (!keep_q) *q0 = (uint16_t)nq0;
            }
        }
    }
}

static void deblock(Ctx *c) {
    for (int pass = 0; pass < 2; pass++) {
        const int vert = pass == 0;
        for (int ye = 0; ye < c->H; ye += 8)
            for (int xe = 0; xe < c->W; xe += 8)
                for (int seg = 0; seg < 2; seg++) {
                    const int x = vert ? xe : xe + 4 * seg, y = vert ? ye + 4 * seg : ye;
                    if (x >= c->W || y >= c->H) continue;
                    if (vert ? !c->ev[G4(c, x, y)] : !c->eh[G4(c, x, y)]) continue;
                    if (vert ? x == 0 : y == 0) continue; /* picture edge */
                    if (edge_excluded(c, x, y, vert)) continue;
                    deblock_segment(c, x, y, vert);
                }
    }
}

/* ---------------------------------------------------------------------------
 * SAO (sao.cpp:20-243): reads the deblocked picture, writes a new one.
 * ------------------------------------------------------------------------- */

static const int kEoDx[4][2] = {{-1, 1}, {0, 0}, {-1, 1}, {1, -1}};
static const int kEoDy[4][2] = {{0, 0}, {-1, 1}, {-1, 1}, {-1, 1}};

/* An edge-offset neighbour (luma coordinates) that SAO may not read from the
 * sample at (x, y) (sao.cpp:142-185): another slice segment whose boundary
 * is crossed without slice_loop_filter_across_slices_enabled_flag (the later
 * segment's flag), or another tile without loop_filter_across_tiles. */
static int sao_excluded(const Ctx *c, int x, int y, int nx, int ny) {
    const int a = ctb_at(c, x, y), b = ctb_at(c, nx, ny);
    const int sa = seg_of(c, a), sb = seg_of(c, b);
    if (sa != sb && !c->seg[sb < sa ? sa : sb].lf_across) return 1;
    if (!c->lf_tiles && tile_of(c, a) != tile_of(c, b)) return 1;
    return 0;
}

static int sao(Ctx *c) {
    if (!c->p->sao_enabled || !c->s->sao) return 0;
    int any = 0;
    for (int i = 0; i < c->wctb * c->hctb; i++)
        for (int ci = 0; ci < 3; ci++) any |= c->s->sao[i].type[ci] != 0;
    if (!any) return 0;
    for (int ci = 0; ci < 3; ci++) {
        const int w = c->pw[ci], h = c->ph[ci], cs = ci ? c->ctb >> 1 : c->ctb;
        uint16_t *src = malloc((size_t)w * h * 2);
        if (!src) return -50;
        memcpy(src, c->pl[ci], (size_t)w * h * 2);
        uint16_t *dst = c->pl[ci];
        for (int ry = 0; ry < c->hctb; ry++)
            for (int rx = 0; rx < c->wctb; rx++) {
                const hevcr_sao *sp = &c->s->sao[ry * c->wctb + rx];
                const int type = sp->type[ci];
                if (!type) continue;
                const int x0 = rx * cs, y0 = ry * cs;
                const int x1 = x0 + cs < w ? x0 + cs : w, y1 = y0 + cs < h ? y0 + cs : h;
                const int sub = ci ? 1 : 0;
                if (type == 2) {
                    const int e = sp->eo_class[ci] & 3;
                    for (int y = y0; y < y1; y++)
                        for (int x = x0; x < x1; x++) {
                            if (no_filter(c, x << sub, y << sub)) continue;
                            const int ax = x + kEoDx[e][0], ay = y + kEoDy[e][0];
                            const int bx = x + kEoDx[e][1], by = y + kEoDy[e][1];
                            if (ax < 0 || ax >= w || ay < 0 || ay >= h || bx < 0 || bx >= w || by < 0 || by >= h) continue;
                            if (sao_excluded(c, x << sub, y << sub, ax << sub, ay << sub) ||
                                sao_excluded(c, x << sub, y << sub, bx << sub, by << sub))
                                continue;
                            const int v = src[(size_t)y * w + x], a = src[(size_t)ay * w + ax], b = src[(size_t)by * w + bx];
                            const int idx = 2 + (v > a) - (v < a) + (v > b) - (v < b);
                            const int off = sp->offset[ci][idx];
                            if (off) dst[(size_t)y * w + x] = (uint16_t)clip3(0, c->maxv, v + off);
                        }
                } else {
                    /* band index relative to sao_band_position, not wrapped (sao.cpp:226-230) */
                    const int shift = c->bd - 5, pos = sp->band_pos[ci];
                    for (int y = y0; y < y1; y++)
                        for (int x = x0; x < x1; x++) {
                            if (no_filter(c, x << sub, y << sub)) continue;
                            const int v = src[(size_t)y * w + x], bi = (v >> shift) - pos;
                            if (bi < 0 || bi > 3) continue;
                            const int off = sp->offset[ci][bi];
                            if (off) dst[(size_t)y * w + x] = (uint16_t)clip3(0, c->maxv, v + off);
                        }
                }
            }
        free(src);
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Entry points
 * ------------------------------------------------------------------------- */

static void copy_planes(const Ctx *c, uint16_t *const *out) {
    if (!out) return;
    for (int i = 0; i < 3; i++)
        if (out[i]) memcpy(out[i], c->pl[i], (size_t)c->pw[i] * c->ph[i] * 2);
}

int oracle_recon_frame(const hevcr_frame_syntax *syn, const uint16_t *const *ref_planes, int num_ref_pics,
                       uint16_t *const *out_recon, uint16_t *const *out_dbk, uint16_t *const *out_sao) {
    if (!syn) return -1;
    if (syn->sig) return -2; /* the restatement reads the dense coefficient pool only */
    Ctx c;
    int rc = ctx_init(&c, syn);
    if (!rc) {
        c.refs = ref_planes;
        c.nrefs = num_ref_pics;
        rc = reconstruct(&c);
    }
    if (!rc) {
        copy_planes(&c, out_recon);
        deblock(&c);
        copy_planes(&c, out_dbk);
        rc = sao(&c);
        copy_planes(&c, out_sao);
    }
    ctx_free(&c);
    return rc;
}

/* Side information only (no sample reconstruction), then the filters on the
 * caller's planes. */
int oracle_filters(const hevcr_frame_syntax *syn, uint16_t *const *planes, int do_deblock, int do_sao) {
    if (!syn || !planes) return -1;
    Ctx c;
    int rc = ctx_init(&c, syn);
    if (rc) {
        ctx_free(&c);
        return rc;
    }
    const hevcr_frame_syntax *s = syn;
    uint32_t pi = 0, ti = 0;
    for (uint32_t ci = 0; ci < s->num_cus && !rc; ci++) {
        const hevcr_cu *cu = &s->cus[ci];
        const int cb = 1 << cu->log2_size;
        if (cu->flags & HEVCR_CU_PCM) {
            pcm_flags(&c, cu);
            for (; ti < s->num_tus && s->tus[ti].cu_idx == ci; ti++) {
            }
            continue;
        }
        cu_side_info(&c, cu);
        if (cu->pred_mode != HEVCR_MODE_INTRA)
            for (int k = 0; k < cu->num_pu && pi < s->num_pus; k++, pi++) {
                const hevcr_pu *pu = &s->pus[pi];
                Motion m;
                memset(&m, 0, sizeof(m));
                for (int l = 0; l < 2; l++) {
                    m.mv[l][0] = pu->mv[l][0];
                    m.mv[l][1] = pu->mv[l][1];
                    m.ref[l] = pu->ref_idx[l];
                }
                m.flags = pu->pred_flags;
                for (int j = 0; j < pu->h; j += 4)
                    for (int i = 0; i < pu->w; i += 4) c.mot[G4(&c, pu->x + i, pu->y + j)] = m;
            }
        for (; ti < s->num_tus && s->tus[ti].cu_idx == ci; ti++) {
            const hevcr_tu *t = &s->tus[ti];
            const int n = 1 << t->log2_size;
            for (int dy = 0; dy < n; dy += 4)
                for (int dx = 0; dx < n; dx += 4) {
                    c.cbf[G4(&c, t->x + dx, t->y + dy)] = (t->cbf & HEVCR_CBF_Y) ? 1 : 0;
                    c.log2tu[G4(&c, t->x + dx, t->y + dy)] = t->log2_size;
                }
            for (int k = 0; k < n; k += 4) {
                c.ev[G4(&c, t->x, t->y + k)] = 1;
                c.eh[G4(&c, t->x + k, t->y)] = 1;
            }
        }
        for (int j = 0; j < cb; j += 8)
            for (int i = 0; i < cb; i += 8) c.cu_qp[G8(&c, cu->x + i, cu->y + j)] = cu->qp_y;
    }
    for (int i = 0; i < 3; i++) memcpy(c.pl[i], planes[i], (size_t)c.pw[i] * c.ph[i] * 2);
    if (do_deblock) deblock(&c);
    if (do_sao) rc = sao(&c);
    for (int i = 0; i < 3; i++) memcpy(planes[i], c.pl[i], (size_t)c.pw[i] * c.ph[i] * 2);
    ctx_free(&c);
    return rc;
}
