/*
 * footile_oracle.c -- CPU oracle (test infrastructure only; see the header).
 *
 * Plain C restatement of footile 0.7.0's sequential fill/stroke algorithm.
 * Each function names the reference file:line it follows.  Built with
 * -O2 -mssse3 -fno-fast-math -ffp-contract=off (oracle/Makefile).
 */
#include "footile_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <tmmintrin.h>

#define FX_ONE 65536
#define FX_HALF 32768
#define FX_FRACT_MASK 0xFFFF

/* Rust release-mode i32 arithmetic wraps; C signed overflow does not. */
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t low32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }
static inline int16_t low16(int32_t v) { return (int16_t)(uint16_t)(uint32_t)v; }

/* ========================================================================
 * Fixed 16.16 -- src/fixed.rs
 * ======================================================================== */
int32_t or_fx_from_i32(int32_t i) { return (int32_t)((uint32_t)i << 16); } /* fixed.rs:74-79 */
int32_t or_fx_to_i32(int32_t f) { return f >> 16; }                       /* fixed.rs:81-86 */

/* fixed.rs:88-93: `(f * 65536.0) as i32`; Rust's `as` saturates, NaN -> 0. */
int32_t or_fx_from_f32(float f)
{
    float v = f * 65536.0f;
    if (v != v)
        return 0;
    if (v >= 2147483648.0f)
        return INT32_MAX;
    if (v < -2147483648.0f)
        return INT32_MIN;
    return (int32_t)v;
}

float or_fx_to_f32(int32_t f) { return (float)f / 65536.0f; } /* fixed.rs:95-100 */
int32_t or_fx_add(int32_t a, int32_t b) { return wadd(a, b); } /* fixed.rs:24-30 */
int32_t or_fx_sub(int32_t a, int32_t b) { return wsub(a, b); } /* fixed.rs:32-38 */

/* fixed.rs:40-47: i64 product, arithmetic >> 16, truncated to i32. */
int32_t or_fx_mul(int32_t a, int32_t b) { return low32(((int64_t)a * (int64_t)b) >> 16); }

/* fixed.rs:49-56: (i64 a << 16) / b, truncating division, truncated to i32.
 * The reference panics on b == 0; no caller on the fill path divides by 0. */
int32_t or_fx_div(int32_t a, int32_t b)
{
    if (b == 0)
        return 0;
    return low32(((int64_t)a * 65536) / (int64_t)b);
}

int32_t or_fx_shl(int32_t a, uint32_t s) { return (int32_t)((uint32_t)a << (s & 31)); } /* fixed.rs:58-64 */
int32_t or_fx_shr(int32_t a, uint32_t s) { return a >> (s & 31); }                       /* fixed.rs:66-72 */
int32_t or_fx_abs(int32_t a) { return a < 0 ? (int32_t)(0u - (uint32_t)a) : a; }       /* fixed.rs:122-124 */
int32_t or_fx_floor(int32_t a) { return (int32_t)((uint32_t)a & ~(uint32_t)FX_FRACT_MASK); } /* fixed.rs:127-129 */
int32_t or_fx_ceil(int32_t a) { return or_fx_floor(wsub(wadd(a, FX_ONE), 1)); }          /* fixed.rs:132-134 */
int32_t or_fx_round(int32_t a) { return or_fx_floor(wadd(a, FX_HALF)); }                 /* fixed.rs:137-139 */
int32_t or_fx_trunc(int32_t a) { return a >= 0 ? or_fx_floor(a) : or_fx_ceil(a); }        /* fixed.rs:142-148 */
int32_t or_fx_fract(int32_t a) { return a & FX_FRACT_MASK; }                             /* fixed.rs:151-153 */
int32_t or_fx_avg(int32_t a, int32_t b) { return wadd(a, b) >> 1; }                      /* fixed.rs:156-158 */

/* fig.rs:116-119: cross-product sign with Fixed multiplies. */
int or_fx_widdershins(int32_t ax, int32_t ay, int32_t bx, int32_t by)
{
    return or_fx_mul(ax, by) > or_fx_mul(bx, ay);
}

/* fig.rs:677-682: round(fcov << 8) as an integer in 0..=256. */
int16_t or_pixel_cov(int32_t fcov)
{
    return (int16_t)or_fx_to_i32(or_fx_round(or_fx_shl(fcov, 8)));
}

/* ========================================================================
 * Row accumulate -- src/imgbuf.rs
 * ======================================================================== */
static inline uint8_t sat_i16_u8(int16_t v) { return v < 0 ? 0 : (v > 255 ? 255 : (uint8_t)v); } /* imgbuf.rs:64-66 */

/* imgbuf.rs:54-61 */
void or_accumulate_non_zero_scalar(uint8_t *dst, int16_t *src, size_t n)
{
    int16_t sum = 0;
    for (size_t i = 0; i < n; i++) {
        sum = low16((int32_t)sum + src[i]);
        src[i] = 0;
        dst[i] = sat_i16_u8(sum);
    }
}

/* imgbuf.rs:157-167 */
void or_accumulate_even_odd_scalar(uint8_t *dst, int16_t *src, size_t n)
{
    int16_t sum = 0;
    for (size_t i = 0; i < n; i++) {
        sum = low16((int32_t)sum + src[i]);
        src[i] = 0;
        int16_t v = (int16_t)(sum & 0xFF);
        int16_t odd = (int16_t)(sum & 0x100);
        int16_t c = low16(v - odd);
        c = c < 0 ? low16(-(int32_t)c) : c;
        dst[i] = sat_i16_u8(c);
    }
}

/* 8-lane inclusive prefix of i16 lanes by byte shifts (imgbuf.rs:103-117). */
static inline __m128i prefix_i16x8(__m128i a)
{
    a = _mm_add_epi16(a, _mm_slli_si128(a, 2));
    a = _mm_add_epi16(a, _mm_slli_si128(a, 4));
    a = _mm_add_epi16(a, _mm_slli_si128(a, 8));
    return a;
}

/* The reference's SSSE3 loop reads and writes whole 8-lane chunks past the
 * row end (imgbuf.rs:78-96); lanes past n never feed lanes below n, so the
 * tail chunk here is staged through a zero-padded local copy instead. */
static void accumulate_ssse3(uint8_t *dst, int16_t *src, size_t n, int even_odd)
{
    const __m128i zero = _mm_setzero_si128();
    const __m128i bcast7 = _mm_set1_epi16(0x0F0E);
    const __m128i m_ff = _mm_set1_epi16(0xFF);
    const __m128i m_100 = _mm_set1_epi16(0x100);
    __m128i carry = zero;
    for (size_t i = 0; i < n; i += 8) {
        size_t k = n - i < 8 ? n - i : 8;
        int16_t s_tmp[8] = {0};
        uint8_t d_tmp[16];
        memcpy(s_tmp, src + i, k * sizeof(int16_t));
        memset(src + i, 0, k * sizeof(int16_t));
        __m128i a = _mm_loadu_si128((const __m128i *)s_tmp);
        a = _mm_add_epi16(prefix_i16x8(a), carry);
        __m128i v = a;
        if (even_odd) {
            __m128i lo = _mm_and_si128(a, m_ff);
            __m128i odd = _mm_and_si128(a, m_100);
            v = _mm_abs_epi16(_mm_sub_epi16(lo, odd));
        }
        _mm_storeu_si128((__m128i *)d_tmp, _mm_packus_epi16(v, v));
        memcpy(dst + i, d_tmp, k);
        carry = _mm_shuffle_epi8(a, bcast7);
    }
}

void or_accumulate_non_zero_ssse3(uint8_t *dst, int16_t *src, size_t n) { accumulate_ssse3(dst, src, n, 0); }
void or_accumulate_even_odd_ssse3(uint8_t *dst, int16_t *src, size_t n) { accumulate_ssse3(dst, src, n, 1); }

/* ========================================================================
 * pix 0.14 subset (external crate, not in the reference tree).
 * UNVERIFIED recollection of pix::chan::Ch8 Mul and pix::ops::SrcOver as used
 * by Pixel::composite_channels_alpha (call sites fig.rs:637-643,655-664):
 *   Ch8 mul: widen both to 12 bits ((v<<4)|(v>>4)), multiply, >> 16.
 *   SrcOver per channel (alpha included): d = s*a + d*(MAX - sa*a), saturating.
 * Pinned only by fig_3x3 (coverage 255, opaque colour).
 * ======================================================================== */
uint8_t or_ch8_mul(uint8_t l, uint8_t r)
{
    uint32_t a = ((uint32_t)l << 4) | ((uint32_t)l >> 4);
    uint32_t b = ((uint32_t)r << 4) | ((uint32_t)r >> 4);
    return (uint8_t)((a * b) >> 16);
}

static inline uint8_t sat_add_u8(uint8_t a, uint8_t b)
{
    unsigned s = (unsigned)a + b;
    return s > 255 ? 255 : (uint8_t)s;
}

static int fmt_bpp(int fmt) { return fmt == OR_RGBA8P ? 4 : (fmt == OR_GRAYA8P ? 2 : 1); }

static void composite_alpha(uint8_t *d, const uint8_t *s, int nch, uint8_t alpha)
{
    uint8_t sa = or_ch8_mul(s[nch - 1], alpha);
    uint8_t inv = (uint8_t)(255 - sa);
    for (int c = 0; c < nch; c++)
        d[c] = sat_add_u8(or_ch8_mul(s[c], alpha), or_ch8_mul(d[c], inv));
}

/* ========================================================================
 * Fig -- src/fig.rs
 * ======================================================================== */
typedef struct { int32_t x, y; } fxpt;
typedef struct { uint32_t start, n_points; int done; } subfig; /* fig.rs:36-43 */

struct or_fig {
    fxpt *pts;
    size_t n, cap;
    subfig *subs;
    size_t nsubs, subcap;
    size_t offered; /* add_point calls, for statistics */
};

#define VID_MAX 65535u /* vid.rs:17 */

or_fig *or_fig_new(void) /* fig.rs:339-344 */
{
    or_fig *f = (or_fig *)calloc(1, sizeof(or_fig));
    if (!f)
        return NULL;
    f->cap = 1024;
    f->pts = (fxpt *)malloc(f->cap * sizeof(fxpt));
    f->subcap = 16;
    f->subs = (subfig *)calloc(f->subcap, sizeof(subfig));
    f->nsubs = 1; /* SubFig(0) */
    return f;
}

void or_fig_free(or_fig *f)
{
    if (!f)
        return;
    free(f->pts);
    free(f->subs);
    free(f);
}

static subfig *sub_current(or_fig *f) { return &f->subs[f->nsubs - 1]; } /* fig.rs:347-354 */

static void sub_add(or_fig *f) /* fig.rs:357-360 */
{
    if (f->nsubs == f->subcap) {
        f->subcap *= 2;
        f->subs = (subfig *)realloc(f->subs, f->subcap * sizeof(subfig));
    }
    subfig s = {(uint32_t)f->n, 0, 0};
    f->subs[f->nsubs++] = s;
}

static int is_coincident(const or_fig *f, fxpt p) /* fig.rs:445-451 */
{
    return f->n > 0 && f->pts[f->n - 1].x == p.x && f->pts[f->n - 1].y == p.y;
}

static void sub_set_done(or_fig *f) /* fig.rs:373-383 */
{
    subfig *s = sub_current(f);
    if (s->n_points > 0) {
        fxpt p = f->pts[s->start];
        if (is_coincident(f, p)) {
            f->n--;
            s->n_points--;
        }
        s->done = 1;
    }
}

static void fig_add_fx(or_fig *f, fxpt p) /* fig.rs:428-442 */
{
    f->offered++;
    if (f->n < VID_MAX) {
        int done = sub_current(f)->done;
        if (done)
            sub_add(f);
        if (done || !is_coincident(f, p)) {
            if (f->n == f->cap) {
                f->cap *= 2;
                f->pts = (fxpt *)realloc(f->pts, f->cap * sizeof(fxpt));
            }
            f->pts[f->n++] = p;
            sub_current(f)->n_points++;
        }
    }
}

void or_fig_add_point(or_fig *f, float x, float y)
{
    fxpt p = {or_fx_from_f32(x), or_fx_from_f32(y)};
    fig_add_fx(f, p);
}

void or_fig_close(or_fig *f) /* fig.rs:457-461 */
{
    if (f->n > 0)
        sub_set_done(f);
}

size_t or_fig_points(const or_fig *f, int32_t *xy, size_t cap)
{
    for (size_t i = 0; i < f->n && i < cap; i++) {
        xy[2 * i] = f->pts[i].x;
        xy[2 * i + 1] = f->pts[i].y;
    }
    return f->n;
}

enum { DIR_FWD = 0, DIR_REV = 1 }; /* fig.rs:30-33 */

static const subfig *sub_at(const or_fig *f, uint32_t vid) /* fig.rs:386-394 */
{
    for (size_t i = 0; i < f->nsubs; i++)
        if (vid < f->subs[i].start + f->subs[i].n_points)
            return &f->subs[i];
    return NULL; /* unreachable in the reference */
}

static uint32_t fig_next(const or_fig *f, uint32_t vid, int dir) /* fig.rs:143-163,397-399 */
{
    const subfig *s = sub_at(f, vid);
    if (dir == DIR_FWD) {
        uint32_t v = vid + 1;
        return v < s->start + s->n_points ? v : s->start;
    }
    if (vid > s->start)
        return vid - 1;
    if (s->n_points > 0)
        return s->start + s->n_points - 1;
    return s->start;
}

static int32_t row_of(int32_t y) { return or_fx_to_i32(y); } /* fig.rs:167-169 */

typedef struct { /* fig.rs:46-66 */
    uint32_t v1;
    int32_t y_upper, y_lower;
    int dir;
    int32_t step_pix, inv_slope, x_bot, min_x, max_x;
} edge;

static int32_t calculate_step(int32_t dx, int32_t dy) /* fig.rs:204-210 */
{
    if (dx != 0) {
        int32_t s = or_fx_abs(or_fx_div(dy, dx));
        return s < FX_ONE ? s : FX_ONE;
    }
    return 0;
}

static edge edge_new(uint32_t v1, fxpt p0, fxpt p1, int dir) /* fig.rs:179-201 */
{
    edge e;
    int32_t dx = wsub(p1.x, p0.x);
    int32_t dy = wsub(p1.y, p0.y);
    e.v1 = v1;
    e.y_upper = p0.y;
    e.y_lower = p1.y;
    e.dir = dir;
    e.step_pix = calculate_step(dx, dy);
    e.inv_slope = or_fx_div(dx, dy);
    int32_t y_bot = wsub(or_fx_floor(wadd(e.y_upper, FX_ONE)), e.y_upper);
    e.x_bot = wadd(p0.x, or_fx_mul(e.inv_slope, y_bot));
    e.min_x = 0;
    e.max_x = 0;
    return e;
}

static int is_ending(const edge *e, int32_t y_row) { return row_of(e->y_lower) == y_row; } /* fig.rs:233-235 */

static int16_t continuing_cov(const edge *e, int32_t y_row) /* fig.rs:252-259 */
{
    return is_ending(e, y_row) ? or_pixel_cov(or_fx_fract(e->y_lower)) : 256;
}

static int16_t starting_cov(const edge *e) /* fig.rs:238-241 */
{
    return low16(continuing_cov(e, row_of(e->y_upper)) - or_pixel_cov(or_fx_fract(e->y_upper)));
}

static void set_x_limits(edge *e, int32_t x0, int32_t y_row) /* fig.rs:269-278 */
{
    int32_t x1 = e->x_bot;
    if (is_ending(e, y_row))
        x1 = wsub(e->x_bot, or_fx_mul(e->inv_slope, wsub(or_fx_ceil(e->y_lower), e->y_lower)));
    e->min_x = x0 < x1 ? x0 : x1;
    e->max_x = x0 < x1 ? x1 : x0;
}

static void x_limits_starting(edge *e) /* fig.rs:244-249 */
{
    int32_t y0 = wsub(FX_ONE, or_fx_fract(e->y_upper));
    set_x_limits(e, wsub(e->x_bot, or_fx_mul(e->inv_slope, y0)), row_of(e->y_upper));
}

static void x_limits_continuing(edge *e, int32_t y_row) /* fig.rs:262-266 */
{
    set_x_limits(e, wsub(e->x_bot, e->inv_slope), y_row);
}

static int32_t step_cov(const edge *e, int32_t r) /* fig.rs:315-321 */
{
    return e->step_pix > 0 ? or_fx_mul(r, e->step_pix) : r;
}

static int32_t first_cov(const edge *e, int32_t full_cov) /* fig.rs:305-312 */
{
    int32_t r;
    if (or_fx_to_i32(e->min_x) == or_fx_to_i32(e->max_x))
        r = or_fx_mul(wsub(FX_ONE, or_fx_fract(or_fx_avg(e->max_x, e->min_x))), full_cov);
    else
        r = or_fx_mul(wsub(FX_ONE, or_fx_fract(e->min_x)), FX_HALF);
    return step_cov(e, r);
}

/* fig.rs:285-302 */
static void scan_area(const edge *e, int dir, int16_t cov, int16_t *area, size_t len)
{
    int16_t ed = e->dir == dir ? 1 : -1;
    int32_t full_cov = or_fx_from_f32((float)cov / 256.0f);
    int32_t x_cov = first_cov(e, full_cov);
    int32_t stp = step_cov(e, FX_ONE);
    int16_t sum_pix = 0;
    for (int32_t x = or_fx_to_i32(e->min_x); x < (int32_t)len; x++) {
        int16_t x_pix = or_pixel_cov(x_cov);
        if (x_pix > cov)
            x_pix = cov;
        int16_t p = low16(x_pix - sum_pix);
        size_t ix = x > 0 ? (size_t)x : 0;
        area[ix] = low16(area[ix] + p * ed);
        sum_pix = low16(sum_pix + p);
        if (sum_pix >= cov)
            break;
        x_cov = wadd(x_cov, stp);
        if (x_cov > FX_ONE)
            x_cov = FX_ONE;
    }
}

typedef struct { /* fig.rs:77-95 */
    const or_fig *fig;
    int rule, dir;
    uint8_t *px;
    uint32_t w, h;
    int fmt;
    const uint8_t *clr;
    int16_t *sgn;
    size_t sgn_len;
    int ssse3;
    edge *edges;
    size_t ne, ecap;
} scanner;

static void add_edge(scanner *s, uint32_t v0, uint32_t v1, int dir) /* fig.rs:589-600 */
{
    edge e = edge_new(v1, s->fig->pts[v0], s->fig->pts[v1], dir);
    int16_t cov = starting_cov(&e);
    if (cov > 0) {
        x_limits_starting(&e);
        scan_area(&e, s->dir, cov, s->sgn, s->sgn_len);
    }
    if (s->ne == s->ecap) {
        s->ecap = s->ecap ? 2 * s->ecap : 16;
        s->edges = (edge *)realloc(s->edges, s->ecap * sizeof(edge));
    }
    s->edges[s->ne++] = e;
}

static void remove_edge(scanner *s, uint32_t v1, int dir) /* fig.rs:603-617 */
{
    for (size_t i = 0; i < s->ne; i++) {
        if (s->edges[i].v1 == v1 && s->edges[i].dir == dir) {
            s->edges[i] = s->edges[--s->ne]; /* swap_remove */
            return;
        }
    }
}

static void update_edges(scanner *s, uint32_t vid, int dir) /* fig.rs:576-586 */
{
    uint32_t v = fig_next(s->fig, vid, dir);
    if (v != vid) {
        int32_t y = s->fig->pts[vid].y;
        int32_t yv = s->fig->pts[v].y;
        if (yv > y)
            add_edge(s, vid, v, dir);
        else if (yv < y)
            remove_edge(s, vid, dir == DIR_FWD ? DIR_REV : DIR_FWD);
    }
}

static void scan_continuing_edges(scanner *s, int32_t y_row) /* fig.rs:557-566 */
{
    for (size_t i = 0; i < s->ne; i++) {
        edge *e = &s->edges[i];
        int16_t cov = continuing_cov(e, y_row);
        if (cov > 0) {
            x_limits_continuing(e, y_row);
            scan_area(e, s->dir, cov, s->sgn, s->sgn_len);
        }
    }
}

static void advance_edges(scanner *s) /* fig.rs:569-573 */
{
    for (size_t i = 0; i < s->ne; i++)
        s->edges[i].x_bot = wadd(s->edges[i].x_bot, s->edges[i].inv_slope);
}

static void rasterize_row(scanner *s, uint8_t *row) /* fig.rs:621-665 */
{
    if (s->fmt == OR_MATTE8) { /* fig.rs:632-635,650-653 */
        if (s->rule == OR_NON_ZERO) {
            if (s->ssse3)
                or_accumulate_non_zero_ssse3(row, s->sgn, s->w);
            else
                or_accumulate_non_zero_scalar(row, s->sgn, s->w);
        } else {
            if (s->ssse3)
                or_accumulate_even_odd_ssse3(row, s->sgn, s->w);
            else
                or_accumulate_even_odd_scalar(row, s->sgn, s->w);
        }
        return;
    }
    int nch = fmt_bpp(s->fmt);
    int16_t sum = 0;
    for (uint32_t x = 0; x < s->w; x++) {
        sum = low16(sum + s->sgn[x]);
        s->sgn[x] = 0;
        int16_t c = sum;
        if (s->rule == OR_EVEN_ODD) {
            int16_t v = (int16_t)(sum & 0xFF);
            int16_t odd = (int16_t)(sum & 0x100);
            c = low16(v - odd);
            c = c < 0 ? low16(-(int32_t)c) : c;
        }
        composite_alpha(row + (size_t)x * nch, s->clr, nch, sat_i16_u8(c));
    }
}

static int cmp_fx_key(const void *a, const void *b)
{
    const int64_t *ka = (const int64_t *)a, *kb = (const int64_t *)b;
    for (int i = 0; i < 3; i++)
        if (ka[i] != kb[i])
            return ka[i] < kb[i] ? -1 : 1;
    return 0;
}

/* fig.rs:480-502 (+ Scanner::scan_vertices fig.rs:536-554) */
int or_fig_fill(const or_fig *f, int rule, uint8_t *px, uint32_t w, uint32_t h, int fmt,
                const uint8_t clr[4], int16_t *sgn, size_t sgn_len, int use_ssse3)
{
    if (sgn_len < w)
        return OR_EINVAL;
    size_t n = f->n;
    if (n == 0)
        return OR_OK;
    if (!sub_current((or_fig *)f)->done)
        return OR_EINVAL; /* fig.rs:492 assert */
    /* Stable sort of vids by (y, x): ties keep vid order (fig.rs:493-494). */
    int64_t *keys = (int64_t *)malloc(n * 3 * sizeof(int64_t));
    if (!keys)
        return OR_ENOMEM;
    for (size_t i = 0; i < n; i++) {
        keys[3 * i] = f->pts[i].y;
        keys[3 * i + 1] = f->pts[i].x;
        keys[3 * i + 2] = (int64_t)i;
    }
    qsort(keys, n, 3 * sizeof(int64_t), cmp_fx_key);
    uint32_t v0 = (uint32_t)keys[2];
    /* get_dir, fig.rs:402-411 */
    fxpt p = f->pts[v0];
    fxpt pf = f->pts[fig_next(f, v0, DIR_FWD)];
    fxpt pr = f->pts[fig_next(f, v0, DIR_REV)];
    int dir = or_fx_widdershins(wsub(pr.x, p.x), wsub(pr.y, p.y), wsub(pf.x, p.x), wsub(pf.y, p.y))
                  ? DIR_FWD
                  : DIR_REV;
    int32_t top_row = row_of(p.y);
    scanner s;
    memset(&s, 0, sizeof s);
    s.fig = f;
    s.rule = rule;
    s.dir = dir;
    s.px = px;
    s.w = w;
    s.h = h;
    s.fmt = fmt;
    s.clr = clr;
    s.sgn = sgn;
    s.sgn_len = sgn_len;
    s.ssse3 = use_ssse3;
    size_t pitch = (size_t)w * fmt_bpp(fmt);
    int32_t y_row = top_row;
    size_t vi = 0;
    /* rows_mut((0, max(top_row,0), w, h)): rows max(top_row,0) .. h-1 */
    for (int64_t r = top_row > 0 ? top_row : 0; r < (int64_t)h; r++) {
        scan_continuing_edges(&s, y_row);
        while (vi < n && row_of((int32_t)keys[3 * vi]) <= y_row) {
            uint32_t vid = (uint32_t)keys[3 * vi + 2];
            vi++;
            update_edges(&s, vid, DIR_FWD);
            update_edges(&s, vid, DIR_REV);
        }
        rasterize_row(&s, px + (size_t)r * pitch);
        advance_edges(&s);
        y_row++;
    }
    free(s.edges);
    free(keys);
    return OR_OK;
}

/* ========================================================================
 * Stroke -- src/stroker.rs (parity unpinned; see the header)
 * ======================================================================== */
typedef struct { float x, y; } pt2;
typedef struct { float x, y, w; } widept; /* geom.rs:9 */
typedef struct { uint16_t start, n_points; int joined, done; } substroke; /* stroker.rs:31-40 */

typedef struct {
    int join;
    float miter;
    float tol_sq;
    widept *pts;
    size_t n, cap;
    substroke *subs;
    size_t nsubs, subcap;
    int err;
} stroke;

typedef struct {
    or_path_op *v;
    size_t n, cap;
    int err;
} opvec;

static void opvec_push(opvec *o, uint32_t tag, float x, float y)
{
    if (o->n == o->cap) {
        o->cap = o->cap ? 2 * o->cap : 256;
        or_path_op *nv = (or_path_op *)realloc(o->v, o->cap * sizeof(or_path_op));
        if (!nv) {
            o->err = OR_ENOMEM;
            return;
        }
        o->v = nv;
    }
    or_path_op op;
    memset(&op, 0, sizeof op);
    op.tag = tag;
    op.v[0] = x;
    op.v[1] = y;
    o->v[o->n++] = op;
}

static void stroke_init(stroke *s, int join, float miter, float tol_sq) /* stroker.rs:114-124 */
{
    memset(s, 0, sizeof *s);
    s->join = join;
    s->miter = miter;
    s->tol_sq = tol_sq;
    s->cap = 1024;
    s->pts = (widept *)malloc(s->cap * sizeof(widept));
    s->subcap = 16;
    s->subs = (substroke *)calloc(s->subcap, sizeof(substroke));
    s->nsubs = 1;
}

static void stroke_free(stroke *s)
{
    free(s->pts);
    free(s->subs);
}

static void stroke_add_point(stroke *s, widept p) /* stroker.rs:204-216 */
{
    if (s->n < VID_MAX) {
        int done = s->subs[s->nsubs - 1].done;
        if (done) { /* sub_add, stroker.rs:164-167 */
            if (s->nsubs == s->subcap) {
                s->subcap *= 2;
                s->subs = (substroke *)realloc(s->subs, s->subcap * sizeof(substroke));
            }
            substroke ns = {(uint16_t)s->n, 0, 0, 0};
            s->subs[s->nsubs++] = ns;
        }
        int coincident = s->n > 0 && s->pts[s->n - 1].x == p.x && s->pts[s->n - 1].y == p.y; /* 219-225 */
        if (done || !coincident) {
            if (s->n == s->cap) {
                s->cap *= 2;
                s->pts = (widept *)realloc(s->pts, s->cap * sizeof(widept));
            }
            s->pts[s->n++] = p;
            s->subs[s->nsubs - 1].n_points++;
        }
    }
}

static void stroke_close(stroke *s, int joined) /* stroker.rs:230-236 */
{
    if (s->n > 0) {
        s->subs[s->nsubs - 1].joined = joined;
        s->subs[s->nsubs - 1].done = 1;
    }
}

static uint16_t substroke_len(const substroke *u) /* stroker.rs:88-96 (u16 wraps) */
{
    if (u->joined)
        return (uint16_t)(u->n_points + 1);
    if (u->n_points > 0)
        return (uint16_t)(u->n_points - 1);
    return 0;
}

static uint16_t substroke_next(const substroke *u, uint16_t vid, int dir) /* stroker.rs:67-85 */
{
    if (dir == DIR_FWD) {
        uint16_t v = (uint16_t)(vid + 1);
        return v < (uint16_t)(u->start + u->n_points) ? v : u->start;
    }
    if (vid > u->start)
        return (uint16_t)(vid - 1);
    return (uint16_t)(u->start + u->n_points - 1);
}

static uint16_t stroke_next(stroke *s, uint16_t vid, int dir) /* stroker.rs:176-192 */
{
    for (size_t i = 0; i < s->nsubs; i++) {
        const substroke *u = &s->subs[i];
        if (vid < (uint16_t)(u->start + u->n_points))
            return substroke_next(u, vid, dir);
    }
    s->err = OR_EINVAL; /* unreachable!() in the reference */
    return vid;
}

/* pointy 0.7 Pt helpers (UNVERIFIED recollection, parity unpinned). */
static pt2 pt_right(pt2 v) { pt2 r = {-v.y, v.x}; return r; }
static pt2 pt_normalize(pt2 v)
{
    float m = hypotf(v.x, v.y);
    pt2 r = {0.0f, 0.0f};
    if (m > 0.0f) {
        r.x = v.x / m;
        r.y = v.y / m;
    }
    return r;
}
static float pt_angle_rel(pt2 a, pt2 b)
{
    const float pi = (float)M_PI;
    float th = atan2f(a.y, a.x) - atan2f(b.y, b.x);
    if (th < -pi)
        return th + pi + pi;
    if (th > pi)
        return th - pi - pi;
    return th;
}
static int line_intersection(pt2 a0, pt2 a1, pt2 b0, pt2 b1, pt2 *out)
{
    pt2 av = {a0.x - a1.x, a0.y - a1.y};
    pt2 bv = {b0.x - b1.x, b0.y - b1.y};
    float den = av.x * bv.y - av.y * bv.x;
    if (den == 0.0f)
        return 0;
    float ca = a0.x * a1.y - a0.y * a1.x;
    float cb = b0.x * b1.y - b0.y * b1.x;
    float xn = bv.x * ca - av.x * cb;
    float yn = bv.y * ca - av.y * cb;
    out->x = xn / den;
    out->y = yn / den;
    return 1;
}

static void stroke_offset(widept p0, widept p1, pt2 *r0, pt2 *r1) /* stroker.rs:301-309 */
{
    pt2 d = {p1.x - p0.x, p1.y - p0.y};
    pt2 vr = pt_normalize(pt_right(d));
    float h0 = p0.w / 2.0f, h1 = p1.w / 2.0f;
    r0->x = p0.x + vr.x * h0;
    r0->y = p0.y + vr.y * h0;
    r1->x = p1.x + vr.x * h1;
    r1->y = p1.y + vr.y * h1;
}

static void stroke_bevel(opvec *o, pt2 a1, pt2 b0) /* stroker.rs:370-373 */
{
    opvec_push(o, OR_OP_LINE, a1.x, a1.y);
    opvec_push(o, OR_OP_LINE, b0.x, b0.y);
}

static void stroke_arc(stroke *s, opvec *o, widept p, pt2 a, pt2 b, int depth) /* stroker.rs:399-416 */
{
    if (s->err)
        return;
    pt2 d = {b.x - a.x, b.y - a.y};
    pt2 vr = pt_normalize(pt_right(d));
    float h = p.w / 2.0f;
    pt2 c = {p.x + vr.x * h, p.y + vr.y * h};
    pt2 ab = {(a.x + b.x) / 2.0f, (a.y + b.y) / 2.0f};
    float dx = c.x - ab.x, dy = c.y - ab.y;
    if (dx * dx + dy * dy <= s->tol_sq) {
        opvec_push(o, OR_OP_LINE, b.x, b.y);
    } else if (depth >= OR_FLATTEN_MAX_DEPTH) {
        s->err = OR_ERANGE;
    } else {
        stroke_arc(s, o, p, a, c, depth + 1);
        stroke_arc(s, o, p, c, b, depth + 1);
    }
}

static void stroke_join(stroke *s, opvec *o, widept p, pt2 a0, pt2 a1, pt2 b0, pt2 b1) /* 323-396 */
{
    pt2 u = {a1.x - a0.x, a1.y - a0.y};
    pt2 v = {b0.x - b1.x, b0.y - b1.y};
    if (s->join == OR_JOIN_MITER) { /* stroker.rs:340-367 */
        float ml = s->miter;
        if (ml > 0.0f) {
            float sm_min = 1.0f / ml;
            float th = pt_angle_rel(u, v);
            float sm = fabsf(sinf(th / 2.0f));
            if (sm >= sm_min && sm < 1.0f) {
                pt2 xp;
                if (line_intersection(a0, a1, b0, b1, &xp)) {
                    opvec_push(o, OR_OP_LINE, xp.x, xp.y);
                    return;
                }
            }
        }
        stroke_bevel(o, a1, b0);
    } else if (s->join == OR_JOIN_BEVEL) {
        stroke_bevel(o, a1, b0);
    } else { /* stroker.rs:380-396 */
        float th = pt_angle_rel(u, v);
        if (th <= 0.0f) {
            stroke_bevel(o, a1, b0);
        } else {
            opvec_push(o, OR_OP_LINE, a1.x, a1.y);
            stroke_arc(s, o, p, a1, b0, 0);
        }
    }
}

static void stroke_side(stroke *s, opvec *o, size_t i, uint16_t start, int dir) /* 265-295 */
{
    int have_xr = 0;
    pt2 xr0 = {0, 0}, xr1 = {0, 0};
    uint16_t v0 = start;
    uint16_t v1 = stroke_next(s, v0, dir);
    int joined = s->subs[i].joined;
    uint16_t count = substroke_len(&s->subs[i]);
    for (uint32_t k = 0; k < count && !s->err; k++) {
        widept p0 = s->pts[v0], p1 = s->pts[v1];
        pt2 pr0, pr1;
        stroke_offset(p0, p1, &pr0, &pr1);
        if (have_xr)
            stroke_join(s, o, p0, xr0, xr1, pr0, pr1);
        else if (!joined)
            opvec_push(o, OR_OP_LINE, pr0.x, pr0.y);
        have_xr = 1;
        xr0 = pr0;
        xr1 = pr1;
        v0 = v1;
        v1 = stroke_next(s, v1, dir);
    }
    if (!joined && have_xr)
        opvec_push(o, OR_OP_LINE, xr1.x, xr1.y);
}

static void stroke_path_ops(stroke *s, opvec *o) /* stroker.rs:239-262 */
{
    for (size_t i = 0; i < s->nsubs && !s->err; i++) {
        if (substroke_len(&s->subs[i]) > 0) {
            uint16_t start = s->subs[i].start;
            uint16_t end = substroke_next(&s->subs[i], start, DIR_REV); /* sub_end 143-146 */
            int joined = s->subs[i].joined;
            stroke_side(s, o, i, start, DIR_FWD);
            if (joined)
                opvec_push(o, OR_OP_CLOSE, 0, 0);
            stroke_side(s, o, i, end, DIR_REV);
            opvec_push(o, OR_OP_CLOSE, 0, 0);
        }
    }
}

/* ========================================================================
 * Plotter -- src/plotter.rs
 * ======================================================================== */
struct or_plotter { /* plotter.rs:38-56 */
    uint32_t w, h;
    int fmt;
    uint8_t *px;
    int16_t *sgn;
    widept pen;
    float tr[6];
    float tol_sq;
    float s_width;
    int join;
    float miter;
    int ssse3;
    size_t last_offered, last_stored;
};

enum { DEST_FIG = 0, DEST_STROKE = 1, DEST_RECORD = 2 };
typedef struct { /* PlotDest, plotter.rs:59-87 */
    int kind;
    or_fig *fig;
    stroke *stk;
    float *rec;
    size_t rec_n, rec_cap;
    int err;
    uint32_t leaves;
} dest;

static void dest_add_point(dest *d, widept p)
{
    if (d->kind == DEST_FIG) {
        or_fig_add_point(d->fig, p.x, p.y);
    } else if (d->kind == DEST_STROKE) {
        stroke_add_point(d->stk, p);
    } else {
        if (d->rec && d->rec_n < d->rec_cap) {
            d->rec[3 * d->rec_n] = p.x;
            d->rec[3 * d->rec_n + 1] = p.y;
            d->rec[3 * d->rec_n + 2] = p.w;
        }
        d->rec_n++;
    }
}

static void dest_close(dest *d, int joined)
{
    if (d->kind == DEST_FIG) {
        or_fig_close(d->fig);
    } else if (d->kind == DEST_STROKE) {
        stroke_close(d->stk, joined);
    } else {
        if (d->rec && d->rec_n < d->rec_cap) {
            d->rec[3 * d->rec_n] = NAN;
            d->rec[3 * d->rec_n + 1] = NAN;
            d->rec[3 * d->rec_n + 2] = joined ? 1.0f : 0.0f;
        }
        d->rec_n++;
    }
}

or_plotter *or_plotter_new(uint32_t w, uint32_t h, int fmt) /* plotter.rs:96-115 */
{
    or_plotter *p = (or_plotter *)calloc(1, sizeof(or_plotter));
    if (!p)
        return NULL;
    p->w = w;
    p->h = h;
    p->fmt = fmt;
    p->px = (uint8_t *)calloc((size_t)w * h * fmt_bpp(fmt) + 1, 1);
    p->sgn = (int16_t *)calloc((size_t)w + 8, sizeof(int16_t));
    p->pen.x = 0.0f;
    p->pen.y = 0.0f;
    p->pen.w = 1.0f;
    p->tr[0] = 1.0f;
    p->tr[4] = 1.0f;
    const float tol = 0.3f;
    p->tol_sq = tol * tol;
    p->s_width = 1.0f;
    p->join = OR_JOIN_MITER;
    p->miter = 4.0f;
    p->ssse3 = 1; /* Cargo feature "simd" is on by default (Cargo.toml:26-28) */
    return p;
}

void or_plotter_free(or_plotter *p)
{
    if (!p)
        return;
    free(p->px);
    free(p->sgn);
    free(p);
}

void or_plotter_set_tolerance(or_plotter *p, float t) /* plotter.rs:133-137 */
{
    float tol = t > 0.01f ? t : 0.01f; /* f32::max: NaN yields the other operand */
    if (t != t)
        tol = 0.01f;
    p->tol_sq = tol * tol;
}

void or_plotter_set_transform(or_plotter *p, const float e[6]) { memcpy(p->tr, e, sizeof p->tr); }
void or_plotter_set_join(or_plotter *p, int kind, float miter) { p->join = kind; p->miter = miter; }
void or_plotter_set_ssse3(or_plotter *p, int on) { p->ssse3 = on; }
float or_plotter_s_width(const or_plotter *p) { return p->s_width; }
uint8_t *or_plotter_pixels(or_plotter *p) { return p->px; }
void or_plotter_clear(or_plotter *p) { memset(p->px, 0, (size_t)p->w * p->h * fmt_bpp(p->fmt)); }

void or_plotter_last_counts(const or_plotter *p, size_t *offered, size_t *stored)
{
    if (offered)
        *offered = p->last_offered;
    if (stored)
        *stored = p->last_stored;
}

static void reset_pen(or_plotter *p) /* plotter.rs:128-130 */
{
    p->pen.x = 0.0f;
    p->pen.y = 0.0f;
    p->pen.w = p->s_width;
}

static widept transform_point(const or_plotter *p, widept q) /* plotter.rs:169-172 */
{
    widept r;
    r.x = p->tr[0] * q.x + p->tr[1] * q.y + p->tr[2];
    r.y = p->tr[3] * q.x + p->tr[4] * q.y + p->tr[5];
    r.w = q.w;
    return r;
}

static widept wmid(widept a, widept b) /* geom.rs:31-35 */
{
    widept r;
    r.x = (a.x + b.x) / 2.0f;
    r.y = (a.y + b.y) / 2.0f;
    r.w = (a.w + b.w) / 2.0f;
    return r;
}

static int within_tol(const or_plotter *p, widept a, widept b) /* plotter.rs:268-276 */
{
    float dx = a.x - b.x, dy = a.y - b.y;
    return dx * dx + dy * dy <= p->tol_sq;
}

static void flat_leaf(dest *d, widept pt)
{
    if (++d->leaves > OR_FLATTEN_MAX_LEAVES) {
        d->err = OR_ERANGE;
        return;
    }
    dest_add_point(d, pt);
}

static void quad_tran(const or_plotter *p, dest *d, widept a, widept b, widept c, int depth) /* 248-265 */
{
    if (d->err)
        return;
    widept ab = wmid(a, b), bc = wmid(b, c);
    widept ab_bc = wmid(ab, bc), ac = wmid(a, c);
    if (within_tol(p, ab_bc, ac)) {
        flat_leaf(d, c);
    } else if (depth >= OR_FLATTEN_MAX_DEPTH) {
        d->err = OR_ERANGE;
    } else {
        quad_tran(p, d, a, ab, ab_bc, depth + 1);
        quad_tran(p, d, ab_bc, bc, c, depth + 1);
    }
}

static void cubic_tran(const or_plotter *p, dest *d, widept pa, widept pb, widept pc, widept pd,
                       int depth) /* plotter.rs:311-332 */
{
    if (d->err)
        return;
    widept ab = wmid(pa, pb), bc = wmid(pb, pc), cd = wmid(pc, pd);
    widept ab_bc = wmid(ab, bc), bc_cd = wmid(bc, cd);
    widept pe = wmid(ab_bc, bc_cd);
    widept ad = wmid(pa, pd);
    if (within_tol(p, pe, ad)) {
        flat_leaf(d, pd);
    } else if (depth >= OR_FLATTEN_MAX_DEPTH) {
        d->err = OR_ERANGE;
    } else {
        cubic_tran(p, d, pa, ab, ab_bc, pe, depth + 1);
        cubic_tran(p, d, pe, bc_cd, cd, pd, depth + 1);
    }
}

static float float_lerp(float a, float b, float t) { return b + (a - b) * t; } /* geom.rs:14-16 */

static void add_ops(or_plotter *p, const or_path_op *ops, size_t n, dest *d) /* plotter.rs:175-224 */
{
    reset_pen(p);
    for (size_t i = 0; i < n && !d->err; i++) {
        const or_path_op *op = &ops[i];
        widept q;
        switch (op->tag) {
        case OR_OP_CLOSE: /* plotter.rs:200-203 */
            dest_close(d, 1);
            reset_pen(p);
            break;
        case OR_OP_MOVE: /* plotter.rs:208-214 */
            q.x = op->v[0];
            q.y = op->v[1];
            q.w = p->s_width;
            dest_close(d, 0);
            dest_add_point(d, transform_point(p, q));
            p->pen = q;
            break;
        case OR_OP_LINE: /* plotter.rs:219-224 */
            q.x = op->v[0];
            q.y = op->v[1];
            q.w = p->s_width;
            dest_add_point(d, transform_point(p, q));
            p->pen = q;
            break;
        case OR_OP_QUAD: { /* plotter.rs:233-242 */
            widept bb = {op->v[0], op->v[1], (p->pen.w + p->s_width) / 2.0f};
            widept cc = {op->v[2], op->v[3], p->s_width};
            d->leaves = 0;
            quad_tran(p, d, transform_point(p, p->pen), transform_point(p, bb), transform_point(p, cc), 0);
            p->pen = cc;
            break;
        }
        case OR_OP_CUBIC: { /* plotter.rs:286-305 */
            float w0 = float_lerp(p->pen.w, p->s_width, 1.0f / 3.0f);
            float w1 = float_lerp(p->pen.w, p->s_width, 2.0f / 3.0f);
            widept bb = {op->v[0], op->v[1], w0};
            widept cc = {op->v[2], op->v[3], w1};
            widept dd = {op->v[4], op->v[5], p->s_width};
            d->leaves = 0;
            cubic_tran(p, d, transform_point(p, p->pen), transform_point(p, bb), transform_point(p, cc),
                       transform_point(p, dd), 0);
            p->pen = dd;
            break;
        }
        case OR_OP_PEN_WIDTH: /* plotter.rs:151-153 */
            p->s_width = op->v[0];
            break;
        default:
            d->err = OR_EINVAL;
            break;
        }
    }
}

int or_plotter_fill(or_plotter *p, int rule, const or_path_op *ops, size_t n, const uint8_t clr[4])
{ /* plotter.rs:339-350 */
    dest d;
    memset(&d, 0, sizeof d);
    d.kind = DEST_FIG;
    d.fig = or_fig_new();
    if (!d.fig)
        return OR_ENOMEM;
    add_ops(p, ops, n, &d);
    if (d.err) {
        or_fig_free(d.fig);
        return d.err;
    }
    or_fig_close(d.fig);
    p->last_offered = d.fig->offered;
    p->last_stored = d.fig->n;
    int rc = or_fig_fill(d.fig, rule, p->px, p->w, p->h, p->fmt, clr, p->sgn, p->w, p->ssse3);
    or_fig_free(d.fig);
    return rc;
}

static int stroke_to_ops(or_plotter *p, const or_path_op *ops, size_t n, opvec *o)
{
    stroke s;
    stroke_init(&s, p->join, p->miter, p->tol_sq);
    dest d;
    memset(&d, 0, sizeof d);
    d.kind = DEST_STROKE;
    d.stk = &s;
    add_ops(p, ops, n, &d);
    int rc = d.err;
    if (!rc) {
        stroke_path_ops(&s, o);
        rc = s.err ? s.err : o->err;
    }
    stroke_free(&s);
    return rc;
}

int or_plotter_stroke(or_plotter *p, const or_path_op *ops, size_t n, const uint8_t clr[4])
{ /* plotter.rs:356-365 */
    opvec o;
    memset(&o, 0, sizeof o);
    int rc = stroke_to_ops(p, ops, n, &o);
    if (!rc)
        rc = or_plotter_fill(p, OR_NON_ZERO, o.v, o.n, clr);
    free(o.v);
    return rc;
}

int or_plotter_stroke_ops(or_plotter *p, const or_path_op *ops, size_t n, or_path_op *out, size_t cap,
                          size_t *count)
{
    opvec o;
    memset(&o, 0, sizeof o);
    int rc = stroke_to_ops(p, ops, n, &o);
    if (!rc) {
        if (out)
            memcpy(out, o.v, (o.n < cap ? o.n : cap) * sizeof(or_path_op));
        if (count)
            *count = o.n;
    }
    free(o.v);
    return rc;
}

int or_plotter_flatten(or_plotter *p, const or_path_op *ops, size_t n, float *xyw, size_t cap,
                       size_t *count)
{
    dest d;
    memset(&d, 0, sizeof d);
    d.kind = DEST_RECORD;
    d.rec = xyw;
    d.rec_cap = cap;
    add_ops(p, ops, n, &d);
    if (count)
        *count = d.rec_n;
    return d.err;
}
