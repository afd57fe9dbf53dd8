/*
 * footile_oracle.h -- CPU oracle for the footile fill/stroke path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of footile 0.7.0's
 * sequential algorithm (Rust, DougLau/footile), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product (footile_amd/, libfootile_hip.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Semantics are the
 * reference's *release* build: i16/i32 arithmetic wraps, `as` casts from f32
 * saturate (NaN -> 0), no FMA contraction (built with -ffp-contract=off).
 *
 * Parity status per part (what pins it):
 *   Fixed arithmetic ......... pinned: src/fixed.rs:166-313 known answers.
 *   Fig scan + accumulate .... pinned: src/fig.rs:691-794 and
 *                              src/imgbuf.rs:205-232 known answers, plus
 *                              closed-form shapes in tests/.
 *   Rgba8p composite ......... parity unpinned beyond fig_3x3 (pix 0.14 Ch8
 *                              multiply / SrcOver restated from recollection).
 *   Flattening ............... pinned by invariants only (collinear cubic ->
 *                              1 segment, symmetric leaf counts, DFS order);
 *                              bit level of pointy midpoint/distance_sq is
 *                              parity unpinned.
 *   Stroker .................. parity unpinned (pointy right/normalize/
 *                              angle_rel/Line::intersection restated from
 *                              recollection; libm sinf/atan2f/hypotf called).
 */
#ifndef FOOTILE_ORACLE_H
#define FOOTILE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Fixed 16.16 (src/fixed.rs) ---------------------------------------- */
int32_t or_fx_from_i32(int32_t i);           /* fixed.rs:74-79  */
int32_t or_fx_to_i32(int32_t f);             /* fixed.rs:81-86  */
int32_t or_fx_from_f32(float f);             /* fixed.rs:88-93  */
float or_fx_to_f32(int32_t f);               /* fixed.rs:95-100 */
int32_t or_fx_add(int32_t a, int32_t b);     /* fixed.rs:24-30  */
int32_t or_fx_sub(int32_t a, int32_t b);     /* fixed.rs:32-38  */
int32_t or_fx_mul(int32_t a, int32_t b);     /* fixed.rs:40-47  */
int32_t or_fx_div(int32_t a, int32_t b);     /* fixed.rs:49-56  */
int32_t or_fx_shl(int32_t a, uint32_t s);    /* fixed.rs:58-64  */
int32_t or_fx_shr(int32_t a, uint32_t s);    /* fixed.rs:66-72  */
int32_t or_fx_abs(int32_t a);                /* fixed.rs:122-124 */
int32_t or_fx_floor(int32_t a);              /* fixed.rs:127-129 */
int32_t or_fx_ceil(int32_t a);               /* fixed.rs:132-134 */
int32_t or_fx_round(int32_t a);              /* fixed.rs:137-139 */
int32_t or_fx_trunc(int32_t a);              /* fixed.rs:142-148 */
int32_t or_fx_fract(int32_t a);              /* fixed.rs:151-153 */
int32_t or_fx_avg(int32_t a, int32_t b);     /* fixed.rs:156-158 */
int or_fx_widdershins(int32_t ax, int32_t ay, int32_t bx, int32_t by); /* fig.rs:116-119 */
int16_t or_pixel_cov(int32_t fcov);          /* fig.rs:677-682 */

/* ---- Row accumulate (src/imgbuf.rs) ------------------------------------ */
/* Scalar and SSSE3 restatements; both zero `src[0..n)` and write `dst[0..n)`. */
void or_accumulate_non_zero_scalar(uint8_t *dst, int16_t *src, size_t n); /* imgbuf.rs:54-61 */
void or_accumulate_non_zero_ssse3(uint8_t *dst, int16_t *src, size_t n);  /* imgbuf.rs:71-117 */
void or_accumulate_even_odd_scalar(uint8_t *dst, int16_t *src, size_t n); /* imgbuf.rs:157-167 */
void or_accumulate_even_odd_ssse3(uint8_t *dst, int16_t *src, size_t n);  /* imgbuf.rs:172-199 */

/* ---- pix subset ---------------------------------------------------------- */
enum { OR_MATTE8 = 0, OR_RGBA8P = 1, OR_GRAYA8P = 2 };
uint8_t or_ch8_mul(uint8_t l, uint8_t r);    /* pix 0.14 Ch8 Mul (unverified recollection) */

/* ---- Path ops (src/path.rs:18-31) --------------------------------------- */
enum {
    OR_OP_CLOSE = 0, OR_OP_MOVE = 1, OR_OP_LINE = 2,
    OR_OP_QUAD = 3, OR_OP_CUBIC = 4, OR_OP_PEN_WIDTH = 5
};
typedef struct { uint32_t tag; float v[6]; } or_path_op;

enum { OR_JOIN_MITER = 0, OR_JOIN_BEVEL = 1, OR_JOIN_ROUND = 2 };
enum { OR_NON_ZERO = 0, OR_EVEN_ODD = 1 };
enum { OR_OK = 0, OR_EINVAL = 1, OR_ENOMEM = 2, OR_ERANGE = 4 };

/* Flattening limits shared with the GPU path: recursion deeper than this, or
 * more leaves than this from one curve, is reported as OR_ERANGE (the
 * reference recurses without bound and overflows its stack on NaN). */
#define OR_FLATTEN_MAX_DEPTH 32
#define OR_FLATTEN_MAX_LEAVES (1u << 20)

/* ---- Fig (src/fig.rs) ---------------------------------------------------- */
typedef struct or_fig or_fig;
or_fig *or_fig_new(void);                                   /* fig.rs:339-344 */
void or_fig_free(or_fig *f);
void or_fig_add_point(or_fig *f, float x, float y);         /* fig.rs:428-442 */
void or_fig_close(or_fig *f);                               /* fig.rs:457-461 */
size_t or_fig_points(const or_fig *f, int32_t *xy, size_t cap); /* stored Fixed points */
/* Fig::fill into a caller raster (fig.rs:480-502).  sgn has sgn_len >= w
 * entries of persistent signed area (zero on entry, zero on exit). */
int or_fig_fill(const or_fig *f, int rule, uint8_t *px, uint32_t w, uint32_t h,
                int fmt, const uint8_t clr[4], int16_t *sgn, size_t sgn_len,
                int use_ssse3);

/* ---- Plotter (src/plotter.rs) -------------------------------------------- */
typedef struct or_plotter or_plotter;
or_plotter *or_plotter_new(uint32_t w, uint32_t h, int fmt);   /* plotter.rs:96-115 */
void or_plotter_free(or_plotter *p);
void or_plotter_set_tolerance(or_plotter *p, float t);         /* plotter.rs:133-137 */
void or_plotter_set_transform(or_plotter *p, const float e[6]);/* plotter.rs:140-143 */
void or_plotter_set_join(or_plotter *p, int kind, float miter);/* plotter.rs:158-161 */
void or_plotter_set_ssse3(or_plotter *p, int on);
float or_plotter_s_width(const or_plotter *p);
uint8_t *or_plotter_pixels(or_plotter *p);
void or_plotter_clear(or_plotter *p);
int or_plotter_fill(or_plotter *p, int rule, const or_path_op *ops, size_t n,
                    const uint8_t clr[4]);                     /* plotter.rs:339-350 */
int or_plotter_stroke(or_plotter *p, const or_path_op *ops, size_t n,
                      const uint8_t clr[4]);                   /* plotter.rs:356-365 */
/* Flattening only: the points Fig::add_point would receive, in order, with
 * close markers as NaN pairs.  Returns the count written (or needed). */
int or_plotter_flatten(or_plotter *p, const or_path_op *ops, size_t n,
                       float *xyw, size_t cap, size_t *count);
/* Stroke outline only: the PathOps Stroke::path_ops returns (stroker.rs:239). */
int or_plotter_stroke_ops(or_plotter *p, const or_path_op *ops, size_t n,
                          or_path_op *out, size_t cap, size_t *count);
/* Fig statistics of the last fill: points offered to add_point before the
 * 65,535 cap, points stored. */
void or_plotter_last_counts(const or_plotter *p, size_t *offered, size_t *stored);

#ifdef __cplusplus
}
#endif
#endif
