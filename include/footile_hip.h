/*
 * footile_hip.h -- C ABI of the MI355X (gfx950) footile fill/stroke rasteriser.
 *
 * This is the drop-in boundary for footile's hot path: `Plotter::fill` and
 * `Plotter::stroke` over a pix raster (Matte8 / Rgba8p / Graya8p).  Each entry
 * point names the reference interface it replaces.  Plain pointers and sizes
 * only; every function returns an ft_status and never aborts.  A handle is used
 * by one host thread at a time and owns one HIP stream of default priority.
 *
 * Reference: footile 0.7.0 (Rust).  Paths below are relative to its tree.
 */
#ifndef FOOTILE_HIP_H
#define FOOTILE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FT_OK = 0,
    FT_EINVAL = 1,   /* null handle/pointer, bad tag, bad enum value */
    FT_ENOMEM = 2,   /* host or device allocation failed */
    FT_EDEVICE = 3,  /* HIP runtime error */
    FT_ERANGE = 4,   /* curve flattening deeper than FT_FLATTEN_MAX_DEPTH or with
                        more than FT_FLATTEN_MAX_LEAVES leaves (the reference
                        recurses without bound, src/plotter.rs:248-332) */
    FT_ENOTSUP = 5   /* raster wider than FT_MAX_WIDTH */
} ft_status;

/* Pixel formats of pix 0.14 usable with Plotter<P> (src/plotter.rs:89-92). */
typedef enum {
    FT_MATTE8 = 0,  /* 1 byte: alpha.  Fill overwrites rows (src/fig.rs:632-635) */
    FT_RGBA8P = 1,  /* 4 bytes: premultiplied R,G,B,A; SrcOver composite */
    FT_GRAYA8P = 2  /* 2 bytes: premultiplied value, alpha; SrcOver composite */
} ft_format;

/* FillRule, src/path.rs:8-14 */
typedef enum { FT_NON_ZERO = 0, FT_EVEN_ODD = 1 } ft_fill_rule;

/* JoinStyle, src/stroker.rs:13-20 */
typedef enum { FT_JOIN_MITER = 0, FT_JOIN_BEVEL = 1, FT_JOIN_ROUND = 2 } ft_join;

/* PathOp, src/path.rs:18-31.  28 bytes, absolute coordinates:
 *   CLOSE      -
 *   MOVE       v[0..1] = point
 *   LINE       v[0..1] = end point
 *   QUAD       v[0..1] = control, v[2..3] = end
 *   CUBIC      v[0..1] = control 0, v[2..3] = control 1, v[4..5] = end
 *   PEN_WIDTH  v[0]    = width                                            */
typedef enum {
    FT_OP_CLOSE = 0, FT_OP_MOVE = 1, FT_OP_LINE = 2,
    FT_OP_QUAD = 3, FT_OP_CUBIC = 4, FT_OP_PEN_WIDTH = 5
} ft_op_tag;

typedef struct {
    uint32_t tag;
    float v[6];
} ft_path_op;

#define FT_FLATTEN_MAX_DEPTH 32
#define FT_FLATTEN_MAX_LEAVES (1u << 20)
#define FT_MAX_POINTS 65535u /* Vid::MAX, src/vid.rs:17 */
#define FT_MAX_WIDTH 32768u /* a raster row's i32 cells (128 KiB at this width) are
                              held in LDS at once; a device whose LDS per block is
                              smaller rejects wide rows with FT_ENOTSUP when the
                              batch is laid out */

/* ---------------------------------------------------------------------------
 * Plotter -- replaces footile::Plotter<P> (src/plotter.rs:38-381).
 * The raster lives in device memory; read/write copy it to/from the host.
 * ------------------------------------------------------------------------- */
typedef struct ft_plotter ft_plotter;

/* Plotter::new(Raster::with_clear(w, h)) -- src/plotter.rs:96-115 */
ft_status ft_plotter_new(uint32_t width, uint32_t height, ft_format fmt, ft_plotter **out);
/* Plotter::new(Raster::with_pixels(w, h, px)) -- src/plotter.rs:96-115 */
ft_status ft_plotter_new_with_pixels(uint32_t width, uint32_t height, ft_format fmt,
                                     const void *pixels, ft_plotter **out);
/* drop(Plotter) */
void ft_plotter_free(ft_plotter *p);
/* Plotter::width / height -- src/plotter.rs:118-125 */
ft_status ft_plotter_size(const ft_plotter *p, uint32_t *width, uint32_t *height);
/* Plotter::set_tolerance -- src/plotter.rs:133-137 */
ft_status ft_plotter_set_tolerance(ft_plotter *p, float t);
/* Plotter::set_transform -- src/plotter.rs:140-143; e = pointy Transform
 * elements: x' = e0*x + e1*y + e2, y' = e3*x + e4*y + e5. */
ft_status ft_plotter_set_transform(ft_plotter *p, const float e[6]);
/* Plotter::set_join -- src/plotter.rs:158-161 (miter_limit used by MITER) */
ft_status ft_plotter_set_join(ft_plotter *p, ft_join kind, float miter_limit);
/* Plotter::fill -- src/plotter.rs:339-350.  clr: the pixel value (Matte8
 * ignores it, as the reference does). */
ft_status ft_plotter_fill(ft_plotter *p, ft_fill_rule rule, const ft_path_op *ops, size_t n,
                          const uint8_t clr[4]);
/* Plotter::stroke -- src/plotter.rs:356-365 */
ft_status ft_plotter_stroke(ft_plotter *p, const ft_path_op *ops, size_t n, const uint8_t clr[4]);
/* Plotter::raster / into_raster: copy the raster to host memory. */
ft_status ft_plotter_read(const ft_plotter *p, void *dst, size_t bytes);
/* Plotter::raster_mut: overwrite the raster from host memory. */
ft_status ft_plotter_write(ft_plotter *p, const void *src, size_t bytes);
/* Zero-copy access for device consumers: raster base and row pitch (bytes). */
ft_status ft_plotter_device_ptr(ft_plotter *p, void **dptr, size_t *pitch);
/* The stroke width state that persists across calls (src/plotter.rs:53,151). */
ft_status ft_plotter_pen_width(const ft_plotter *p, float *width);
/* Figure statistics of the last fill: points offered to Fig::add_point and
 * points stored after the 65,535-point cap (src/fig.rs:428-442). */
ft_status ft_plotter_last_counts(const ft_plotter *p, uint64_t *offered, uint64_t *stored);

/* ---------------------------------------------------------------------------
 * Fig -- debug entry for footile's crate-private Fig (src/fig.rs:339-502), so
 * the reference's Fig-level known answers replay on the GPU.  Points are
 * Fig::add_point calls; close_after[i] != 0 calls Fig::close after point i.
 * The raster is host memory, updated in place.
 * ------------------------------------------------------------------------- */
ft_status ft_fig_fill(ft_fill_rule rule, const float *xy, const uint8_t *close_after, size_t n,
                      ft_format fmt, uint32_t width, uint32_t height, void *pixels,
                      const uint8_t clr[4]);

/* ---------------------------------------------------------------------------
 * Batch -- many independent fills in one device-resident launch sequence.
 * No reference counterpart: it runs N Plotter::fill calls (src/plotter.rs:
 * 339-350), each into its own raster region, with ops uploaded once.
 *
 * Concurrency: the prefix scans of a batch or plotter run hand out their
 * work in block order over a grid the device holds at once, which assumes no
 * other such run's scans wait beside them.  Within a process the library
 * makes such runs of different batches and plotters take turns on a device
 * (a launch first waits on the host for the other one's stream), so they do
 * not overlap; to overlap launches, pipeline one batch (ft_batch_set_pipeline).
 * Pipelined batches and ft_batch_run_noalloc take arrival tickets instead and
 * carry no such assumption.  Should another process's work starve a run's
 * scan blocks anyway, the run fails with FT_EDEVICE after a bounded wait
 * instead of hanging.
 * ------------------------------------------------------------------------- */
typedef struct {
    uint64_t op_begin;   /* first op of this path in the batch op array */
    uint32_t op_count;
    uint32_t rule;       /* ft_fill_rule */
    uint64_t dst;        /* device address of the raster's first pixel */
    uint32_t width, height;
    uint32_t pitch;      /* bytes between rows */
    uint32_t fmt;        /* ft_format */
    uint8_t clr[4];
    float transform[6];
    float tolerance;     /* as given to set_tolerance (floored at 0.01) */
    float pen_width;     /* s_width at the start of the fill */
    uint32_t fresh;      /* 1: region starts cleared (rows above the path are
                            written as zeros, as a fresh Raster::with_clear) */
} ft_fill_desc;

typedef struct ft_batch ft_batch;

ft_status ft_batch_new(const ft_path_op *ops, size_t n_ops, const ft_fill_desc *paths,
                       size_t n_paths, ft_batch **out);
void ft_batch_free(ft_batch *b);
/* Enqueue the whole batch on the batch's stream (asynchronous).  The
 * launch's status stays on the device until the next ft_batch_sync; a launch
 * that fails leaves the launches after it no-ops until then. */
ft_status ft_batch_launch(ft_batch *b);
/* Wait for the launches since the last sync; returns their status (FT_ERANGE
 * etc., FT_ENOMEM when a capacity ran out: ft_batch_run grows it). */
ft_status ft_batch_sync(ft_batch *b);
/* Launch + sync; grows workspace and relaunches once if a capacity ran out. */
ft_status ft_batch_run(ft_batch *b);
/* Per-path results of the last run (arrays of n_paths, any may be NULL):
 * points offered to Fig::add_point, points stored, non-horizontal edges, and
 * the figure's top row (row_of of the smallest y; INT32_MAX if empty). */
ft_status ft_batch_counts(ft_batch *b, uint64_t *offered, uint64_t *stored, uint64_t *edges,
                          int32_t *top_rows);
/* Raster geometry of a batch: row bands of at least band_rows_min rows
 * (1..256) are the unit of edge binning, and each band is rasterised in
 * groups of rows whose i32 cells fit lds_cells (1024..32768, a multiple of
 * 16, within the device's LDS per block); one block per group.  A path of at
 * most tile_groups such groups (0: none) skips the bands: the atlas tile
 * scheduler, one persistent launch per pixel format, rasters its groups from
 * the path's own edges, largest work items first.  Setting it re-lays the
 * batch out.  No reference counterpart: the autotuner (footile_amd.tune)
 * picks it per workload. */
typedef struct ft_tuning {
    uint32_t band_rows_min;
    uint32_t lds_cells;
    uint32_t dst_prefetch; /* 1: raster blocks of Rgba8p/Graya8p read their
                              destination before the cell pass */
    uint32_t tile_groups;
} ft_tuning;
ft_status ft_batch_set_tuning(ft_batch *b, const ft_tuning *t);
ft_status ft_batch_get_tuning(const ft_batch *b, ft_tuning *t);
/* Graph mode: ft_batch_launch captures the launch sequence into a HIP graph
 * once and replays it (re-captured after any re-layout or capacity growth);
 * the raster kernel is then not timed separately. */
ft_status ft_batch_set_graph(ft_batch *b, int on);
/* Pipelined launches (no counterpart in the reference, which renders one fill
 * at a time): with depth 2 to 4 the batch keeps depth workspaces of the same
 * layout, and consecutive ft_batch_launch calls take them in turn, so that
 * the front ends of later launches (flattening through binning, each on a
 * stream of its own) overlap each other and the raster of an earlier one.  The
 * rasters run on the batch's stream in launch order, so every launch blends
 * onto the result of the one before, as without pipelining.  Depth 0 or 1:
 * one workspace, no overlap.  Each extra workspace costs a set of device
 * buffers; not combinable with graph mode (FT_EINVAL). */
ft_status ft_batch_set_pipeline(ft_batch *b, int depth);
/* Work statistics of the raster: runs the batch once and returns the
 * (edge, row) pairs its cell pass walked and the nonzero signed-area cell
 * deltas it added (one LDS atomic each). */
ft_status ft_batch_raster_stats(ft_batch *b, uint64_t *pairs, uint64_t *cell_adds);
/* The batch's HIP stream (hipStream_t) for event timing by the caller. */
ft_status ft_batch_stream(ft_batch *b, void **stream);
/* Event timing on the batch stream: when on, every launch records HIP events
 * around the whole run and around the raster kernel; ft_batch_timing returns
 * the sums (ms) and the number of timed runs, and resets them. */
ft_status ft_batch_set_timing(ft_batch *b, int on);
ft_status ft_batch_timing(ft_batch *b, double *total_ms, double *raster_ms, uint64_t *runs);
/* Batch fill without allocation.  No footile counterpart: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The same fills as ft_batch_new +
 * ft_batch_run + ft_batch_counts (offered / stored may be NULL), from host
 * ops and descriptors into the device rasters the descriptors name, with
 * every device buffer the kernels need beyond those carved from one fixed
 * array of FT_NOALLOC_SCRATCH_BYTES per device, allocated when the library's
 * code loads on that device and held until the process ends.  No hipMalloc,
 * hipFree or pinned host allocation.  Runs on `stream` (a hipStream_t; NULL:
 * the null stream) and returns when the batch is done.  More than
 * FT_NOALLOC_MAX_OPS ops or FT_NOALLOC_MAX_PATHS paths, a layout that does
 * not fit the scratch, or a run whose events or band pairs outgrow it:
 * FT_ENOMEM.  Calls on one device share its scratch, which keeps the last
 * call's contents (each run clears what must start at zero); calls from
 * threads of this process take turns on a lock, and other users of the same
 * array on that device would race. */
#define FT_NOALLOC_SCRATCH_BYTES (64u << 20)
#define FT_NOALLOC_MAX_OPS (1u << 15)
#define FT_NOALLOC_MAX_PATHS (1u << 14)
ft_status ft_batch_run_noalloc(const ft_path_op *ops, size_t n_ops, const ft_fill_desc *paths, size_t n_paths,
                               void *stream, uint64_t *offered, uint64_t *stored);
/* Device allocation helpers for callers without a runtime of their own. */
ft_status ft_device_alloc(size_t bytes, void **dptr);
ft_status ft_device_free(void *dptr);
ft_status ft_memcpy_h2d(void *dst, const void *src, size_t bytes);
ft_status ft_memcpy_d2h(void *dst, const void *src, size_t bytes);
ft_status ft_memset_d(void *dst, int value, size_t bytes);

/* ------------------------------------------------------------------------
 * Differentiable fill.  No footile counterpart: footile's fill is not
 * differentiable.  A float32 analytic-coverage forward and a hand-written
 * adjoint over a fixed scene layout, for callers that optimise path control
 * points and colours (the forward/backward pair of an autograd function).
 *
 * Layout: n_images images of width x height premultiplied RGBA float32
 * pixels; n_paths paths in composite order (path_image non-decreasing), path
 * k having path_cubics[k] >= 1 cubics.  Parameters (device float32):
 * points = per path (1 + 3 n) (x, y) pairs back to back (start point, then
 * 3 control points per cubic); colours = [n_paths][4] premultiplied RGBA.
 * Each cubic is flattened at t = j/16 (j = 1..16) and the outline closes back
 * to the start point; a pixel's coverage is min(|A|, 1), A the integral of
 * the winding number over the pixel square; paths composite SrcOver in order
 * over a transparent canvas.  canvas / target / grad_canvas are
 * [n_images][height][width][4].  Every call is synchronous.
 * ------------------------------------------------------------------------ */
typedef struct ft_diff ft_diff;
/* ft_diff_layout.flags */
#define FT_DIFF_KEEP_COVERAGE 1u /* opt-in: keep each path's coverage from the
                                    forward for the backward (4 B per path and
                                    pixel, see ft_diff_backward) instead of
                                    recomputing it */
typedef struct ft_diff_layout {
    uint32_t n_images, width, height;
    uint32_t n_paths;
    const uint32_t *path_image;  /* host [n_paths] */
    const uint32_t *path_cubics; /* host [n_paths] */
    uint32_t flags;              /* 0 or FT_DIFF_KEEP_COVERAGE; other bits: FT_EINVAL */
} ft_diff_layout;
ft_status ft_diff_new(const ft_diff_layout *layout, ft_diff **out);
void ft_diff_free(ft_diff *d);
ft_status ft_diff_info(const ft_diff *d, uint64_t *n_points, uint64_t *n_paths, uint64_t *canvas_floats);
/* canvas = render(points, colours) */
ft_status ft_diff_forward(ft_diff *d, const float *points, const float *colours, float *canvas);
/* loss = mean((canvas - target)^2); grad_canvas (may be NULL) = d loss / d canvas */
ft_status ft_diff_loss(ft_diff *d, const float *canvas, const float *target, float *grad_canvas, double *loss);
/* Vector-Jacobian product: grad_canvas -> d/d points, d/d colours.  Replays
 * the forward into canvas checkpoints every 8 paths (16 B per pixel of every
 * image per checkpoint, max_paths / 8 rounded up of them, max_paths the most
 * paths of one image) and recomputes each path's coverage while it walks the
 * paths backwards.  With FT_DIFF_KEEP_COVERAGE (and when the buffer fits in
 * 16 GiB) the replay also keeps every path's coverage, max_paths x 4 B per
 * pixel of every image, and the walk reads it instead. */
ft_status ft_diff_backward(ft_diff *d, const float *points, const float *colours, const float *grad_canvas,
                           float *grad_points, float *grad_colours);
/* forward + loss + backward (one optimisation step's gradient) with one host
 * wait: the forward leaves the checkpoints and geometry the backward needs,
 * so the backward does not replay it, and the loss is reduced on the device */
ft_status ft_diff_step(ft_diff *d, const float *points, const float *colours, const float *target, float *canvas,
                       float *grad_canvas, float *grad_points, float *grad_colours, double *loss);

/* Library version string. */
const char *ft_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FOOTILE_HIP_H */
