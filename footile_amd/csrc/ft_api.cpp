// ft_api.cpp -- the C ABI (include/footile_hip.h) over the device pipeline.
//
// ft_plotter mirrors footile::Plotter<P> (src/plotter.rs:38-381): it owns its
// raster (here in device memory), its tolerance, transform, join style and
// the stroke width that persists across calls.  fill and stroke run the whole
// rasteriser on the GPU, stroke's outline construction (src/stroker.rs)
// included: it runs between the device flattening and the device fill.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <string.h>

#include <new>
#include <vector>

#include "../../include/footile_hip.h"
#include "ft_pipeline.h"

using ft::Workspace;

struct ft_plotter {
    Workspace ws;
    uint32_t width = 0, height = 0;
    ft_format fmt = FT_MATTE8;
    uint32_t bpp = 1;
    size_t pitch = 0;
    uint8_t *d_px = nullptr;
    float tolerance = 0.3f;  // plotter.rs:97
    float transform[6] = {1, 0, 0, 0, 1, 0};
    float s_width = 1.0f;    // plotter.rs:112
    ft_join join = FT_JOIN_MITER;
    float miter = 4.0f;      // plotter.rs:113
    uint64_t last_offered = 0, last_stored = 0;
};

struct ft_batch {
    Workspace ws;
    // Pipelined launches (ft_batch_set_pipeline): ws and depth - 1 twins
    // hold the same layout and take turns; every raster runs on ws.stream,
    // in launch order.
    static constexpr int PIPE_MAX = 4;
    Workspace twin[PIPE_MAX - 1];
    int depth = 1;
    int turn = 0;                  // the slot of the next launch (0: ws)
    bool live[PIPE_MAX] = {};      // launched since the last sync
    Workspace &slot(int k) { return k ? twin[k - 1] : ws; }
    size_t n_paths = 0;
    std::vector<ft_path_op> ops;  // host copies, for a re-layout after a tuning change
    std::vector<ft_fill_desc> descs;
};

// The twins of a pipelined batch take over the main workspace's capacities
// and kept-flag form after the main one grew or changed them.
static ft_status twin_follow(ft_batch *b) {
    for (int k = 1; k < b->depth; ++k)
        if (ft_status s = ft::ws_reserve_like(b->slot(k), b->ws)) return s;
    return FT_OK;
}

// Drop the twins (the caller has drained the batch's stream).
static void twins_free(ft_batch *b) {
    for (int k = 1; k < b->depth; ++k) ft::ws_free(b->slot(k));
    ft::ws_pipe_detach(b->ws);
    b->depth = 1;
}

namespace {

uint32_t bpp_of(ft_format f) { return f == FT_RGBA8P ? 4u : (f == FT_GRAYA8P ? 2u : 1u); }

bool valid_ops(const ft_path_op *ops, size_t n) {
    if (n && !ops) return false;
    for (size_t i = 0; i < n; ++i)
        if (ops[i].tag > FT_OP_PEN_WIDTH) return false;
    return true;
}

// s_width after the ops run: the last PenWidth wins (plotter.rs:151-153).
float final_pen_width(const ft_path_op *ops, size_t n, float w) {
    for (size_t i = n; i-- > 0;)
        if (ops[i].tag == FT_OP_PEN_WIDTH) return ops[i].v[0];
    return w;
}

ft_fill_desc plotter_desc(const ft_plotter *p, ft_fill_rule rule, size_t n, const uint8_t clr[4]) {
    ft_fill_desc d;
    memset(&d, 0, sizeof d);
    d.op_begin = 0;
    d.op_count = (uint32_t)n;
    d.rule = (uint32_t)rule;
    d.dst = (uint64_t)(uintptr_t)p->d_px;
    d.width = p->width;
    d.height = p->height;
    d.pitch = (uint32_t)p->pitch;
    d.fmt = (uint32_t)p->fmt;
    if (clr) memcpy(d.clr, clr, 4);
    memcpy(d.transform, p->transform, sizeof d.transform);
    d.tolerance = p->tolerance;
    d.pen_width = p->s_width;
    d.fresh = 0;
    return d;
}

ft_status plotter_fill_ops(ft_plotter *p, ft_fill_rule rule, const ft_path_op *ops, size_t n,
                           const uint8_t clr[4]) {
    if (n > 0x7FFFFFF0u) return FT_EINVAL;
    ft_fill_desc d = plotter_desc(p, rule, n, clr);
    // small figures: two launches (the full pipeline otherwise)
    bool handled = false;
    uint64_t offered = 0, stored = 0;
    ft_status s = ft::ws_small_fill(p->ws, ops, n, d, &handled, &offered, &stored);
    if (s) return s;
    if (handled) {
        p->last_offered = offered;
        p->last_stored = stored;
        return FT_OK;
    }
    s = ft::ws_set_batch(p->ws, ops, n, &d, 1);
    if (s) return s;
    s = ft::ws_run(p->ws, ft::MODE_FILL, true);
    if (s) return s;
    p->last_offered = p->ws.h_pstate[0].offered;
    p->last_stored = p->ws.h_pstate[0].stored;
    return FT_OK;
}

}  // namespace

extern "C" {

const char *ft_version(void) { return "footile-hip 0.1.0 (gfx950)"; }

ft_status ft_plotter_new(uint32_t width, uint32_t height, ft_format fmt, ft_plotter **out) {
    return ft_plotter_new_with_pixels(width, height, fmt, nullptr, out);
}

ft_status ft_plotter_new_with_pixels(uint32_t width, uint32_t height, ft_format fmt, const void *pixels,
                                     ft_plotter **out) {
    if (!out) return FT_EINVAL;
    *out = nullptr;
    if ((unsigned)fmt > FT_GRAYA8P) return FT_EINVAL;
    if (width > FT_MAX_WIDTH) return FT_ENOTSUP;
    ft_plotter *p = new (std::nothrow) ft_plotter();
    if (!p) return FT_ENOMEM;
    p->width = width;
    p->height = height;
    p->fmt = fmt;
    p->bpp = bpp_of(fmt);
    p->pitch = (size_t)width * p->bpp;
    ft_status s = ft::ws_init(p->ws);
    if (s) {
        ft::ws_free(p->ws);
        delete p;
        return s;
    }
    const size_t bytes = p->pitch * height;
    if (hipMalloc(reinterpret_cast<void **>(&p->d_px), bytes ? bytes : 1) != hipSuccess) {
        ft::ws_free(p->ws);
        delete p;
        return FT_ENOMEM;
    }
    hipError_t e = pixels ? hipMemcpy(p->d_px, pixels, bytes, hipMemcpyHostToDevice)
                          : hipMemset(p->d_px, 0, bytes ? bytes : 1);
    if (e != hipSuccess) {
        ft_plotter_free(p);
        return FT_EDEVICE;
    }
    *out = p;
    return FT_OK;
}

void ft_plotter_free(ft_plotter *p) {
    if (!p) return;
    if (p->d_px) (void)hipFree(p->d_px);
    ft::ws_free(p->ws);
    delete p;
}

ft_status ft_plotter_size(const ft_plotter *p, uint32_t *width, uint32_t *height) {
    if (!p) return FT_EINVAL;
    if (width) *width = p->width;
    if (height) *height = p->height;
    return FT_OK;
}

ft_status ft_plotter_set_tolerance(ft_plotter *p, float t) {
    if (!p) return FT_EINVAL;
    p->tolerance = t > 0.01f ? t : 0.01f;  // t.max(0.01), plotter.rs:134 (NaN -> 0.01)
    return FT_OK;
}

ft_status ft_plotter_set_transform(ft_plotter *p, const float e[6]) {
    if (!p || !e) return FT_EINVAL;
    memcpy(p->transform, e, sizeof p->transform);
    return FT_OK;
}

ft_status ft_plotter_set_join(ft_plotter *p, ft_join kind, float miter_limit) {
    if (!p || (unsigned)kind > FT_JOIN_ROUND) return FT_EINVAL;
    p->join = kind;
    p->miter = miter_limit;
    return FT_OK;
}

ft_status ft_plotter_fill(ft_plotter *p, ft_fill_rule rule, const ft_path_op *ops, size_t n, const uint8_t clr[4]) {
    if (!p || (unsigned)rule > FT_EVEN_ODD || !valid_ops(ops, n)) return FT_EINVAL;
    ft_status s = plotter_fill_ops(p, rule, ops, n, clr);
    if (s == FT_OK) p->s_width = final_pen_width(ops, n, p->s_width);
    return s;
}

// Plotter::stroke (plotter.rs:356-365): the ops are flattened with widths,
// the stroker stages build the outline, and the outline is filled NonZero --
// all in one device run (two launches for a small figure).
ft_status ft_plotter_stroke(ft_plotter *p, const ft_path_op *ops, size_t n, const uint8_t clr[4]) {
    if (!p || !valid_ops(ops, n)) return FT_EINVAL;
    if (n > 0x7FFFFFF0u) return FT_EINVAL;
    ft_fill_desc d = plotter_desc(p, FT_NON_ZERO, n, clr);
    p->ws.join = (uint32_t)p->join;
    p->ws.miter = p->miter;
    p->ws.stroke_tol_sq = p->tolerance * p->tolerance;
    bool handled = false;
    uint64_t offered = 0, stored = 0;
    ft_status s = ft::ws_small_fill(p->ws, ops, n, d, &handled, &offered, &stored, ft::MODE_STROKE);
    if (s) return s;
    if (handled) {
        p->last_offered = offered;
        p->last_stored = stored;
    } else {
        s = ft::ws_set_batch(p->ws, ops, n, &d, 1, ft::PATH_WHOLE);
        if (s) return s;
        s = ft::ws_run(p->ws, ft::MODE_STROKE, true);
        if (s) return s;
        p->last_offered = p->ws.h_pstate[0].offered;
        p->last_stored = p->ws.h_pstate[0].stored;
    }
    p->s_width = final_pen_width(ops, n, p->s_width);
    return FT_OK;
}

ft_status ft_plotter_read(const ft_plotter *p, void *dst, size_t bytes) {
    if (!p || (!dst && bytes)) return FT_EINVAL;
    const size_t n = p->pitch * p->height;
    if (bytes < n) return FT_EINVAL;
    if (n && hipMemcpy(dst, p->d_px, n, hipMemcpyDeviceToHost) != hipSuccess) return FT_EDEVICE;
    return FT_OK;
}

ft_status ft_plotter_write(ft_plotter *p, const void *src, size_t bytes) {
    if (!p || (!src && bytes)) return FT_EINVAL;
    const size_t n = p->pitch * p->height;
    if (bytes < n) return FT_EINVAL;
    if (n && hipMemcpy(p->d_px, src, n, hipMemcpyHostToDevice) != hipSuccess) return FT_EDEVICE;
    return FT_OK;
}

ft_status ft_plotter_device_ptr(ft_plotter *p, void **dptr, size_t *pitch) {
    if (!p) return FT_EINVAL;
    if (dptr) *dptr = p->d_px;
    if (pitch) *pitch = p->pitch;
    return FT_OK;
}

ft_status ft_plotter_pen_width(const ft_plotter *p, float *width) {
    if (!p || !width) return FT_EINVAL;
    *width = p->s_width;
    return FT_OK;
}

ft_status ft_plotter_last_counts(const ft_plotter *p, uint64_t *offered, uint64_t *stored) {
    if (!p) return FT_EINVAL;
    if (offered) *offered = p->last_offered;
    if (stored) *stored = p->last_stored;
    return FT_OK;
}

ft_status ft_fig_fill(ft_fill_rule rule, const float *xy, const uint8_t *close_after, size_t n, ft_format fmt,
                      uint32_t width, uint32_t height, void *pixels, const uint8_t clr[4]) {
    if ((unsigned)rule > FT_EVEN_ODD || (unsigned)fmt > FT_GRAYA8P) return FT_EINVAL;
    if ((n && (!xy || !close_after)) || (!pixels && width && height)) return FT_EINVAL;
    if (n && !close_after[n - 1]) return FT_EINVAL;  // Fig::fill asserts the sub is done (fig.rs:492)
    if (width > FT_MAX_WIDTH) return FT_ENOTSUP;
    std::vector<ft_path_op> ops;
    ops.reserve(n + n / 2 + 1);
    for (size_t i = 0; i < n; ++i) {
        ft_path_op op{};
        op.tag = FT_OP_LINE;  // Fig::add_point
        op.v[0] = xy[2 * i];
        op.v[1] = xy[2 * i + 1];
        ops.push_back(op);
        if (close_after[i]) {
            ft_path_op c{};
            c.tag = FT_OP_CLOSE;  // Fig::close
            ops.push_back(c);
        }
    }
    Workspace ws;
    ft_status s = ft::ws_init(ws);
    uint8_t *d_px = nullptr;
    const uint32_t bpp = bpp_of(fmt);
    const size_t bytes = (size_t)width * bpp * height;
    if (!s && hipMalloc(reinterpret_cast<void **>(&d_px), bytes ? bytes : 1) != hipSuccess) s = FT_ENOMEM;
    if (!s && bytes && hipMemcpy(d_px, pixels, bytes, hipMemcpyHostToDevice) != hipSuccess) s = FT_EDEVICE;
    if (!s) {
        ft_fill_desc d;
        memset(&d, 0, sizeof d);
        d.op_count = (uint32_t)ops.size();
        d.rule = (uint32_t)rule;
        d.dst = (uint64_t)(uintptr_t)d_px;
        d.width = width;
        d.height = height;
        d.pitch = width * bpp;
        d.fmt = (uint32_t)fmt;
        if (clr) memcpy(d.clr, clr, 4);
        d.transform[0] = 1.0f;
        d.transform[4] = 1.0f;
        d.tolerance = 0.3f;
        d.pen_width = 1.0f;
        // raw points: Fig::add_point receives the caller's f32 values
        s = ft::ws_set_batch(ws, ops.data(), ops.size(), &d, 1, ft::PATH_RAW);
        if (!s) s = ft::ws_run(ws, ft::MODE_FILL, false);
        if (!s && bytes && hipMemcpy(pixels, d_px, bytes, hipMemcpyDeviceToHost) != hipSuccess) s = FT_EDEVICE;
    }
    if (d_px) (void)hipFree(d_px);
    ft::ws_free(ws);
    return s;
}

ft_status ft_batch_new(const ft_path_op *ops, size_t n_ops, const ft_fill_desc *paths, size_t n_paths,
                       ft_batch **out) {
    if (!out || (n_paths && !paths) || !valid_ops(ops, n_ops)) return FT_EINVAL;
    *out = nullptr;
    ft_batch *b = new (std::nothrow) ft_batch();
    if (!b) return FT_ENOMEM;
    ft_status s = ft::ws_init(b->ws);
    if (!s) s = ft::ws_set_batch(b->ws, ops, n_ops, paths, n_paths);
    if (s) {
        ft::ws_free(b->ws);
        delete b;
        return s;
    }
    b->n_paths = n_paths;
    b->ops.assign(ops, ops + n_ops);
    b->descs.assign(paths, paths + n_paths);
    *out = b;
    return FT_OK;
}

ft_status ft_batch_set_tuning(ft_batch *b, const ft_tuning *t) {
    if (!b || !t) return FT_EINVAL;
    if (t->band_rows_min < 1 || t->band_rows_min > 256 || t->lds_cells < 1024 || t->lds_cells > 32768 ||
        t->lds_cells % ft::CELL_ALIGN)
        return FT_EINVAL;
    ft::Tuning tu;
    tu.band_rows_min = t->band_rows_min;
    tu.lds_cells = t->lds_cells;
    tu.dst_prefetch = t->dst_prefetch ? 1u : 0u;
    tu.tile_groups = t->tile_groups;
    const ft::Tuning old = b->ws.tuning;
    b->ws.tuning = tu;
    ft_status s = ft::ws_set_batch(b->ws, b->ops.data(), b->ops.size(), b->descs.data(), b->descs.size());
    if (s) {
        b->ws.tuning = old;
        (void)ft::ws_set_batch(b->ws, b->ops.data(), b->ops.size(), b->descs.data(), b->descs.size());
        return s;
    }
    for (int k = 1; k < b->depth && !s; ++k) {
        b->slot(k).tuning = tu;
        s = ft::ws_set_batch(b->slot(k), b->ops.data(), b->ops.size(), b->descs.data(), b->descs.size());
    }
    return s ? s : twin_follow(b);
}

ft_status ft_batch_get_tuning(const ft_batch *b, ft_tuning *t) {
    if (!b || !t) return FT_EINVAL;
    t->band_rows_min = b->ws.tuning.band_rows_min;
    t->lds_cells = b->ws.tuning.lds_cells;
    t->dst_prefetch = b->ws.tuning.dst_prefetch;
    t->tile_groups = b->ws.tuning.tile_groups;
    return FT_OK;
}

ft_status ft_batch_set_graph(ft_batch *b, int on) {
    if (!b || (on && b->depth > 1)) return FT_EINVAL;
    b->ws.graph = on != 0;
    return FT_OK;
}

void ft_batch_free(ft_batch *b) {
    if (!b) return;
    if (b->depth > 1) {
        (void)hipStreamSynchronize(b->ws.stream);
        twins_free(b);  // (their raster stream is ws's)
    }
    ft::ws_free(b->ws);
    delete b;
}

ft_status ft_batch_set_pipeline(ft_batch *b, int depth) {
    if (!b || depth < 0 || depth > ft_batch::PIPE_MAX || (depth > 1 && b->ws.graph)) return FT_EINVAL;
    if (depth < 1) depth = 1;
    if (depth == b->depth) return FT_OK;
    // every launch so far ends on ws.stream (its front end precedes its raster)
    if (hipStreamSynchronize(b->ws.stream) != hipSuccess) return FT_EDEVICE;
    for (bool &l : b->live) l = false;
    b->turn = 0;
    twins_free(b);
    if (depth == 1) return FT_OK;
    ft_status s = ft::ws_pipe_attach(b->ws, b->ws.stream);
    for (int k = 1; k < depth && !s; ++k) {
        ft::Workspace &w = b->slot(k);
        b->depth = k + 1;  // (twins_free frees this one on failure)
        s = ft::ws_init(w);
        if (!s) s = ft::ws_pipe_attach(w, b->ws.stream);
        if (!s) {
            w.tuning = b->ws.tuning;
            w.timing = b->ws.timing;
            s = ft::ws_set_batch(w, b->ops.data(), b->ops.size(), b->descs.data(), b->descs.size());
        }
        if (!s) s = ft::ws_reserve_like(w, b->ws);
    }
    if (s) {
        (void)hipStreamSynchronize(b->ws.stream);
        twins_free(b);
    }
    return s;
}

ft_status ft_batch_launch(ft_batch *b) {
    if (!b) return FT_EINVAL;
    const int k = b->turn;
    b->turn = (k + 1) % b->depth;
    const ft_status s = ft::ws_launch(b->slot(k), ft::MODE_FILL);
    if (!s) b->live[k] = true;
    return s;
}

ft_status ft_batch_sync(ft_batch *b) {
    if (!b) return FT_EINVAL;
    // (a pipelined batch: every workspace's raster and status copy ends on
    // ws.stream, so the first finish waits for all of them)
    ft_status s = FT_OK;
    uint32_t ov = 0;
    for (int k = 0; k < b->depth; ++k) {
        if (!(k == 0 || b->live[k])) continue;
        ft::Workspace &w = b->slot(k);
        const ft_status sk = ft::ws_finish(w, true);
        if (sk == (ft_status)0x100)
            ov |= w.h_status->overflow;
        else if (sk && !s)
            s = sk;
    }
    for (bool &l : b->live) l = false;
    b->turn = 0;
    if (s) return s;
    // a run whose kept flags needed the ingest pass runs again with it
    if (ov && !(ov & ~ft::OV_LONG_WALK)) {
        s = ft::ws_run(b->ws, ft::MODE_FILL, true);
        return s ? s : twin_follow(b);
    }
    return ov ? FT_ENOMEM : FT_OK;
}

ft_status ft_batch_run(ft_batch *b) {
    if (!b) return FT_EINVAL;
    const ft_status s = ft::ws_run(b->ws, ft::MODE_FILL, true);
    return s ? s : twin_follow(b);
}

ft_status ft_batch_counts(ft_batch *b, uint64_t *offered, uint64_t *stored, uint64_t *edges,
                          int32_t *top_rows) {
    if (!b) return FT_EINVAL;
    if (edges) {
        const ft_status st = ft::ws_count_edges(b->ws);
        if (st) return st;
    }
    for (size_t i = 0; i < b->n_paths; ++i) {
        const ft::PathState &s = b->ws.h_pstate[i];
        if (offered) offered[i] = s.offered;
        if (stored) stored[i] = s.stored;
        if (edges) edges[i] = s.n_edges;
        if (top_rows)
            top_rows[i] = s.topkey == ~0ull ? INT32_MAX
                                             : ((int32_t)((uint32_t)(s.topkey >> 32) ^ 0x80000000u)) >> 16;
    }
    return FT_OK;
}

ft_status ft_batch_raster_stats(ft_batch *b, uint64_t *pairs, uint64_t *cell_adds) {
    if (!b) return FT_EINVAL;
    return ft::ws_raster_stats(b->ws, pairs, cell_adds);
}

ft_status ft_batch_stream(ft_batch *b, void **stream) {
    if (!b || !stream) return FT_EINVAL;
    *stream = b->ws.stream;
    return FT_OK;
}

ft_status ft_batch_set_timing(ft_batch *b, int on) {
    if (!b) return FT_EINVAL;
    for (int k = 0; k < ft_batch::PIPE_MAX; ++k) b->slot(k).timing = on != 0;
    return FT_OK;
}

ft_status ft_batch_timing(ft_batch *b, double *total_ms, double *raster_ms, uint64_t *runs) {
    if (!b) return FT_EINVAL;
    double total = 0.0, raster = 0.0;
    uint64_t n = 0;
    for (int k = 0; k < ft_batch::PIPE_MAX; ++k) {
        ft::Workspace &w = b->slot(k);
        total += w.t_total_ms;
        raster += w.t_raster_ms;
        n += w.n_timed;
        w.t_total_ms = w.t_raster_ms = 0.0;
        w.n_timed = 0;
    }
    if (total_ms) *total_ms = total;
    if (raster_ms) *raster_ms = raster;
    if (runs) *runs = n;
    return FT_OK;
}

ft_status ft_device_alloc(size_t bytes, void **dptr) {
    if (!dptr) return FT_EINVAL;
    return hipMalloc(dptr, bytes ? bytes : 1) == hipSuccess ? FT_OK : FT_ENOMEM;
}

ft_status ft_device_free(void *dptr) { return hipFree(dptr) == hipSuccess ? FT_OK : FT_EDEVICE; }

ft_status ft_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? FT_OK : FT_EDEVICE;
}

ft_status ft_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? FT_OK : FT_EDEVICE;
}

ft_status ft_memset_d(void *dst, int value, size_t bytes) {
    return hipMemset(dst, value, bytes) == hipSuccess ? FT_OK : FT_EDEVICE;
}

}  // extern "C"
