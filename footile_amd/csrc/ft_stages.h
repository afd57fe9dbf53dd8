// ft_stages.h -- per-item logic of every pipeline stage, host- and
// device-compilable.
//
// The gfx950 kernels (ft_pipeline.hip) are thin loops around these functions;
// tests/cpu_mirror runs the very same functions sequentially on the host
// under AddressSanitizer and compares the result with the CPU oracle.
// Atomic read-modify-writes go through a policy type `A` (device atomics on
// the GPU, plain operations in the sequential host build).
#pragma once

#include <stdint.h>

#include "ft_device.h"
#include "ft_types.h"

namespace ft {

// ---------------------------------------------------------------------------
// Flattening (src/plotter.rs:163-332, src/geom.rs:9-35)
// ---------------------------------------------------------------------------
struct WP {
    float x, y, w;
};

// WidePt::midpoint: position midpoint (a+b)/2 per axis, width (w0+w1)/2.
FT_HD inline WP wmid(WP a, WP b) {
    WP r;
    r.x = (a.x + b.x) / 2.0f;
    r.y = (a.y + b.y) / 2.0f;
    r.w = (a.w + b.w) / 2.0f;
    return r;
}

// Plotter::is_within_tolerance2: distance_sq <= tol_sq (plotter.rs:273-276)
FT_HD inline bool within_tol(WP a, WP b, float tol_sq) {
    float dx = a.x - b.x;
    float dy = a.y - b.y;
    return dx * dx + dy * dy <= tol_sq;
}

// Plotter::transform_point (plotter.rs:169-172): width is not transformed.
FT_HD inline WP xform(const PathDev &P, float x, float y, float w) {
    WP r;
    if (P.flags & PATH_RAW) {
        r.x = x;
        r.y = y;
    } else {
        r.x = P.tr[0] * x + P.tr[1] * y + P.tr[2];
        r.y = P.tr[3] * x + P.tr[4] * y + P.tr[5];
    }
    r.w = w;
    return r;
}

struct Quad {
    WP a, b, c;
};
struct Cubic {
    WP a, b, c, d;
};

// One subdivision step of quad_to_tran (plotter.rs:255-258).
FT_HD inline bool split(const Quad &q, float tol_sq, Quad &L, Quad &R) {
    WP ab = wmid(q.a, q.b);
    WP bc = wmid(q.b, q.c);
    WP ab_bc = wmid(ab, bc);
    WP ac = wmid(q.a, q.c);
    L.a = q.a; L.b = ab; L.c = ab_bc;
    R.a = ab_bc; R.b = bc; R.c = q.c;
    return within_tol(ab_bc, ac, tol_sq);
}

// One subdivision step of cubic_to_tran (plotter.rs:319-326).
FT_HD inline bool split(const Cubic &q, float tol_sq, Cubic &L, Cubic &R) {
    WP ab = wmid(q.a, q.b);
    WP bc = wmid(q.b, q.c);
    WP cd = wmid(q.c, q.d);
    WP ab_bc = wmid(ab, bc);
    WP bc_cd = wmid(bc, cd);
    WP pe = wmid(ab_bc, bc_cd);
    WP ad = wmid(q.a, q.d);
    L.a = q.a; L.b = ab; L.c = ab_bc; L.d = pe;
    R.a = pe; R.b = bc_cd; R.c = cd; R.d = q.d;
    return within_tol(pe, ad, tol_sq);
}

FT_HD inline WP end_of(const Quad &q) { return q.c; }
FT_HD inline WP end_of(const Cubic &q) { return q.d; }

// Depth-first walk of the subdivision tree in the recursion's order (left
// subtree, then right).  `bits` holds the left/right choices from the root;
// moving to a right sibling replays those choices from the root, so every
// node is computed by exactly the float operations the recursion performs.
// Returns 0, or FT_ERANGE when a node `max_depth` levels below `root` is
// still out of tolerance or `leaves` exceeds FT_FLATTEN_MAX_LEAVES.
template <class Curve, class Emit>
FT_HD uint32_t flatten_walk(const Curve &root, float tol_sq, uint32_t max_depth, uint32_t &leaves, Emit &&emit) {
    Curve cur = root;
    uint32_t depth = 0, bits = 0;
    for (;;) {
        Curve L, R;
        if (split(cur, tol_sq, L, R)) {
            if (++leaves > FT_FLATTEN_MAX_LEAVES) return FT_ERANGE;
            emit(end_of(cur));
            while (depth > 0 && (bits & 1u)) {
                bits >>= 1;
                --depth;
            }
            if (depth == 0) return 0;
            bits |= 1u;
            cur = root;
            for (int lv = (int)depth - 1; lv >= 0; --lv) {
                Curve l2, r2;
                split(cur, tol_sq, l2, r2);
                cur = ((bits >> lv) & 1u) ? r2 : l2;
            }
        } else {
            if (depth >= max_depth) return FT_ERANGE;
            cur = L;
            bits <<= 1;
            ++depth;
        }
    }
}

// A curve is flattened by 2^sp lanes (OP_LANES in the pipeline, sp =
// FLAT_SPLIT).  Lane s takes the subtree sp levels below the root whose
// left/right choices are the bits of s, most significant first; a node that
// is already within tolerance above that depth is a leaf of the lowest lane
// beneath it, and the other lanes beneath it emit nothing.  Concatenated in
// lane order the lanes' points are the recursion's points in order
// (plotter.rs:250-262,314-330), whatever sp (< FT_FLATTEN_MAX_DEPTH) is.
constexpr uint32_t FLAT_SPLIT = 3;
constexpr uint32_t OP_LANES = 1u << FLAT_SPLIT;

template <class Curve, class Emit>
FT_HD uint32_t flatten_share(const Curve &root, float tol_sq, uint32_t s, uint32_t &leaves, Emit &&emit,
                             uint32_t sp = FLAT_SPLIT) {
    Curve cur = root;
    for (uint32_t lv = 0; lv < sp; ++lv) {
        Curve L, R;
        if (split(cur, tol_sq, L, R)) {
            if ((s & ((1u << (sp - lv)) - 1u)) == 0) {
                ++leaves;
                emit(end_of(cur));
            }
            return 0;
        }
        cur = ((s >> (sp - 1 - lv)) & 1u) ? R : L;
    }
    return flatten_walk(cur, tol_sq, FT_FLATTEN_MAX_DEPTH - sp, leaves, emit);
}

// Index of the op that last set the pen before op i (plotter.rs:128-130,
// 164-166): PenWidth ops do not move the pen.  Returns < begin at path start.
FT_HD inline int64_t pen_event_before(const ft_path_op *ops, uint32_t i, uint32_t begin) {
    int64_t j = (int64_t)i - 1;
    while (j >= (int64_t)begin && ops[j].tag == TAG_PEN_WIDTH) --j;
    return j;
}

// s_width in effect just before op k: the last PenWidth op of the path before
// k, else the width the fill started with (plotter.rs:53,151-153).
FT_HD inline float width_before(const ft_path_op *ops, const uint32_t *lastpw, int64_t k, const PathDev &P) {
    if (!lastpw || k < (int64_t)P.op_begin) return P.s_width0;
    uint32_t l = lastpw[k];
    if (l == 0 || l - 1 < P.op_begin || (int64_t)(l - 1) >= k) return P.s_width0;
    return ops[l - 1].v[0];
}

// Pen (untransformed WidePt) before op i.
FT_HD inline WP pen_before(const ft_path_op *ops, const uint32_t *lastpw, uint32_t i, const PathDev &P) {
    int64_t j = pen_event_before(ops, i, P.op_begin);
    WP pen;
    pen.x = 0.0f;
    pen.y = 0.0f;
    if (j < (int64_t)P.op_begin) {
        pen.w = P.s_width0;
        return pen;
    }
    pen.w = width_before(ops, lastpw, j, P);
    const ft_path_op &o = ops[j];
    switch (o.tag) {
    case TAG_MOVE:
    case TAG_LINE: pen.x = o.v[0]; pen.y = o.v[1]; break;
    case TAG_QUAD: pen.x = o.v[2]; pen.y = o.v[3]; break;
    case TAG_CUBIC: pen.x = o.v[4]; pen.y = o.v[5]; break;
    default: break;  // Close resets the pen to the origin
    }
    return pen;
}

FT_HD inline Quad quad_of(const ft_path_op &o, WP pen, float sw, const PathDev &P) {
    Quad q;  // plotter.rs:233-240
    q.a = xform(P, pen.x, pen.y, pen.w);
    q.b = xform(P, o.v[0], o.v[1], (pen.w + sw) / 2.0f);
    q.c = xform(P, o.v[2], o.v[3], sw);
    return q;
}

FT_HD inline Cubic cubic_of(const ft_path_op &o, WP pen, float sw, const PathDev &P) {
    // float_lerp(a, b, t) = b + (a - b) * t   (geom.rs:14-16, plotter.rs:294-295)
    const float t1 = 1.0f / 3.0f, t2 = 2.0f / 3.0f;
    float w0 = sw + (pen.w - sw) * t1;
    float w1 = sw + (pen.w - sw) * t2;
    Cubic q;
    q.a = xform(P, pen.x, pen.y, pen.w);
    q.b = xform(P, o.v[0], o.v[1], w0);
    q.c = xform(P, o.v[2], o.v[3], w1);
    q.d = xform(P, o.v[4], o.v[5], sw);
    return q;
}

// ---------------------------------------------------------------------------
// Stage 1: events per op (plotter.rs:175-224 + Plotter::fill's final close,
// plotter.rs:347).  Close: 1 close; Move: 1 close + 1 point; Line: 1 point;
// Quad / Cubic: one point per subdivision leaf; PenWidth: nothing.  An op
// is handled by OP_LANES lanes: lane 0 takes the non-curve events, each lane
// its share of a curve, and the last lane the path's final close, so the
// lanes' events in lane order are the op's events in order.
// ---------------------------------------------------------------------------
struct OpCount {
    uint32_t ne, nc, err, leaves;
    uint32_t flags;  // OPC_CURVE: a fill's curve lane; OPC_LAST_CLOSE: it emits the path's final close
};
enum : uint32_t { OPC_CURVE = 1u, OPC_LAST_CLOSE = 2u };

// Fill runs may keep each curve lane's first LEAF_CACHE points (as the
// Fixed points add_point stores) in `cache` (point k at cache[k * stride]),
// so that emitting the lane is a copy unless it has more leaves.
constexpr uint32_t LEAF_CACHE = 8;

// sp / last_lane: a curve's lanes split its tree sp levels down, and the
// op's last lane (which emits the path's final close) is last_lane; the
// pipeline gives every op OP_LANES lanes.
FT_HD inline OpCount count_op(const ft_path_op *ops, const uint32_t *op_path, const PathDev *paths, uint32_t i,
                              uint32_t s, uint32_t mode, const uint32_t *lastpw, I2 *cache = nullptr,
                              size_t stride = 0, uint32_t sp = FLAT_SPLIT, uint32_t last_lane = OP_LANES - 1,
                              uint32_t cap = LEAF_CACHE) {
    OpCount c = {0, 0, 0, 0, 0};
    const uint32_t p = op_path[i];
    if (p == NO_IDX) return c;
    const PathDev &P = paths[p];
    const ft_path_op o = ops[i];
    const bool lead = s == 0;
    switch (o.tag) {
    case TAG_CLOSE: c.ne = lead ? 1 : 0; c.nc = lead ? 1 : 0; break;
    case TAG_MOVE: c.ne = lead ? 2 : 0; c.nc = lead ? 1 : 0; break;
    case TAG_LINE: c.ne = lead ? 1 : 0; break;
    case TAG_QUAD:
    case TAG_CUBIC: {
        const WP pen = pen_before(ops, lastpw, i, P);
        const float sw = width_before(ops, lastpw, i, P);
        I2 *const keep = mode == MODE_FILL ? cache : nullptr;
        auto cnt = [&](WP q) {
            const uint32_t k = c.leaves - 1;  // flatten counts a leaf before handing it over
            if (keep && k < cap) {
                I2 v;
                v.x = fx_from_f32(q.x);
                v.y = fx_from_f32(q.y);
                keep[k * stride] = v;
            }
        };
        const uint32_t rc = o.tag == TAG_QUAD ? flatten_share(quad_of(o, pen, sw, P), P.tol_sq, s, c.leaves, cnt, sp)
                                              : flatten_share(cubic_of(o, pen, sw, P), P.tol_sq, s, c.leaves, cnt, sp);
        if (rc)
            c.err = rc;
        else
            c.ne = c.leaves;
        c.flags |= mode == MODE_FILL ? OPC_CURVE : 0u;
        break;
    }
    case TAG_PEN_WIDTH: break;
    default: c.err = lead ? FT_EINVAL : 0u; break;
    }
    if (mode == MODE_FILL && i + 1 == P.op_begin + P.op_count && s == last_lane) {
        ++c.ne;
        ++c.nc;
        c.flags |= OPC_LAST_CLOSE;
    }
    return c;
}

FT_HD inline unsigned long long pack_count(OpCount c) { return ((unsigned long long)c.ne << 32) | c.nc; }

// ---------------------------------------------------------------------------
// Stage 2: emit events.  Fill mode writes Fixed points (f32 -> Fixed as
// Fig::add_point does, fig.rs:435-436) and the run (sub-figure) id of every
// event; stroke mode writes (x, y, width, kind) for the host stroker.
// ---------------------------------------------------------------------------
struct EventSink {
    uint32_t mode;
    uint64_t e;
    uint32_t run, p;
    I2 *ev_xy;
    F4 *ev_f;
    uint32_t *ev_meta;
    int32_t *run_first;
    uint32_t *run_close;  // the run's close event
    uint32_t *run_path, *run_start, *run_end;
    uint64_t cap_events;  // writes at or past these limits are dropped and
    uint32_t cap_runs;    // flagged in `bad`
    uint32_t bad;

    FT_HD void point(WP q) {
        const bool ok = e < cap_events;
        if (ok & (mode == MODE_FILL)) {
            I2 v;
            v.x = fx_from_f32(q.x);
            v.y = fx_from_f32(q.y);
            ev_xy[e] = v;
            ev_meta[e] = run;
        }
        if (ok & (mode != MODE_FILL)) {
            F4 v;
            v.x = q.x;
            v.y = q.y;
            v.z = q.w;
            v.w = 0.0f;
            ev_f[e] = v;
        }
        bad |= ok ? 0u : 1u;
        ++e;
    }
    FT_HD void point_fx(I2 v) {  // a fill point already converted (count_op's cache)
        const bool ok = e < cap_events;
        if (ok) {
            ev_xy[e] = v;
            ev_meta[e] = run;
        }
        bad |= ok ? 0u : 1u;
        ++e;
    }
    FT_HD void close(bool joined) {
        const bool ok = (e < cap_events) & (run < cap_runs);
        if (ok & (mode == MODE_FILL)) {
            ev_meta[e] = run | CLOSE_BIT;
            run_first[run] = -1;
            run_close[run] = (uint32_t)e;
            run_path[run] = p;
            run_start[run] = NO_IDX;
            run_end[run] = 0;
        }
        if (ok & (mode != MODE_FILL)) {
            F4 v;
            v.x = 0.0f;
            v.y = 0.0f;
            v.z = joined ? 1.0f : 0.0f;
            v.w = joined ? 2.0f : 1.0f;
            ev_f[e] = v;
        }
        bad |= ok ? 0u : 1u;
        ++run;
        ++e;
    }
};

// Writes lane s's events of op i at offset `off` (events << 32 | closes
// before them).
// `cache` / `cached`: the lane's points kept by count_op (point k at
// cache[k * stride]) and their number (used in fill runs when it is at most
// the cache's capacity `cap`).
FT_HD inline uint32_t emit_op(const ft_path_op *ops, const uint32_t *op_path, const PathDev *paths, uint32_t i,
                              uint32_t s, uint32_t mode, const uint32_t *lastpw, unsigned long long off,
                              EventSink em, const I2 *cache = nullptr, uint32_t cached = 0, size_t stride = 0,
                              uint32_t sp = FLAT_SPLIT, uint32_t last_lane = OP_LANES - 1,
                              uint32_t cap = LEAF_CACHE) {
    const uint32_t p = op_path[i];
    if (p == NO_IDX) return 0;
    const PathDev &P = paths[p];
    const ft_path_op o = ops[i];
    const bool lead = s == 0;
    em.e = off >> 32;
    em.run = (uint32_t)off;
    em.p = p;
    switch (o.tag) {
    case TAG_CLOSE:  // plotter.rs:200-203
        if (lead) em.close(true);
        break;
    case TAG_MOVE:  // plotter.rs:208-214
        if (lead) {
            em.close(false);
            em.point(xform(P, o.v[0], o.v[1], width_before(ops, lastpw, i, P)));
        }
        break;
    case TAG_LINE:  // plotter.rs:219-224
        if (lead) em.point(xform(P, o.v[0], o.v[1], width_before(ops, lastpw, i, P)));
        break;
    case TAG_QUAD:
    case TAG_CUBIC: {
        if (cache && mode == MODE_FILL && cached <= cap) {
            for (uint32_t k = 0; k < cached; ++k) em.point_fx(cache[k * stride]);
            break;
        }
        const WP pen = pen_before(ops, lastpw, i, P);
        const float sw = width_before(ops, lastpw, i, P);
        uint32_t leaves = 0;
        auto put = [&](WP q) { em.point(q); };
        if (o.tag == TAG_QUAD)
            flatten_share(quad_of(o, pen, sw, P), P.tol_sq, s, leaves, put, sp);
        else
            flatten_share(cubic_of(o, pen, sw, P), P.tol_sq, s, leaves, put, sp);
        break;
    }
    default: break;
    }
    if (mode == MODE_FILL && i + 1 == P.op_begin + P.op_count && s == last_lane) em.close(true);
    return em.bad;
}

// ---------------------------------------------------------------------------
// Stages 3-5 are written branch-free: every value is computed for every lane
// (with in-range stand-in indices) and only the stores and atomics are
// predicated.  No value then takes different definitions along divergent
// paths, a shape the gfx950 compiler has been seen to miscompile (a register
// live for lanes inactive in a nested divergent block was reused inside that
// block).  Every index read from memory is also checked against the run's
// limits; a failed check skips the access and sets `bad`.
// ---------------------------------------------------------------------------
struct Limits {
    uint32_t n_ops, n_paths, n_bands;
    uint32_t n_events;    // events written by this run
    uint32_t n_runs;      // runs (= closes) of this run
    uint32_t cap_points;  // capacity of the point / edge arrays
    uint32_t n_pairs;     // (edge, band) pairs of this run
};

FT_HD inline bool i2_eq(I2 a, I2 b) { return (a.x == b.x) & (a.y == b.y); }

// Stage 3: Fig::add_point / Fig::close (fig.rs:373-383,428-451).  Within a
// run (events between closes) a point is pushed if it is the run's first or
// differs in Fixed from the previous point event; closing pops the last
// pushed point when it equals the run's first point.
struct Pushed {
    uint32_t run;
    bool is_pt, first, pushed, ok;
};

FT_HD inline Pushed pushed_event(uint32_t e, const Limits &L, const uint32_t *meta, const I2 *xy) {
    const uint32_t ep = e > 0 ? e - 1 : 0;
    const uint32_t m = meta[e], mp = meta[ep];
    const I2 a = xy[e], b = xy[ep];
    Pushed r;
    r.run = m & ~CLOSE_BIT;
    r.ok = r.run < L.n_runs;
    r.is_pt = r.ok & ((m & CLOSE_BIT) == 0);
    r.first = (e == 0) | ((mp & CLOSE_BIT) != 0);
    r.pushed = r.first | !i2_eq(a, b);
    return r;
}

// Records the run's first point and `last[e]` = e + 1 for a pushed point
// event, else 0.  The inclusive max-scan of `last` at a run's close event is
// then 1 + the run's last pushed point (pushed points rise with e, and the
// run's first point is always pushed).
// ingest_mark returns last[e] and, in first_run, the run whose first point is
// e (NO_IDX: none); ingest_event stores both.
FT_HD inline uint32_t ingest_mark(uint32_t e, const Limits &L, const uint32_t *meta, const I2 *xy,
                                  const uint32_t *run_path, uint32_t &bad, uint32_t &first_run) {
    const Pushed q = pushed_event(e, L, meta, xy);
    const uint32_t p = run_path[q.ok ? q.run : 0];
    first_run = (q.is_pt & q.first) ? q.run : NO_IDX;
    bad |= (q.ok & (p < L.n_paths)) ? 0u : 1u;
    return (q.is_pt & q.pushed) ? e + 1 : 0u;
}

FT_HD inline void ingest_event(uint32_t e, const Limits &L, const uint32_t *meta, const I2 *xy, int32_t *run_first,
                               const uint32_t *run_path, uint32_t *last, uint32_t &bad) {
    uint32_t first_run;
    last[e] = ingest_mark(e, L, meta, xy, run_path, bad, first_run);
    if (first_run != NO_IDX) run_first[first_run] = (int32_t)e;
}

// 1 if point event e survives the closing pop.  `last_scan` is the inclusive
// max-scan of ingest_event's `last`.
FT_HD inline uint32_t kept_event(uint32_t e, const Limits &L, const uint32_t *meta, const I2 *xy,
                                 const int32_t *run_first, const uint32_t *run_close, const uint32_t *last_scan,
                                 uint32_t &bad) {
    const Pushed q = pushed_event(e, L, meta, xy);
    const uint32_t run = q.ok ? q.run : 0;
    const int32_t rf = run_first[run];
    const uint32_t c = run_close[run];
    const bool cok = c < L.n_events;
    const int32_t rl = (int32_t)last_scan[cok ? c : e] - 1;
    const bool rok = cok & (rf >= 0) & (rl >= rf) & ((uint32_t)rf < L.n_events) & ((uint32_t)rl < L.n_events);
    const I2 f = xy[rok ? (uint32_t)rf : e], l = xy[rok ? (uint32_t)rl : e];
    const bool pop = i2_eq(f, l) & ((int32_t)e == rl);
    bad |= ((q.is_pt & q.pushed & !rok) ? 1u : 0u) | (q.ok ? 0u : 1u);
    return (q.is_pt & q.pushed & rok & !pop) ? 1u : 0u;
}

// Points a run's closing walk visits at most in kept_event_local.
constexpr uint32_t KEPT_WALK = 32;

// kept_event without the ingest pass and its running maximum.  The runs are
// consecutive event ranges, each ending in its close, so run r's first event
// follows run r - 1's close; the run's first point is that event unless the
// run is only its close.  The run's last pushed point is found by walking
// back from the close over unpushed points (repeats of the point before),
// visiting at most `walk` points: a longer walk sets `long_walk` and keeps
// nothing (the caller then runs the ingest pass instead).  `first_of`
// receives e's run when e is the run's first point (else NO_IDX), for
// run_first.
FT_HD inline uint32_t kept_event_local(uint32_t e, const Limits &L, const uint32_t *meta, const I2 *xy,
                                       const uint32_t *run_close, uint32_t walk, uint32_t &bad, uint32_t &long_walk,
                                       uint32_t &first_of) {
    const Pushed q = pushed_event(e, L, meta, xy);
    const uint32_t run = q.ok ? q.run : 0;
    const uint32_t c = run_close[run];
    const uint32_t pc = run > 0 ? run_close[run - 1] : NO_IDX;
    const uint32_t fe = pc + 1;  // 0 for run 0
    const bool cok = (c < L.n_events) & ((run == 0) | (pc < c));
    const bool has = cok & (fe < c);  // the run has points
    // the last pushed point: the highest k in [fe, c) with k == fe or a point
    // unlike the one before
    uint32_t rl = NO_IDX, lw = 0;
    if (has) {
        uint32_t k = c - 1;
        for (uint32_t s = 0;; ++s) {
            const bool pushed = (k == fe) || !i2_eq(xy[k], xy[k - 1]);
            if (pushed) {
                rl = k;
                break;
            }
            if (s + 1 >= walk) {
                lw = 1;
                break;
            }
            --k;
        }
    }
    const bool rok = has & (rl != NO_IDX);
    const I2 f = xy[rok ? fe : e], l = xy[rok ? rl : e];
    const bool pop = i2_eq(f, l) & (e == rl);
    long_walk |= lw;
    first_of = (q.is_pt & has & (e == fe)) ? run : NO_IDX;
    bad |= ((q.is_pt & q.pushed & !rok & !lw) ? 1u : 0u) | (q.ok ? 0u : 1u) | ((q.ok & !cok) ? 1u : 0u);
    return (q.is_pt & q.pushed & rok & !pop) ? 1u : 0u;
}

// Local point index of a kept point with global index idx, or NO_IDX when the
// 65,535 cap drops it; `repop` tells whether the cap's re-pop removed 65534.
FT_HD inline uint32_t capped_local(uint32_t idx, uint32_t base, uint32_t repop) {
    const uint32_t local = idx - base;
    const bool keep = (local < FT_MAX_POINTS) & !((local == FT_MAX_POINTS - 1) & (repop != 0));
    return keep ? local : NO_IDX;
}

FT_HD inline uint32_t path_event_begin(const PathDev &P, const unsigned long long *op_off) {
    return (uint32_t)(op_off[P.op_begin] >> 32);
}

FT_HD inline uint32_t path_event_end(const PathDev &P, const unsigned long long *op_off, uint32_t n_ops,
                                     uint32_t n_events) {
    const uint32_t k = P.op_begin + P.op_count;
    return k < n_ops ? (uint32_t)(op_off[k] >> 32) : n_events;
}

FT_HD inline unsigned long long point_key(I2 q) {
    return ((unsigned long long)((uint32_t)q.y ^ 0x80000000u) << 32) |
           (unsigned long long)((uint32_t)q.x ^ 0x80000000u);
}

FT_HD inline int32_t top_row_of(unsigned long long key) {
    return ((int32_t)((uint32_t)(key >> 32) ^ 0x80000000u)) >> 16;
}

// Stage 4a, per path: the path's first point index, the 65,535-point cap
// (vid.rs:17, fig.rs:429-431) and the offered / stored counts.  Once the
// stored count reaches the cap every later point is ignored; the run holding
// the cap point then closes against its truncated last point (fig.rs:373-383),
// which pops local point 65534 when it equals the run's first point.  Every
// later run only adds a point and pops it again.  The stored points are the
// kept indices [pt_base, pt_base + stored).
struct CapSpan {
    uint32_t eb, ee;     // the path's events
    uint32_t base;       // kept index of its first point
    uint32_t offered, stored;
    bool over;           // stored reached the cap: cap_point / cap_repop decide
};

FT_HD inline CapSpan cap_span(uint32_t p, const Limits &L, const PathDev *paths, const unsigned long long *op_off,
                              const uint32_t *kept, const uint32_t *idx_of, uint32_t &bad) {
    const PathDev &P = paths[p];
    CapSpan c;
    c.eb = c.ee = c.base = c.offered = c.stored = 0;
    c.over = false;
    const bool whole = (P.flags & PATH_WHOLE) != 0;
    if (whole || (P.op_count && P.op_begin < L.n_ops)) {
        const uint32_t k = P.op_begin + P.op_count;
        const uint32_t eb = whole ? 0u : path_event_begin(P, op_off);
        const uint32_t ee = whole ? L.n_events : path_event_end(P, op_off, L.n_ops, L.n_events);
        const uint32_t cb = whole ? 0u : (uint32_t)op_off[P.op_begin];
        const uint32_t ce = whole ? L.n_runs : (k < L.n_ops ? (uint32_t)op_off[k] : L.n_runs);
        if (eb < ee && ee <= L.n_events && cb <= ce && ce - cb <= ee - eb) {
            c.eb = eb;
            c.ee = ee;
            c.offered = (ee - eb) - (ce - cb);  // point events: events minus closes
            c.base = idx_of[eb];
            c.stored = idx_of[ee - 1] + kept[ee - 1] - c.base;
            c.over = c.stored >= FT_MAX_POINTS;
        } else if (eb > ee || ee > L.n_events || cb > ce) {
            bad = 1;
        }
    }
    return c;
}

// The cap point is the kept event with local index 65534: the first event e of
// the path with cap_after(e) (a monotone predicate over the path's events).
FT_HD inline bool cap_after(uint32_t e, const CapSpan &c, const uint32_t *kept, const uint32_t *idx_of) {
    return idx_of[e] + kept[e] > c.base + FT_MAX_POINTS - 1;
}

// 1 if the close of the cap point's run pops it: the run's next pushed point
// is offered and dropped, and the close compares the run's first point with
// the cap point.
FT_HD inline uint32_t cap_repop(uint32_t ea, const CapSpan &c, const Limits &L, const uint32_t *meta, const I2 *xy,
                                const int32_t *run_first) {
    const uint32_t run = meta[ea] & ~CLOSE_BIT;
    uint32_t e2 = ea + 1;  // the next pushed point event of the path
    while (e2 < c.ee) {
        const Pushed q = pushed_event(e2, L, meta, xy);
        if (q.is_pt & q.pushed) break;
        ++e2;
    }
    if (e2 < c.ee && meta[e2] == run && run < L.n_runs) {
        const int32_t rf = run_first[run];
        if (rf >= 0 && (uint32_t)rf < L.n_events && i2_eq(xy[rf], xy[ea])) return 1;
    }
    return 0;
}

FT_HD inline void cap_store(uint32_t p, const CapSpan &c, uint32_t repop, PathState *ps) {
    ps[p].pt_base = c.base;
    ps[p].repop = repop;
    ps[p].offered = c.offered;
    ps[p].stored = c.over ? FT_MAX_POINTS - repop : c.stored;
}

// The whole stage for one path, sequentially (binary search for the cap point).
FT_HD inline void cap_path(uint32_t p, const Limits &L, const PathDev *paths, const unsigned long long *op_off,
                           const uint32_t *meta, const I2 *xy, const uint32_t *kept, const uint32_t *idx_of,
                           const int32_t *run_first, PathState *ps, uint32_t &bad) {
    const CapSpan c = cap_span(p, L, paths, op_off, kept, idx_of, bad);
    uint32_t repop = 0;
    if (c.over) {
        uint32_t lo = c.eb, hi = c.ee;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (cap_after(mid, c, kept, idx_of))
                hi = mid;
            else
                lo = mid + 1;
        }
        repop = cap_repop(lo, c, L, meta, xy, run_first);
    }
    cap_store(p, c, repop, ps);
}

// Stage 4b: compaction into the per-path point arrays (the Fig's `points`).
// A kept point beyond the cap gets pt_run = NO_IDX; a close event sets its
// run's end (one past its last stored point).
struct TopCand {
    uint32_t path;  // NO_IDX: not a stored point
    unsigned long long key;
};

// `top`, if given, receives the stored point's path and (y, x) key (path
// NO_IDX when the event stores no point), for the per-path top-left point.
FT_HD inline void compact_event(uint32_t e, const Limits &L, const uint32_t *meta, const I2 *xy,
                                const uint32_t *kept, const uint32_t *idx_of, const int32_t *run_first,
                                const uint32_t *run_path, uint32_t *run_start, uint32_t *run_end, I2 *pts,
                                uint32_t *pt_run, const PathState *ps, uint32_t &bad, TopCand *top = nullptr) {
    const uint32_t m = meta[e];
    const uint32_t run = m & ~CLOSE_BIT;
    const bool rok = run < L.n_runs;
    const uint32_t p = run_path[rok ? run : 0];
    const bool ok = rok & (p < L.n_paths);
    const PathState s = ps[ok ? p : 0];
    const uint32_t idx = idx_of[e];
    const bool kp = ok & (kept[e] != 0);
    const bool keep = kp & (capped_local(idx, s.pt_base, s.repop) != NO_IDX);
    const bool iok = idx < L.cap_points;
    const I2 q = xy[e];
    const int32_t rf = run_first[rok ? run : 0];
    const uint32_t cap_end = s.pt_base + s.stored;
    if (keep & iok) pts[idx] = q;
    if (kp & iok) pt_run[idx] = keep ? run : NO_IDX;
    if (keep & iok & ((int32_t)e == rf)) run_start[run] = idx;
    if (ok & ((m & CLOSE_BIT) != 0)) run_end[run] = idx < cap_end ? idx : cap_end;
    bad |= (ok ? 0u : 1u) | ((kp & !iok) ? 1u : 0u);
    if (top) {
        top->path = (keep & iok) ? p : NO_IDX;
        top->key = point_key(q);
    }
}

// Stage 4c: path and (y, x) key of stored point i, for the per-path top-left
// point (Fig::sort_points' first vertex, fig.rs:467-478).
FT_HD inline TopCand top_cand(uint32_t i, const Limits &L, const I2 *pts, const uint32_t *pt_run,
                              const uint32_t *run_path, uint32_t &bad) {
    const uint32_t run = pt_run[i];
    const bool rok = run < L.n_runs;
    const uint32_t p = run_path[rok ? run : 0];
    const bool ok = rok & (p < L.n_paths);
    bad |= ((run != NO_IDX) & !ok) ? 1u : 0u;
    TopCand c;
    c.path = ok ? p : NO_IDX;
    c.key = point_key(pts[i]);
    return c;
}

// Scan rows of an edge mapped to raster rows (fig.rs:496-500,536-553): scan
// row y_row = raster row + min(top_row, 0); rows run from max(top_row, 0) to
// the raster bottom.  Returns false when the edge touches no raster row.
FT_HD inline bool edge_rows(const EdgeRec &E, int32_t top_row, uint32_t H, int32_t &rr0, int32_t &rr1) {
    const int64_t first = top_row > 0 ? top_row : 0;
    const int64_t shift = top_row < 0 ? top_row : 0;
    int64_t a = (int64_t)fx_int(E.y_upper) - shift;
    int64_t b = (int64_t)fx_int(E.y_lower) - shift;
    a = a < first ? first : a;
    b = b > (int64_t)H - 1 ? (int64_t)H - 1 : b;
    rr0 = (int32_t)a;
    rr1 = (int32_t)b;
    return (first < (int64_t)H) & (a <= b);
}

// Stored point of event e, if any (branch-free).
struct PointRef {
    uint32_t path, idx, run;
    bool valid;
};

FT_HD inline PointRef point_of(uint32_t e, const Limits &L, const uint32_t *meta, const uint32_t *kept,
                               const uint32_t *idx_of, const uint32_t *run_path, const PathState *ps, uint32_t &bad) {
    const uint32_t run = meta[e] & ~CLOSE_BIT;
    const bool rok = run < L.n_runs;
    const uint32_t p = run_path[rok ? run : 0];
    const bool ok = rok & (p < L.n_paths);
    const PathState s = ps[ok ? p : 0];
    const uint32_t idx = idx_of[e];
    const bool iok = idx < L.cap_points;
    PointRef r;
    r.path = ok ? p : NO_IDX;
    r.idx = idx;
    r.run = run;
    r.valid = ok & iok & (kept[e] != 0) & (capped_local(idx, s.pt_base, s.repop) != NO_IDX);
    bad |= (ok ? 0u : 1u) | (((kept[e] != 0) & !iok) ? 1u : 0u);
    return r;
}

// Edge::new (fig.rs:179-210) for the segment q -> q1 (q1.y != q.y).
FT_HD inline EdgeRec make_edge(I2 q, I2 q1) {
    const bool fwd = q1.y > q.y;  // Forward: descends in vertex order
    const I2 up = fwd ? q : q1, lo = fwd ? q1 : q;
    const int32_t dx = wsub(lo.x, up.x), dy = wsub(lo.y, up.y);
    const int32_t dyd = dy != 0 ? dy : 1, dxd = dx != 0 ? dx : 1;  // no division by zero
    int32_t step = fx_abs(fx_div(dy, dxd));  // Edge::calculate_step
    step = step > FX_ONE ? FX_ONE : step;
    EdgeRec E;
    E.inv = fx_div(dx, dyd);
    E.step = dx != 0 ? step : 0;
    E.y_upper = up.y;
    E.y_lower = lo.y;
    E.x_bot0 = wadd(up.x, fx_mul(E.inv, wsub(fx_floor(wadd(up.y, FX_ONE)), up.y)));
    E.flags = 1u | (fwd ? 2u : 0u);
    return E;
}

// The global row bands [g0, g1] an edge touches.
FT_HD inline bool edge_bands(const EdgeRec &E, const PathDev &P, unsigned long long topkey, const Limits &L,
                             uint32_t &g0, uint32_t &g1) {
    int32_t r0, r1;
    const bool rows = edge_rows(E, top_row_of(topkey), P.height, r0, r1) & ((E.flags & 1u) != 0) &
                      (P.n_bands != 0);  // tile-mode paths have no bands
    const uint32_t br = P.band_rows ? P.band_rows : 1;
    g0 = P.band_base + (rows ? (uint32_t)r0 / br : 0u);
    g1 = P.band_base + (rows ? (uint32_t)r1 / br : 0u);
    return rows & (g1 < L.n_bands) & (g0 <= g1);
}

// Stage 5: one edge per stored point: the segment to its successor within the
// sub-figure (SubFig::next, fig.rs:143-163).  Edge::new runs once per edge.
// The caller records the edge's band range [g0, g1] for binning.
struct EdgeOut {
    uint32_t path;    // path of the stored point (NO_IDX: none)
    uint32_t g0, g1;  // band range, when `banded`
    bool valid;       // a stored point
    bool has;         // it starts a non-horizontal edge
    bool banded;      // that edge touches a raster row
};

// A path's top key may be kept as `reps` partial minima (the device
// compaction spreads its atomics over them); top_fold is their minimum.
constexpr uint32_t TOP_REPS = 8;
FT_HD inline unsigned long long top_fold(const unsigned long long *rep, uint32_t reps) {
    unsigned long long m = rep[0];
    for (uint32_t k = 1; k < reps; ++k) m = rep[k] < m ? rep[k] : m;
    return m;
}

// edge_point in two parts: edge_eval only loads (so that a thread's several
// points can have their loads in flight together), edge_commit stores.
// With `top_rep` the path's top key is folded from its partial minima (and
// stored into PathState::topkey by the points at it), else read from `ps`.
struct EdgeEval {
    EdgeOut out;
    EdgeRec E;
    unsigned long long topkey;  // the path's top key
    uint32_t top_local;         // local index of the point when it is at the path's top key
    bool top;
    bool fold;                  // topkey was folded from partial minima
};

FT_HD inline EdgeEval edge_eval(uint32_t i, const Limits &L, const uint32_t *pt_run, const uint32_t *run_path,
                                const uint32_t *run_start, const uint32_t *run_end, const PathDev *paths,
                                const I2 *pts, const PathState *ps, uint32_t &bad,
                                const unsigned long long *top_rep = nullptr, uint32_t top_reps = 1) {
    // point i is stored iff pt_run[i] names a run (compact_event writes NO_IDX for capped points)
    const uint32_t run0 = pt_run[i];
    const bool rok = run0 < L.n_runs;
    const uint32_t p0 = run_path[rok ? run0 : 0];
    const bool valid = rok & (p0 < L.n_paths);
    bad |= ((run0 != NO_IDX) & !valid) ? 1u : 0u;
    const uint32_t p = valid ? p0 : 0, run = valid ? run0 : 0;
    const PathState s = ps[p];
    const uint32_t rs = run_start[run], re = run_end[run];
    const uint32_t nx = (i + 1 < re) ? i + 1 : rs;
    const bool nok = nx < L.cap_points;
    const I2 q = pts[i], q1 = pts[nok ? nx : i];
    const bool has = valid & nok & (nx != i) & (q1.y != q.y);
    bad |= (valid & !nok) ? 1u : 0u;
    EdgeEval r;
    r.E = make_edge(q, q1);
    if (!has) r.E.flags = 0;
    r.fold = top_rep != nullptr;
    r.topkey = r.fold ? top_fold(top_rep + (size_t)p * top_reps, top_reps) : s.topkey;
    r.top = valid & (point_key(q) == r.topkey);
    r.top_local = i - s.pt_base;
    r.out.banded = edge_bands(r.E, paths[p], r.topkey, L, r.out.g0, r.out.g1) & has;
    r.out.path = valid ? p : NO_IDX;
    r.out.valid = valid;
    r.out.has = has;
    return r;
}

template <class A>
FT_HD inline void edge_commit(uint32_t i, const EdgeEval &r, EdgeRec *edges, PathState *ps) {
    if (r.top) A::min_u32(&ps[r.out.path].topidx, r.top_local);
    if (r.top & r.fold) ps[r.out.path].topkey = r.topkey;  // the same value from every point at it
    if (r.out.valid) edges[i] = r.E;
}

template <class A>
FT_HD inline EdgeOut edge_point(uint32_t i, const Limits &L, const uint32_t *pt_run, const uint32_t *run_path,
                                const uint32_t *run_start, const uint32_t *run_end, const PathDev *paths,
                                const I2 *pts, EdgeRec *edges, PathState *ps, uint32_t &bad,
                                const unsigned long long *top_rep = nullptr, uint32_t top_reps = 1) {
    const EdgeEval r =
        edge_eval(i, L, pt_run, run_path, run_start, run_end, paths, pts, ps, bad, top_rep, top_reps);
    edge_commit<A>(i, r, edges, ps);
    return r.out;
}

// Stage 5b: the band range k_edges recorded for point i, for scattering the
// edge into the band lists (band_list[band_off[g] + slot] = i, g0 <= g <= g1).
struct BinItem {
    uint32_t g0, g1;  // band range
    bool banded;
};

FT_HD inline BinItem bin_item(uint32_t i, const U2 *pt_band) {
    const U2 r = pt_band[i];
    BinItem b;
    b.g0 = r.x;
    b.g1 = r.y;
    b.banded = r.y >= r.x;
    return b;
}

// The record bin_item reads: [g0, g1], or an empty range.
FT_HD inline U2 band_record(const EdgeOut &o) {
    U2 r;
    r.x = o.banded ? o.g0 : 1u;
    r.y = o.banded ? o.g1 : 0u;
    return r;
}

// ---------------------------------------------------------------------------
// Stage 6: raster.
// ---------------------------------------------------------------------------

// Winding direction at the top-left vertex: Fig::get_dir (fig.rs:402-411).
FT_HD inline bool path_forward(const PathState &s, const Limits &L, const I2 *pts, const uint32_t *pt_run,
                               const uint32_t *run_start, const uint32_t *run_end, uint32_t &bad) {
    const uint32_t i = s.pt_base + s.topidx;
    const bool iok = i < L.cap_points;
    const uint32_t ii = iok ? i : 0;
    const uint32_t rn = pt_run[ii];
    const bool rok = rn < L.n_runs;
    const uint32_t rs = run_start[rok ? rn : 0], re = run_end[rok ? rn : 0];
    const uint32_t nf = (ii + 1 < re) ? ii + 1 : rs;
    const uint32_t nr = (ii > rs) ? ii - 1 : re - 1;
    const bool ok = iok & rok & (nf < L.cap_points) & (nr < L.cap_points);
    bad |= ok ? 0u : 1u;
    const I2 q = pts[ii], qf = pts[ok ? nf : ii], qr = pts[ok ? nr : ii];
    return widdershins(wsub(qr.x, q.x), wsub(qr.y, q.y), wsub(qf.x, q.x), wsub(qf.y, q.y));
}

// Signed-area cell deltas of one edge on one scan row: the per-row form of
// Edge::{starting_cov, continuing_cov, calculate_x_limits_*, scan_area,
// first_cov, step_cov} (fig.rs:238-321).  Cells left of x = 0 land in cell 0;
// cells at x >= W are dropped (fig.rs:292-295).
//
// The reference walks the row pixel by pixel: at step k the covered
// fraction is x_pix(k) = min(pixel_cov(min(xc0 + k * st, ONE)), cov), and
// pixel x = x_s + j receives x_pix(k_s + j) - x_pix(k_s + j - 1), up to and
// including the first pixel where x_pix reaches cov, or the raster's right
// edge.  row_setup describes the row; row_walk produces its cells in order.
struct RowCells {
    int64_t xc0, st;  // x_cov at step k: min(xc0 + k * st, ONE)
    int32_t cov;      // the row's coverage, 1..256
    int32_t x_s;      // first cell
    int32_t k_s;      // step of the first cell
    int32_t n;        // cells before the right edge (0: none); the walk ends earlier once x_pix reaches cov
};

FT_HD inline RowCells row_setup(const EdgeRec &E, int32_t sr, int32_t W) {
    RowCells c;
    c.xc0 = 0;
    c.st = FX_ONE;
    c.x_s = 0;
    c.k_s = 0;
    c.n = 0;
    const int32_t r0 = fx_int(E.y_upper), r1 = fx_int(E.y_lower);
    const bool starting = sr == r0, ending = sr == r1;
    int32_t cov = ending ? pixel_cov(fx_fract(E.y_lower)) : 256;
    if (starting) cov -= pixel_cov(fx_fract(E.y_upper));
    c.cov = cov;
    if (cov <= 0) return c;
    const int32_t x_bot = (int32_t)((uint32_t)E.x_bot0 + (uint32_t)(sr - r0) * (uint32_t)E.inv);
    const int32_t x0 = starting ? wsub(x_bot, fx_mul(E.inv, wsub(FX_ONE, fx_fract(E.y_upper)))) : wsub(x_bot, E.inv);
    const int32_t x1 = ending ? wsub(x_bot, fx_mul(E.inv, wsub(fx_ceil(E.y_lower), E.y_lower))) : x_bot;
    const int32_t min_x = x0 < x1 ? x0 : x1, max_x = x0 < x1 ? x1 : x0;
    const int32_t min_pix = fx_int(min_x), max_pix = fx_int(max_x);
    const int32_t full = cov << 8;  // Fixed::from(cov as f32 / 256.0), exact
    const int32_t r = (min_pix == max_pix) ? fx_mul(wsub(FX_ONE, fx_fract(fx_avg(max_x, min_x))), full)
                                           : fx_mul(wsub(FX_ONE, fx_fract(min_x)), FX_HALF);
    c.xc0 = E.step > 0 ? fx_mul(r, E.step) : r;
    c.st = E.step > 0 ? E.step : FX_ONE;
    if (min_pix >= W) return c;
    // every step at x <= 0 adds to cell 0: their sum is x_pix at x = 0
    const int64_t k0 = min_pix < 0 ? -(int64_t)min_pix : 0;
    c.k_s = (int32_t)(k0 < 65537 ? k0 : 65537);
    c.x_s = min_pix < 0 ? 0 : min_pix;
    c.n = W - c.x_s;
    return c;
}

// The row's cells in order: f(j, delta of cell x_s + j) for j = 0, 1, ...
// up to the first cell whose covered fraction reaches cov, or j = n - 1.
// x_cov = min(xc0 + k * st, ONE) is walked from k = k_s in 32 bits: it
// starts at or below ONE (xc0 >= 0), stays at ONE once capped, and below the
// cap one more step (0 < st <= ONE) fits; on [0, ONE], pixel_cov(u) is
// (u + 128) >> 8, which reaches 256 >= cov at ONE, so the walk ends.
template <class F>
FT_HD inline void row_walk(const RowCells &c, F &&f) {
    const int64_t xk = c.xc0 + (int64_t)c.k_s * c.st;
    int32_t xc = xk < FX_ONE ? (int32_t)xk : FX_ONE;
    const int32_t st = (int32_t)c.st;
    int32_t prev = 0;
    for (int32_t j = 0; j < c.n; ++j) {
        const int32_t xp = (xc + 128) >> 8;
        const int32_t cur = xp < c.cov ? xp : c.cov;
        xc = xc < FX_ONE - st ? xc + st : FX_ONE;
        f(j, cur - prev);
        if (cur >= c.cov) break;
        prev = cur;
    }
}

// row_walk in closed form, so that one long walk can be split over many
// lanes.  The walk's x_cov at cell j is min(xk + j * st, ONE) with
// xk = xc0 + k_s * st, and it ends at the first cell whose (x_cov + 128) >> 8
// reaches cov, i.e. whose xk + j * st >= cov * 256 - 128 (below ONE, as
// cov <= 256), or at cell n - 1.
// The number of cells row_walk visits.
FT_HD inline int32_t row_cell_count(const RowCells &c) {
    if (c.n <= 0) return 0;
    const int64_t xk = c.xc0 + (int64_t)c.k_s * c.st;
    const int64_t t = (int64_t)c.cov * 256 - 128;
    const int64_t last = xk >= t ? 0 : (t - xk + c.st - 1) / c.st;
    return last + 1 < (int64_t)c.n ? (int32_t)(last + 1) : c.n;
}

// row_cell_count(c) > m, without the division: cell m - 1 does not end the
// walk.
FT_HD inline bool row_cells_exceed(const RowCells &c, int32_t m) {
    return (c.n > m) & ((m <= 0) | (c.xc0 + ((int64_t)c.k_s + m - 1) * c.st < (int64_t)c.cov * 256 - 128));
}

// The covered fraction after cell j (row_walk's `cur`), and cell j's delta.
FT_HD inline int32_t row_cell_cov(const RowCells &c, int32_t j) {
    const int64_t x = c.xc0 + ((int64_t)c.k_s + j) * c.st;
    const int32_t xc = x < FX_ONE ? (int32_t)x : FX_ONE;
    const int32_t xp = (xc + 128) >> 8;
    return xp < c.cov ? xp : c.cov;
}

FT_HD inline int32_t row_cell_delta(const RowCells &c, int32_t j) {
    return row_cell_cov(c, j) - (j > 0 ? row_cell_cov(c, j - 1) : 0);
}

template <class A>
FT_HD inline void edge_row_cells(const EdgeRec &E, int32_t sr, int32_t ed, int32_t W, int32_t *cells) {
    const RowCells c = row_setup(E, sr, W);
    int32_t *const at = cells + c.x_s;
    row_walk(c, [&](int32_t j, int32_t d) {
        if (d) A::add_cell(&at[j], d * ed);
    });
}

FT_HD inline uint32_t map_coverage(int32_t sum, uint32_t rule) {
    return rule ? coverage_even_odd(sum) : coverage_non_zero(sum);
}

// SrcOver with coverage alpha a for one premultiplied pixel of N channels
// (pix Pixel::composite_channels_alpha; alpha is the last channel).
// sw: source channels widened by ch8_widen.
template <int N>
FT_HD inline void composite_px(uint8_t *d, const uint32_t *sw, uint32_t a) {
    const uint32_t aw = ch8_widen(a);
    const uint32_t sa = ch8_mul_w(sw[N - 1], aw);
    const uint32_t invw = ch8_widen(255u - sa);
    for (int c = 0; c < N; ++c) d[c] = (uint8_t)sat_add_u8(ch8_mul_w(sw[c], aw), ch8_mul_w(ch8_widen(d[c]), invw));
}

// The same over a zero destination (a fresh raster): d * x = 0.
template <int N>
FT_HD inline void composite_px_fresh(uint8_t *d, const uint32_t *sw, uint32_t a) {
    const uint32_t aw = ch8_widen(a);
    for (int c = 0; c < N; ++c) d[c] = (uint8_t)ch8_mul_w(sw[c], aw);
}

// Band geometry.  A band is the unit of edge binning; its rows are rasterised
// in groups whose i32 cells fit the LDS budget, one raster block per group.
// Bands are at least `band_rows_min` rows tall so that wide rasters (one row
// per group) do not list every edge once per row.  Both figures are tunable
// per batch (Tuning); the defaults suit the measured workloads.
constexpr uint32_t LDS_CELLS = 4096;  // default i32 cells per raster block (16 KiB)
constexpr uint32_t BAND_ROWS_MAX = 64;
constexpr uint32_t BAND_ROWS_MIN = 16;
constexpr uint32_t CELL_ALIGN = 16;   // row stride of cells in LDS

struct Tuning {
    uint32_t band_rows_min = BAND_ROWS_MIN;  // 1 .. 256
    uint32_t lds_cells = LDS_CELLS;          // 1024 .. 32768, a multiple of CELL_ALIGN (and fitting LDS)
    uint32_t dst_prefetch = 1;               // raster blocks load their destination chunks first
    uint32_t tile_groups = 0;                // paths of at most this many LDS groups go to the tile
                                             // scheduler (0: none)
};

// Row stride of the LDS cells: the width rounded up to CELL_ALIGN (the row
// pass reads whole 16-cell chunks, and cells past the width read as 0), plus
// 4 so that the same column of consecutive rows falls in different LDS banks
// (a 32-bank LDS sees a multiple of 32 words as the same bank; the (edge,
// row) pairs of one edge sit in consecutive rows at nearby columns).
FT_HD inline uint32_t cell_stride(uint32_t width) {
    return (width + CELL_ALIGN - 1) / CELL_ALIGN * CELL_ALIGN + 4u;
}

// rows: band height; group: rows per LDS group; cells: LDS cells per block.
FT_HD inline void band_geometry(uint32_t width, uint32_t height, const Tuning &t, uint32_t *rows, uint32_t *group,
                                uint32_t *n_bands, uint32_t *cells) {
    const uint32_t wp = cell_stride(width);
    uint32_t g = wp ? t.lds_cells / wp : BAND_ROWS_MAX;
    if (g > BAND_ROWS_MAX) g = BAND_ROWS_MAX;
    if (g < 1) g = 1;
    uint32_t r = g < t.band_rows_min ? t.band_rows_min : g;
    if (height && r > height) r = height;
    if (r < 1) r = 1;
    if (g > r) g = r;
    *rows = r;
    *group = g;
    *n_bands = height ? (height + r - 1) / r : 0;
    *cells = g * wp;
}

// Device form of one batch path (ft_fill_desc -> PathDev); band_base is set
// by the caller.
inline PathDev path_from_desc(const ft_fill_desc &d, uint32_t extra_flags, const Tuning &t = Tuning()) {
    PathDev P;
    P.op_begin = (uint32_t)d.op_begin;
    P.op_count = d.op_count;
    P.rule = d.rule;
    P.fmt = d.fmt;
    P.dst = d.dst;
    P.width = d.width;
    P.height = d.height;
    P.pitch = d.pitch;
    P.clr = (uint32_t)d.clr[0] | ((uint32_t)d.clr[1] << 8) | ((uint32_t)d.clr[2] << 16) | ((uint32_t)d.clr[3] << 24);
    for (int k = 0; k < 6; ++k) P.tr[k] = d.transform[k];
    const float tol = d.tolerance > 0.01f ? d.tolerance : 0.01f;  // set_tolerance, plotter.rs:133-137
    P.tol_sq = tol * tol;
    P.s_width0 = d.pen_width;
    uint32_t rows, group, nb, cells;
    band_geometry(d.width, d.height, t, &rows, &group, &nb, &cells);
    if (!d.width) nb = 0;
    P.band_rows = rows;
    P.group_rows = group;
    P.band_base = 0;
    P.n_bands = nb;
    P.flags = (d.fresh ? PATH_FRESH : 0u) | extra_flags;
    return P;
}

}  // namespace ft
