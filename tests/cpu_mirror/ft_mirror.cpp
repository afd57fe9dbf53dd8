// ft_mirror.cpp -- sequential host run of the device pipeline's stage code.
//
// TEST INFRASTRUCTURE.  Runs the functions of footile_amd/csrc/ft_stages.h --
// the same per-item code the gfx950 kernels execute -- one item at a time on
// the CPU, with every intermediate buffer sized exactly to its logical length,
// so that an AddressSanitizer build flags any index the device code would
// take past the data it wrote.  The kernels' wavefront-level parts (scans,
// LDS row scan) are replaced by plain sequential loops.
#include <stdint.h>
#include <string.h>

#include <math.h>
#include <vector>

#include "../../footile_amd/csrc/ft_stages.h"
#include "../../footile_amd/csrc/ft_stroke_dev.h"

using namespace ft;

namespace {

struct HostAtom {
    static void max_i32(int32_t *p, int32_t v) {
        if (v > *p) *p = v;
    }
    static void max_u32(uint32_t *p, uint32_t v) {
        if (v > *p) *p = v;
    }
    static void min_u32(uint32_t *p, uint32_t v) {
        if (v < *p) *p = v;
    }
    static uint32_t add_u32(uint32_t *p, uint32_t v) {
        uint32_t o = *p;
        *p = o + v;
        return o;
    }
    static void add_cell(int32_t *p, int32_t v) { *p = (int32_t)((uint32_t)*p + (uint32_t)v); }
};

// Every array holds at least one element: the stage code reads index 0 as an
// in-range stand-in for lanes with nothing to do, as the device buffers allow.
template <class T>
std::vector<T> buf(size_t n) {
    return std::vector<T>(n ? n : 1);
}

struct HostMath {  // the C library's float functions, as the oracle calls them
    static float atan2(float y, float x) { return atan2f(y, x); }
    static float sin(float t) { return sinf(t); }
};

// The stroker stages in order, on the host: flattened stroke events -> the
// outline's fill events (and run tables).
int stroke_outline(const F4 *ev, uint32_t n_ev, const PathDev &P, uint32_t join, float miter, uint32_t &n_out,
                   uint32_t &n_closes_out, std::vector<I2> &ev_xy, std::vector<uint32_t> &ev_meta,
                   std::vector<int32_t> &run_first, std::vector<uint32_t> &run_close,
                   std::vector<uint32_t> &run_path, std::vector<uint32_t> &run_start, std::vector<uint32_t> &run_end) {
    // S1: flags, their scans, compaction, sub-strokes
    std::vector<uint32_t> sidx(n_ev ? n_ev : 1), subi(n_ev ? n_ev : 1), lastc(n_ev ? n_ev : 1);
    uint32_t pushed = 0, subs = 0, lc = 0;
    for (uint32_t e = 0; e < n_ev; ++e) {
        const StrokeFlags f = stroke_flags(e, ev);
        sidx[e] = pushed;
        subi[e] = subs;
        pushed += f.pushed ? 1 : 0;
        subs += f.first ? 1 : 0;
        if (!f.is_pt) lc = e + 1;
        lastc[e] = lc;
    }
    const uint32_t n_kept = pushed < FT_MAX_POINTS ? pushed : FT_MAX_POINTS;
    auto spts = buf<F4>(n_kept);
    auto sub_start = buf<uint32_t>(subs), sub_event = buf<uint32_t>(subs);
    for (uint32_t e = 0; e < n_ev; ++e) {
        const StrokeFlags f = stroke_flags(e, ev);
        if (f.pushed && sidx[e] < FT_MAX_POINTS) spts[sidx[e]] = ev[e];
        if (f.first) {
            sub_start[subi[e]] = sidx[e];
            sub_event[subi[e]] = e;
        }
    }
    std::vector<StrokeSub> S;
    for (uint32_t s = 0; s < subs; ++s) {
        if (sub_start[s] >= FT_MAX_POINTS) break;
        const bool more = s + 1 < subs;
        const uint32_t a1 = more ? sub_start[s + 1] : n_kept;
        const bool last_kept = !more || a1 >= FT_MAX_POINTS;
        const uint32_t end = a1 < n_kept ? a1 : n_kept;
        const uint32_t range_end = last_kept ? n_ev : sub_event[s + 1];
        const uint32_t c = range_end ? lastc[range_end - 1] : 0;
        StrokeSub sub;
        sub.start = sub_start[s];
        sub.n = end - sub.start;
        sub.joined = c > 0 && c - 1 >= sub_event[s] && ev[c - 1].w == 2.0f;
        S.push_back(sub);
    }
    // S2: the outline, item by item (sub-stroke, side, segment), in output order
    const float tol_sq = P.tol_sq;
    uint32_t count_e = 0, count_c = 0, err = 0;
    for (const StrokeSub &sub : S)
        for (int side = 0; side < 2; ++side)
            for (uint32_t k = 0; k <= sub.n; ++k)
                err |= stroke_item<HostMath>(
                    sub, side == 0, k, spts.data(), join, miter, tol_sq, [&](SV2) { ++count_e; },
                    [&]() {
                        ++count_e;
                        ++count_c;
                    });
    if (err) return (int)err;
    n_out = count_e + 1;  // + Plotter::fill's final close
    n_closes_out = count_c + 1;
    ev_xy = buf<I2>(n_out);
    ev_meta = buf<uint32_t>(n_out);
    run_first = buf<int32_t>(n_closes_out);
    run_close = buf<uint32_t>(n_closes_out);
    run_path = buf<uint32_t>(n_closes_out);
    run_start = buf<uint32_t>(n_closes_out);
    run_end = buf<uint32_t>(n_closes_out);
    EventSink em{};
    em.mode = MODE_FILL;
    em.ev_xy = ev_xy.data();
    em.ev_meta = ev_meta.data();
    em.run_first = run_first.data();
    em.run_close = run_close.data();
    em.run_path = run_path.data();
    em.run_start = run_start.data();
    em.run_end = run_end.data();
    em.cap_events = n_out;
    em.cap_runs = n_closes_out;
    em.e = 0;
    em.run = 0;
    em.p = 0;
    for (const StrokeSub &sub : S)
        for (int side = 0; side < 2; ++side)
            for (uint32_t k = 0; k <= sub.n; ++k)
                stroke_item<HostMath>(
                    sub, side == 0, k, spts.data(), join, miter, tol_sq,
                    [&](SV2 q) { em.point(xform(P, q.x, q.y, 0.0f)); }, [&]() { em.close(true); });
    em.close(true);
    return em.bad ? FT_EDEVICE : 0;
}

}  // namespace

// The reference's pixel walk along one scan row (fig.rs:285-321), step by
// step, as a check of the set-up in row_setup and the walk in row_walk.
static void walk_row(const EdgeRec &E, int32_t sr, int32_t ed, int32_t W, int32_t *cells) {
    const int32_t r0 = fx_int(E.y_upper), r1 = fx_int(E.y_lower);
    const bool starting = sr == r0, ending = sr == r1;
    int32_t cov = ending ? pixel_cov(fx_fract(E.y_lower)) : 256;
    if (starting) cov -= pixel_cov(fx_fract(E.y_upper));
    if (cov <= 0) return;
    const int32_t x_bot = (int32_t)((uint32_t)E.x_bot0 + (uint32_t)(sr - r0) * (uint32_t)E.inv);
    const int32_t x0 = starting ? wsub(x_bot, fx_mul(E.inv, wsub(FX_ONE, fx_fract(E.y_upper)))) : wsub(x_bot, E.inv);
    const int32_t x1 = ending ? wsub(x_bot, fx_mul(E.inv, wsub(fx_ceil(E.y_lower), E.y_lower))) : x_bot;
    const int32_t min_x = x0 < x1 ? x0 : x1, max_x = x0 < x1 ? x1 : x0;
    const int32_t min_pix = fx_int(min_x), max_pix = fx_int(max_x);
    const int32_t full = cov << 8;
    const int32_t r = (min_pix == max_pix) ? fx_mul(wsub(FX_ONE, fx_fract(fx_avg(max_x, min_x))), full)
                                           : fx_mul(wsub(FX_ONE, fx_fract(min_x)), FX_HALF);
    int32_t x_cov = E.step > 0 ? fx_mul(r, E.step) : r;  // first_cov, then step_cov
    int32_t prev = 0;
    for (int64_t x = min_pix; x < W; ++x) {
        int32_t xp = pixel_cov(x_cov > FX_ONE ? FX_ONE : x_cov);
        xp = xp < cov ? xp : cov;
        const int32_t d = xp - prev;
        if (d) cells[x < 0 ? 0 : x] += d * ed;
        if (xp >= cov) break;
        prev = xp;
        const int32_t stp = E.step > 0 ? E.step : FX_ONE;
        x_cov = x_cov > FX_ONE ? x_cov : wadd(x_cov, stp);
    }
}

extern "C" {

// Random quads and cubics flattened by 2^sp lanes (flatten_share, sp = 0..6)
// against the whole tree walked from the root in one lane: the lanes' points
// in lane order must be the same points in the same order, whatever sp.
// Returns the number of mismatching (curve, sp) cases.
int ftm_split_check(uint64_t seed, int iters) {
    uint64_t z = seed;
    auto next = [&]() {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t t = z;
        t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
        t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
        return t ^ (t >> 31);
    };
    auto unif = [&](float lo, float hi) { return lo + (hi - lo) * (float)((next() >> 11) * (1.0 / 9007199254740992.0)); };
    int bad = 0;
    for (int it = 0; it < iters; ++it) {
        const float ext = (next() & 3) ? 256.0f : 4096.0f;
        const float tol = (next() & 1) ? 0.3f : 0.01f;
        auto wp = [&]() {
            WP p;
            p.x = unif(-16.0f, ext);
            p.y = unif(-16.0f, ext);
            p.w = unif(0.5f, 8.0f);
            return p;
        };
        const bool cubic = (next() & 1) != 0;
        Cubic c;
        c.a = wp();
        c.b = wp();
        c.c = wp();
        c.d = wp();
        Quad q;
        q.a = c.a;
        q.b = c.b;
        q.c = c.c;
        std::vector<WP> want, got;
        uint32_t leaves = 0, rc = 0;
        auto put_w = [&](WP p) { want.push_back(p); };
        rc = cubic ? flatten_walk(c, tol * tol, FT_FLATTEN_MAX_DEPTH, leaves, put_w)
                   : flatten_walk(q, tol * tol, FT_FLATTEN_MAX_DEPTH, leaves, put_w);
        for (uint32_t sp = 0; sp <= 6; ++sp) {
            got.clear();
            uint32_t rs = 0;
            for (uint32_t lane = 0; lane < (1u << sp); ++lane) {
                uint32_t lv = 0;
                auto put_g = [&](WP p) { got.push_back(p); };
                rs |= cubic ? flatten_share(c, tol * tol, lane, lv, put_g, sp)
                            : flatten_share(q, tol * tol, lane, lv, put_g, sp);
            }
            bool same = (rs != 0) == (rc != 0) && (rc != 0 || got.size() == want.size());
            for (size_t k = 0; same && rc == 0 && k < got.size(); ++k)
                same = got[k].x == want[k].x && got[k].y == want[k].y && got[k].w == want[k].w;
            if (!same) ++bad;
        }
    }
    return bad;
}

// Random edges and rows: cells from edge_row_cells (row_setup + row_walk) against
// walk_row (step by step).  Returns the number of mismatching rows.
int ftm_row_cells_check(uint64_t seed, int iters) {
    uint64_t z = seed;
    auto next = [&]() {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t t = z;
        t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
        t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
        return t ^ (t >> 31);
    };
    int bad = 0;
    std::vector<int32_t> a(64), b(64), c3(64);
    for (int it = 0; it < iters; ++it) {
        const int32_t W = 1 + (int32_t)(next() % 48);
        // a segment between two Fixed points within [-24, 72) px, mostly short
        const int32_t span = (next() & 3) ? 6 : 80;
        const int32_t ax = (int32_t)(next() % (uint64_t)(span << 16)) - (24 << 16);
        const int32_t ay = (int32_t)(next() % (uint64_t)(span << 16));
        const int32_t bx = (next() & 7) ? ax + (int32_t)(next() % (uint64_t)(span << 17)) - (span << 16)
                                        : (int32_t)(next() % (uint64_t)(96 << 16)) - (24 << 16);
        const int32_t by = (next() & 7) ? ay + (int32_t)(next() % (uint64_t)(span << 16)) + 1 : ay + 1;
        I2 p, q;
        p.x = ax;
        p.y = ay;
        q.x = bx;
        q.y = by;
        const EdgeRec E = make_edge(p, q);
        if (!(E.flags & 1u)) continue;
        const int32_t ed = (next() & 1) ? 1 : -1;
        for (int32_t sr = fx_int(E.y_upper); sr <= fx_int(E.y_lower); ++sr) {
            std::fill(a.begin(), a.end(), 0);
            std::fill(b.begin(), b.end(), 0);
            edge_row_cells<HostAtom>(E, sr, ed, W, a.data());
            walk_row(E, sr, ed, W, b.data());
            if (a != b) ++bad;
            // the closed form (row_cell_count / row_cell_delta) cell by cell
            // against the serial walk, and the division-free length test
            std::fill(c3.begin(), c3.end(), 0);
            const RowCells rc = row_setup(E, sr, W);
            const int32_t m = row_cell_count(rc);
            int32_t walked = 0;
            bool same = true;
            row_walk(rc, [&](int32_t j, int32_t d) {
                same = same && j == walked && j < m && d == row_cell_delta(rc, j);
                ++walked;
            });
            if (!same || walked != m) ++bad;
            for (int32_t k : {0, 1, 2, 5, 16, 40})
                if (row_cells_exceed(rc, k) != (m > k)) ++bad;
            for (int32_t j = 0; j < m; ++j) c3[rc.x_s + j] += row_cell_delta(rc, j) * ed;
            if (c3 != a) ++bad;
        }
    }
    return bad;
}

// Run a batch (host raster pointers in desc.dst).  mode: MODE_FILL, or
// MODE_STROKE_FLATTEN to only produce (x, y, w, kind) events into ev_out
// (capacity ev_cap, count in *n_ev).  Per-path offered/stored counts are
// written when the arrays are non-null.  Returns an ft_status.
int ftm_run(const ft_path_op *ops, size_t n_ops, const ft_fill_desc *descs, size_t n_paths, uint32_t mode,
            uint32_t extra_flags, uint64_t *offered, uint64_t *stored, F4 *ev_out, size_t ev_cap, size_t *n_ev,
            uint32_t join, float miter) {
    std::vector<uint32_t> op_path(n_ops, NO_IDX);
    std::vector<PathDev> paths(n_paths);
    std::vector<uint32_t> band_path;
    for (size_t k = 0; k < n_paths; ++k) {
        const ft_fill_desc &d = descs[k];
        if (d.op_begin > n_ops || d.op_count > n_ops - d.op_begin) return FT_EINVAL;
        for (uint64_t i = d.op_begin; i < d.op_begin + d.op_count; ++i) {
            if (op_path[i] != NO_IDX || ops[i].tag > FT_OP_PEN_WIDTH) return FT_EINVAL;
            op_path[i] = (uint32_t)k;
        }
        paths[k] = path_from_desc(d, extra_flags);
        paths[k].band_base = (uint32_t)band_path.size();
        band_path.insert(band_path.end(), paths[k].n_bands, (uint32_t)k);
    }
    // stroke widths: last PenWidth index (+1) before each op
    std::vector<uint32_t> lastpw;
    if (mode != MODE_FILL) {
        lastpw.assign(n_ops, 0);
        uint32_t m = 0;
        for (size_t i = 0; i < n_ops; ++i) {
            lastpw[i] = m;
            if (ops[i].tag == TAG_PEN_WIDTH) m = (uint32_t)i + 1;
        }
    }
    // stage 1: counts + exclusive scan; fill runs keep each lane's first
    // LEAF_CACHE points (point k of item t at leaf[k * n_items + t]), as the
    // device count pass does, and the emit copies lanes with no more leaves
    const size_t n_items = n_ops * OP_LANES;
    auto leaf = buf<I2>(mode == MODE_FILL ? n_items * LEAF_CACHE : 0);
    std::vector<uint32_t> lane_leaves(n_items);
    std::vector<unsigned long long> op_off(n_ops), sub_off(n_items);
    unsigned long long tot = 0;
    for (size_t i = 0; i < n_ops; ++i) {
        op_off[i] = tot;
        uint32_t leaves = 0;
        for (uint32_t s = 0; s < OP_LANES; ++s) {
            const size_t t = i * OP_LANES + s;
            OpCount c = count_op(ops, op_path.data(), paths.data(), (uint32_t)i, s, mode,
                                 mode == MODE_FILL ? nullptr : lastpw.data(),
                                 mode == MODE_FILL ? leaf.data() + t : nullptr, n_items);
            if (c.err) return (int)c.err;
            lane_leaves[t] = c.leaves;
            sub_off[i * OP_LANES + s] = tot - op_off[i];
            tot += pack_count(c);
            leaves += c.leaves;
        }
        if (leaves > FT_FLATTEN_MAX_LEAVES) return FT_ERANGE;
    }
    uint32_t n_events = (uint32_t)(tot >> 32), n_closes = (uint32_t)tot;  // the outline's, for a stroke
    // stage 2: emit
    auto ev_xy = buf<I2>(mode == MODE_FILL ? n_events : 0);
    auto ev_f = buf<F4>(mode == MODE_FILL ? 0 : n_events);
    auto ev_meta = buf<uint32_t>(mode == MODE_FILL ? n_events : 0);
    auto run_first = buf<int32_t>(n_closes);
    auto run_close = buf<uint32_t>(n_closes);
    auto run_path = buf<uint32_t>(n_closes), run_start = buf<uint32_t>(n_closes), run_end = buf<uint32_t>(n_closes);
    EventSink sink{};
    sink.mode = mode;
    sink.ev_xy = ev_xy.data();
    sink.ev_f = ev_f.data();
    sink.ev_meta = ev_meta.data();
    sink.run_first = run_first.data();
    sink.run_close = run_close.data();
    sink.run_path = run_path.data();
    sink.run_start = run_start.data();
    sink.run_end = run_end.data();
    sink.cap_events = n_events;
    sink.cap_runs = n_closes;
    uint32_t bad = 0;
    for (size_t i = 0; i < n_ops; ++i)
        for (uint32_t s = 0; s < OP_LANES; ++s)
            bad |= emit_op(ops, op_path.data(), paths.data(), (uint32_t)i, s, mode,
                           mode == MODE_FILL ? nullptr : lastpw.data(), op_off[i] + sub_off[i * OP_LANES + s], sink,
                           mode == MODE_FILL ? leaf.data() + i * OP_LANES + s : nullptr, lane_leaves[i * OP_LANES + s],
                           n_items);
    if (bad) return FT_EDEVICE;
    if (mode == MODE_STROKE_FLATTEN) {
        if (n_ev) *n_ev = n_events;
        if (ev_out) memcpy(ev_out, ev_f.data(), (n_events < ev_cap ? n_events : ev_cap) * sizeof(F4));
        return FT_OK;
    }
    if (mode == MODE_STROKE) {
        uint32_t n_out = 0, n_cl = 0;
        const int rc = stroke_outline(ev_f.data(), n_events, paths[0], join, miter, n_out, n_cl, ev_xy, ev_meta,
                                      run_first, run_close, run_path, run_start, run_end);
        if (rc) return rc;
        n_events = n_out;
        n_closes = n_cl;
    }
    Limits L{};
    L.n_ops = (uint32_t)n_ops;
    L.n_paths = (uint32_t)n_paths;
    L.n_bands = (uint32_t)band_path.size();
    L.n_events = n_events;
    L.n_runs = n_closes;
    // stage 3: ingest + kept + exclusive scan
    auto ps = buf<PathState>(n_paths);
    for (PathState &s : ps) {
        s.topkey = NO_KEY;
        s.topidx = NO_IDX;
        s.offered = s.repop = s.pt_base = s.n_edges = s.stored = 0;
    }
    auto last = buf<uint32_t>(n_events);
    for (uint32_t e = 0; e < n_events; ++e)
        ingest_event(e, L, ev_meta.data(), ev_xy.data(), run_first.data(), run_path.data(), last.data(), bad);
    for (uint32_t e = 1; e < n_events; ++e)  // inclusive running maximum
        if (last[e - 1] > last[e]) last[e] = last[e - 1];
    auto kept = buf<uint32_t>(n_events), idx = buf<uint32_t>(n_events);
    uint32_t n_kept = 0;
    for (uint32_t e = 0; e < n_events; ++e) {
        kept[e] = kept_event(e, L, ev_meta.data(), ev_xy.data(), run_first.data(), run_close.data(), last.data(), bad);
        idx[e] = n_kept;
        n_kept += kept[e];
    }
    // the device's usual form (no ingest pass: each run's first point and
    // closing walk found locally) must agree wherever its walk is short
    for (uint32_t e = 0; e < n_events; ++e) {
        uint32_t lw = 0, first_of = NO_IDX, b2 = 0;
        const uint32_t k = kept_event_local(e, L, ev_meta.data(), ev_xy.data(), run_close.data(), KEPT_WALK, b2, lw,
                                            first_of);
        if (!lw && (k != kept[e] || b2 != 0)) bad = 1;
        if (first_of != NO_IDX && run_first[first_of] != (int32_t)e) bad = 1;
    }
    L.cap_points = n_kept;
    // stage 4: cap, compaction, top point
    for (uint32_t p = 0; p < n_paths; ++p)
        cap_path(p, L, paths.data(), op_off.data(), ev_meta.data(), ev_xy.data(), kept.data(), idx.data(),
                 run_first.data(), ps.data(), bad);
    auto pts = buf<I2>(n_kept);
    auto pt_run = buf<uint32_t>(n_kept);
    for (uint32_t e = 0; e < n_events; ++e)
        compact_event(e, L, ev_meta.data(), ev_xy.data(), kept.data(), idx.data(), run_first.data(), run_path.data(),
                      run_start.data(), run_end.data(), pts.data(), pt_run.data(), ps.data(), bad);
    for (uint32_t i = 0; i < n_kept; ++i) {
        const TopCand c = top_cand(i, L, pts.data(), pt_run.data(), run_path.data(), bad);
        if (c.path != NO_IDX && c.key < ps[c.path].topkey) ps[c.path].topkey = c.key;
    }
    // stage 5: edges, band counts, scan, bin
    const size_t n_bands = band_path.size();
    auto edges = buf<EdgeRec>(n_kept);
    auto band_cnt = buf<uint32_t>(n_bands), band_off = buf<uint32_t>(n_bands), band_cur = buf<uint32_t>(n_bands);
    auto pt_band = buf<U2>(n_kept);
    for (uint32_t i = 0; i < n_kept; ++i) {
        const EdgeOut r = edge_point<HostAtom>(i, L, pt_run.data(), run_path.data(), run_start.data(),
                                               run_end.data(), paths.data(), pts.data(), edges.data(), ps.data(),
                                               bad);
        pt_band[i] = band_record(r);
        if (r.has) ps[r.path].n_edges++;
        if (r.banded) {
            band_cnt[r.g0] += 1u;
            if (r.g1 + 1 < n_bands) band_cnt[r.g1 + 1] -= 1u;
        }
    }
    uint32_t acc = 0;
    for (size_t b = 0; b < n_bands; ++b) {  // band differences -> edges per band
        acc += band_cnt[b];
        band_cnt[b] = acc;
    }
    uint32_t n_pairs = 0;
    for (size_t b = 0; b < n_bands; ++b) {
        band_off[b] = n_pairs;
        n_pairs += band_cnt[b];
    }
    L.n_pairs = n_pairs;
    auto band_list = buf<uint32_t>(n_pairs);
    for (uint32_t i = 0; i < n_kept; ++i) {
        const BinItem it = bin_item(i, pt_band.data());
        if (!it.banded) continue;
        for (uint32_t g = it.g0; g <= it.g1; ++g) {
            const uint32_t at = band_off[g] + band_cur[g]++;
            if (at < n_pairs)
                band_list[at] = i;
            else
                bad = 1;
        }
    }
    if (bad) return FT_EDEVICE;
    // stage 6: raster, one band at a time (k_raster with a sequential row scan)
    for (size_t gb = 0; gb < n_bands; ++gb) {
        const uint32_t p = band_path[gb];
        const PathDev &P = paths[p];
        const uint32_t R = P.band_rows, H = P.height, W = P.width;
        const uint32_t y0 = (uint32_t)(gb - P.band_base) * R;
        const uint32_t y1 = y0 + R < H ? y0 + R : H;
        const PathState &S = ps[p];
        const bool has_pts = S.topkey != NO_KEY && S.topidx != NO_IDX;
        const int32_t top_row = has_pts ? top_row_of(S.topkey) : 0;
        const int64_t first = !has_pts ? (int64_t)H : (top_row > 0 ? top_row : 0);
        const uint32_t bpp = P.fmt == FMT_RGBA8P ? 4 : (P.fmt == FMT_GRAYA8P ? 2 : 1);
        uint8_t *const base = reinterpret_cast<uint8_t *>(P.dst);
        if (P.flags & PATH_FRESH) {
            const uint32_t ze = first < (int64_t)y1 ? (uint32_t)first : y1;
            for (uint32_t r = y0; r < ze; ++r) memset(base + (size_t)r * P.pitch, 0, (size_t)W * bpp);
        }
        const uint32_t ry0 = first > (int64_t)y0 ? (uint32_t)first : y0;
        if ((int64_t)ry0 >= (int64_t)y1) continue;
        const uint32_t Wp = cell_stride(W);
        const uint32_t rows = y1 - ry0;
        std::vector<int32_t> cells((size_t)rows * Wp, 0);
        const int fwd = path_forward(S, L, pts.data(), pt_run.data(), run_start.data(), run_end.data(), bad) ? 1 : 0;
        const int32_t shift = top_row < 0 ? top_row : 0;
        for (uint32_t k = 0; k < band_cnt[gb]; ++k) {
            const EdgeRec E = edges[band_list[band_off[gb] + k]];
            const int32_t ed = ((int)((E.flags >> 1) & 1u) == fwd) ? 1 : -1;
            int64_t a = (int64_t)fx_int(E.y_upper) - shift, b = (int64_t)fx_int(E.y_lower) - shift;
            if (a < ry0) a = ry0;
            if (b > (int64_t)y1 - 1) b = (int64_t)y1 - 1;
            for (int64_t r = a; r <= b; ++r)
                edge_row_cells<HostAtom>(E, (int32_t)(r + shift), ed, (int32_t)W, cells.data() + (size_t)(r - ry0) * Wp);
        }
        uint32_t sw[4];
        for (int c = 0; c < 4; ++c) sw[c] = ch8_widen((P.clr >> (8 * c)) & 0xFFu);
        const bool fresh = (P.flags & PATH_FRESH) != 0;
        for (uint32_t r = 0; r < rows; ++r) {
            uint8_t *row = base + (size_t)(ry0 + r) * P.pitch;
            uint32_t sum = 0;
            for (uint32_t x = 0; x < W; ++x) {
                sum += (uint32_t)cells[(size_t)r * Wp + x];
                const uint32_t a = map_coverage((int32_t)sum, P.rule);
                if (P.fmt == FMT_MATTE8)
                    row[x] = (uint8_t)a;
                else if (P.fmt == FMT_RGBA8P)
                    fresh ? composite_px_fresh<4>(row + 4 * x, sw, a) : composite_px<4>(row + 4 * x, sw, a);
                else
                    fresh ? composite_px_fresh<2>(row + 2 * x, sw, a) : composite_px<2>(row + 2 * x, sw, a);
            }
        }
    }
    for (size_t k = 0; k < n_paths; ++k) {
        if (offered) offered[k] = ps[k].offered;
        if (stored) stored[k] = ps[k].stored;
    }
    return bad ? FT_EDEVICE : FT_OK;
}

}  // extern "C"
