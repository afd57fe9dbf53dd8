// ft_stroke_dev.h -- the stroker (src/stroker.rs) as data-parallel stages,
// host- and device-compilable like ft_stages.h.
//
// Input: the flattened stroke events (x, y, width, kind) of one path.
// Stroke::add_point / close (stroker.rs:197-233) become flags and scans: a
// point is stored if it starts a run (events between closes) or differs from
// the previous point; each run with a stored point is a sub-stroke; a close
// sets `joined` of the last sub-stroke that has points, so a sub's joined
// flag is that of the last close before the next sub starts.  65,535 points
// at most are stored (stroker.rs:205-206).
//
// Output: Stroke::path_ops (stroker.rs:239-416) as fill events, directly:
// every (sub-stroke, side, segment) item emits its offset line, join (miter,
// bevel or a round arc) and closing ops, and Plotter::fill (plotter.rs:
// 339-350) sees them as Line/Close ops of one path, transformed again (the
// reference transforms stroke points twice) and closed once more at the end.
//
// pointy's helpers are restated as the host stroker and the oracle state
// them: right = (-y, x); normalize divides by hypot; angle_rel is the
// difference of atan2 angles folded into [-pi, pi]; Line::intersection is the
// cross-product form.  hypot is glibc 2.35's hypotf exactly (the float result
// of the double-precision root); atan2 and sin come from the policy `M`,
// because the device cannot call glibc: they only decide which join is drawn.
#pragma once

#include <math.h>
#include <stdint.h>

#include "ft_stages.h"

namespace ft {

struct SV2 {
    float x, y;
};

FT_HD inline SV2 sv_sub(SV2 a, SV2 b) {
    SV2 r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}

FT_HD inline SV2 sv_right(SV2 v) {
    SV2 r;
    r.x = -v.y;
    r.y = v.x;
    return r;
}

FT_HD inline float sv_hypot(float x, float y) {
    const double dx = x, dy = y;
    return (float)sqrt(dx * dx + dy * dy);
}

FT_HD inline SV2 sv_normalize(SV2 v) {
    const float m = sv_hypot(v.x, v.y);
    SV2 r;
    r.x = m > 0.0f ? v.x / m : 0.0f;
    r.y = m > 0.0f ? v.y / m : 0.0f;
    return r;
}

template <class M>
FT_HD inline float sv_angle_rel(SV2 a, SV2 b) {
    const float pi = 3.14159265358979323846f;
    const float th = M::atan2(a.y, a.x) - M::atan2(b.y, b.x);
    if (th < -pi) return th + pi + pi;
    if (th > pi) return th - pi - pi;
    return th;
}

FT_HD inline bool sv_intersection(SV2 a0, SV2 a1, SV2 b0, SV2 b1, SV2 &out) {
    const SV2 av = sv_sub(a0, a1), bv = sv_sub(b0, b1);
    const float den = av.x * bv.y - av.y * bv.x;
    if (den == 0.0f) return false;
    const float ca = a0.x * a1.y - a0.y * a1.x;
    const float cb = b0.x * b1.y - b0.y * b1.x;
    out.x = (bv.x * ca - av.x * cb) / den;
    out.y = (bv.y * ca - av.y * cb) / den;
    return true;
}

// Stage S1: per flattened event.
struct StrokeFlags {
    bool is_pt;   // a point (else a close)
    bool first;   // first point of its run
    bool pushed;  // stored by Stroke::add_point (before the cap)
};

FT_HD inline StrokeFlags stroke_flags(uint32_t e, const F4 *ev) {
    const F4 a = ev[e], b = ev[e > 0 ? e - 1 : 0];
    StrokeFlags f;
    f.is_pt = a.w == 0.0f;
    const bool prev_pt = (e > 0) & (b.w == 0.0f);
    f.first = f.is_pt & !prev_pt;
    f.pushed = f.is_pt & (f.first | !((a.x == b.x) & (a.y == b.y)));
    return f;
}

// Sub-strokes: start (first stored point), point count, joined flag.
struct StrokeSub {
    uint32_t start, n;
    bool joined;
};

// SubStroke::len (stroker.rs:88-96), u16 arithmetic as the reference's Vid.
FT_HD inline uint32_t stroke_sub_len(const StrokeSub &s) {
    if (s.joined) return (uint32_t)(uint16_t)(s.n + 1);
    return s.n > 0 ? s.n - 1 : 0;
}

// Vertex of segment k on a side (stroker.rs:67-85): forward from the first
// point, reverse from the last.
FT_HD inline uint32_t stroke_vertex(const StrokeSub &s, bool fwd, uint32_t k) {
    const uint32_t m = k % s.n;
    return fwd ? s.start + m : s.start + (s.n - 1 - m);
}

struct Offset {
    SV2 r0, r1;  // the segment offset to its right by half of each end's width
};

// stroke_offset (stroker.rs:301-309)
FT_HD inline Offset stroke_offset(F4 p0, F4 p1) {
    SV2 d;
    d.x = p1.x - p0.x;
    d.y = p1.y - p0.y;
    const SV2 vr = sv_normalize(sv_right(d));
    const float h0 = p0.z / 2.0f, h1 = p1.z / 2.0f;
    Offset o;
    o.r0.x = p0.x + vr.x * h0;
    o.r0.y = p0.y + vr.y * h0;
    o.r1.x = p1.x + vr.x * h1;
    o.r1.y = p1.y + vr.y * h1;
    return o;
}

// Stroke::arc (stroker.rs:399-416) as a depth-first walk in the recursion's
// order: leaf = the node's end point; a node (a, b) splits at the point c on
// the circle of radius width/2 around p.  Returns 0, or FT_ERANGE past depth
// FT_FLATTEN_MAX_DEPTH or FT_FLATTEN_MAX_LEAVES leaves.
template <class Emit>
FT_HD uint32_t stroke_arc(F4 p, SV2 a0, SV2 b0, float tol_sq, uint32_t &leaves, Emit &&emit) {
    const float h = p.z / 2.0f;
    SV2 a = a0, b = b0;
    uint32_t depth = 0, bits = 0;
    auto mid_of = [&](SV2 u, SV2 v) {
        const SV2 vr = sv_normalize(sv_right(sv_sub(v, u)));
        SV2 c;
        c.x = p.x + vr.x * h;
        c.y = p.y + vr.y * h;
        return c;
    };
    for (;;) {
        const SV2 c = mid_of(a, b);
        const float mx = (a.x + b.x) / 2.0f, my = (a.y + b.y) / 2.0f;
        const float dx = c.x - mx, dy = c.y - my;
        if (dx * dx + dy * dy <= tol_sq) {
            if (++leaves > FT_FLATTEN_MAX_LEAVES) return FT_ERANGE;
            emit(b);
            while (depth > 0 && (bits & 1u)) {
                bits >>= 1;
                --depth;
            }
            if (depth == 0) return 0;
            bits |= 1u;
            a = a0;
            b = b0;
            for (int lv = (int)depth - 1; lv >= 0; --lv) {
                const SV2 cc = mid_of(a, b);
                if ((bits >> lv) & 1u)
                    a = cc;
                else
                    b = cc;
            }
        } else {
            if (depth >= FT_FLATTEN_MAX_DEPTH) return FT_ERANGE;
            b = c;
            bits <<= 1;
            ++depth;
        }
    }
}

// Ops of Stroke::join (stroker.rs:330-396) at p between the offset segments
// (a0, a1) and (b0, b1), through `line(x, y)`.
template <class M, class Line>
FT_HD uint32_t stroke_join(uint32_t join, float miter, float tol_sq, F4 p, SV2 a0, SV2 a1, SV2 b0, SV2 b1,
                           uint32_t &leaves, Line &&line) {
    if (join == FT_JOIN_MITER) {
        if (miter > 0.0f) {
            const float sm_min = 1.0f / miter;
            const float th = sv_angle_rel<M>(sv_sub(a1, a0), sv_sub(b0, b1));
            const float sm = fabsf(M::sin(th / 2.0f));
            if ((sm >= sm_min) & (sm < 1.0f)) {
                SV2 xp;
                if (sv_intersection(a0, a1, b0, b1, xp)) {
                    line(xp);
                    return 0;
                }
            }
        }
        line(a1);
        line(b0);
        return 0;
    }
    if (join == FT_JOIN_BEVEL) {
        line(a1);
        line(b0);
        return 0;
    }
    const float th = sv_angle_rel<M>(sv_sub(a1, a0), sv_sub(b0, b1));
    if (th <= 0.0f) {
        line(a1);
        line(b0);
        return 0;
    }
    line(a1);
    return stroke_arc(p, a1, b0, tol_sq, leaves, line);
}

// Item t of the outline: side `fwd` of sub-stroke s, segment k (k <= s.n).
// Calls line(SV2) / close() in output order; returns 0 or FT_ERANGE.
// Items are laid out per sub-stroke: (n + 1) forward slots, then (n + 1)
// reverse slots, so item order is output order.
template <class M, class Line, class Close>
FT_HD uint32_t stroke_item(const StrokeSub &s, bool fwd, uint32_t k, const F4 *pts, uint32_t join, float miter,
                           float tol_sq, Line &&line, Close &&close) {
    const uint32_t len = stroke_sub_len(s);
    if (k >= len) return 0;
    const uint32_t v0 = stroke_vertex(s, fwd, k), v1 = stroke_vertex(s, fwd, k + 1);
    const F4 p0 = pts[v0], p1 = pts[v1];
    const Offset cur = stroke_offset(p0, p1);
    uint32_t err = 0, leaves = 0;
    if (k == 0) {
        if (!s.joined) line(cur.r0);
    } else {
        const uint32_t u0 = stroke_vertex(s, fwd, k - 1);
        const Offset prev = stroke_offset(pts[u0], p0);
        err = stroke_join<M>(join, miter, tol_sq, p0, prev.r0, prev.r1, cur.r0, cur.r1, leaves, line);
    }
    if (k + 1 == len) {
        if (!s.joined) line(cur.r1);
        if (!fwd || s.joined) close();
    }
    return err;
}

}  // namespace ft
