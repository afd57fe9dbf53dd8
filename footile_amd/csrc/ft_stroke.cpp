// ft_stroke.cpp -- host stroker (src/stroker.rs), fed by the device flattening.
//
// Vertex ids are u16 as in the reference (src/vid.rs), including its
// wrapping arithmetic; the point cap is 65,535 (stroker.rs:205-206).
// pointy 0.7 helpers used by the reference (Pt::right, normalize, angle_rel,
// Line::intersection) are restated here; libm sinf/atan2f/hypotf are called.
#include "ft_stroke.h"

#include <math.h>

namespace ft {

namespace {

struct V2 {
    float x, y;
};

inline V2 sub(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
// pointy Pt::right: the perpendicular (-y, x)
inline V2 right(V2 v) { return {-v.y, v.x}; }
// pointy Pt::normalize: v / |v| (hypot), zero vector stays zero
inline V2 normalize(V2 v) {
    float m = hypotf(v.x, v.y);
    if (m > 0.0f) return {v.x / m, v.y / m};
    return {0.0f, 0.0f};
}
// pointy Pt::angle_rel: difference of atan2 angles folded into [-pi, pi]
inline float angle_rel(V2 a, V2 b) {
    const float pi = 3.14159265358979323846f;
    float th = atan2f(a.y, a.x) - atan2f(b.y, b.x);
    if (th < -pi) return th + pi + pi;
    if (th > pi) return th - pi - pi;
    return th;
}
// pointy Line::intersection of the infinite lines a0-a1 and b0-b1
inline bool intersection(V2 a0, V2 a1, V2 b0, V2 b1, V2 &out) {
    V2 av = sub(a0, a1), bv = sub(b0, b1);
    float den = av.x * bv.y - av.y * bv.x;
    if (den == 0.0f) return false;
    float ca = a0.x * a1.y - a0.y * a1.x;
    float cb = b0.x * b1.y - b0.y * b1.x;
    out.x = (bv.x * ca - av.x * cb) / den;
    out.y = (bv.y * ca - av.y * cb) / den;
    return true;
}

inline void push_line(std::vector<ft_path_op> &out, float x, float y) {
    ft_path_op op{};
    op.tag = FT_OP_LINE;
    op.v[0] = x;
    op.v[1] = y;
    out.push_back(op);
}

inline void push_close(std::vector<ft_path_op> &out) {
    ft_path_op op{};
    op.tag = FT_OP_CLOSE;
    out.push_back(op);
}

}  // namespace

Stroker::Stroker(ft_join join, float miter_limit, float tol_sq) : join_(join), miter_(miter_limit), tol_sq_(tol_sq) {
    pts_.reserve(1024);
    subs_.push_back(Sub{0, 0, false, false});
}

void Stroker::add_point(WidePoint p) {
    if (pts_.size() >= FT_MAX_POINTS) return;
    const bool done = subs_.back().done;
    if (done) subs_.push_back(Sub{(uint16_t)pts_.size(), 0, false, false});
    const bool coincident = !pts_.empty() && pts_.back().x == p.x && pts_.back().y == p.y;
    if (done || !coincident) {
        pts_.push_back(p);
        subs_.back().n_points = (uint16_t)(subs_.back().n_points + 1);
    }
}

void Stroker::close(bool joined) {
    if (!pts_.empty()) {
        subs_.back().joined = joined;
        subs_.back().done = true;
    }
}

// SubStroke::len (stroker.rs:88-96); u16 arithmetic wraps.
uint16_t Stroker::sub_len(const Sub &s) const {
    if (s.joined) return (uint16_t)(s.n_points + 1);
    if (s.n_points > 0) return (uint16_t)(s.n_points - 1);
    return 0;
}

// SubStroke::next (stroker.rs:67-85)
uint16_t Stroker::sub_next(const Sub &s, uint16_t vid, bool fwd) const {
    if (fwd) {
        uint16_t v = (uint16_t)(vid + 1);
        return v < (uint16_t)(s.start + s.n_points) ? v : s.start;
    }
    if (vid > s.start) return (uint16_t)(vid - 1);
    return (uint16_t)(s.start + s.n_points - 1);
}

// Stroke::next via sub_at (stroker.rs:176-192)
uint16_t Stroker::next(uint16_t vid, bool fwd) {
    for (const Sub &s : subs_)
        if (vid < (uint16_t)(s.start + s.n_points)) return sub_next(s, vid, fwd);
    err_ = FT_EINVAL;
    return vid;
}

void Stroker::arc(std::vector<ft_path_op> &out, WidePoint p, float ax, float ay, float bx, float by, int depth) {
    if (err_) return;  // stroker.rs:399-416
    V2 vr = normalize(right(sub(V2{bx, by}, V2{ax, ay})));
    float h = p.w / 2.0f;
    V2 c{p.x + vr.x * h, p.y + vr.y * h};
    V2 ab{(ax + bx) / 2.0f, (ay + by) / 2.0f};
    float dx = c.x - ab.x, dy = c.y - ab.y;
    if (dx * dx + dy * dy <= tol_sq_) {
        push_line(out, bx, by);
    } else if (depth >= FT_FLATTEN_MAX_DEPTH) {
        err_ = FT_ERANGE;
    } else {
        arc(out, p, ax, ay, c.x, c.y, depth + 1);
        arc(out, p, c.x, c.y, bx, by, depth + 1);
    }
}

void Stroker::join(std::vector<ft_path_op> &out, WidePoint p, float a0x, float a0y, float a1x, float a1y, float b0x,
                   float b0y, float b1x, float b1y) {
    V2 a0{a0x, a0y}, a1{a1x, a1y}, b0{b0x, b0y}, b1{b1x, b1y};
    if (join_ == FT_JOIN_MITER) {  // stroker.rs:340-367
        if (miter_ > 0.0f) {
            float sm_min = 1.0f / miter_;
            float th = angle_rel(sub(a1, a0), sub(b0, b1));
            float sm = fabsf(sinf(th / 2.0f));
            if (sm >= sm_min && sm < 1.0f) {
                V2 xp;
                if (intersection(a0, a1, b0, b1, xp)) {
                    push_line(out, xp.x, xp.y);
                    return;
                }
            }
        }
        push_line(out, a1.x, a1.y);
        push_line(out, b0.x, b0.y);
    } else if (join_ == FT_JOIN_BEVEL) {  // stroker.rs:370-373
        push_line(out, a1.x, a1.y);
        push_line(out, b0.x, b0.y);
    } else {  // stroker.rs:380-396
        float th = angle_rel(sub(a1, a0), sub(b0, b1));
        if (th <= 0.0f) {
            push_line(out, a1.x, a1.y);
            push_line(out, b0.x, b0.y);
        } else {
            push_line(out, a1.x, a1.y);
            arc(out, p, a1.x, a1.y, b0.x, b0.y, 0);
        }
    }
}

// Stroke::stroke_side (stroker.rs:265-295) with stroke_offset (301-309).
void Stroker::stroke_side(std::vector<ft_path_op> &out, size_t i, uint16_t start, bool fwd) {
    bool have = false;
    V2 xr0{0, 0}, xr1{0, 0};
    uint16_t v0 = start;
    uint16_t v1 = next(v0, fwd);
    const bool joined = subs_[i].joined;
    const uint16_t count = sub_len(subs_[i]);
    for (uint32_t k = 0; k < count && !err_; ++k) {
        const WidePoint p0 = pts_[v0], p1 = pts_[v1];
        V2 vr = normalize(right(sub(V2{p1.x, p1.y}, V2{p0.x, p0.y})));
        float h0 = p0.w / 2.0f, h1 = p1.w / 2.0f;
        V2 pr0{p0.x + vr.x * h0, p0.y + vr.y * h0};
        V2 pr1{p1.x + vr.x * h1, p1.y + vr.y * h1};
        if (have)
            join(out, p0, xr0.x, xr0.y, xr1.x, xr1.y, pr0.x, pr0.y, pr1.x, pr1.y);
        else if (!joined)
            push_line(out, pr0.x, pr0.y);
        have = true;
        xr0 = pr0;
        xr1 = pr1;
        v0 = v1;
        v1 = next(v1, fwd);
    }
    if (!joined && have) push_line(out, xr1.x, xr1.y);
}

ft_status Stroker::path_ops(std::vector<ft_path_op> &out) {
    for (size_t i = 0; i < subs_.size() && !err_; ++i) {  // stroker.rs:239-262
        if (sub_len(subs_[i]) == 0) continue;
        const uint16_t start = subs_[i].start;
        const uint16_t end = sub_next(subs_[i], start, false);
        const bool joined = subs_[i].joined;
        stroke_side(out, i, start, true);
        if (joined) push_close(out);
        stroke_side(out, i, end, false);
        push_close(out);
    }
    return err_;
}

}  // namespace ft
