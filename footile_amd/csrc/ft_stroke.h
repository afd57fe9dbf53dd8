// ft_stroke.h -- host stroker: flattened wide points -> outline PathOps.
//
// Restates footile's Stroke (src/stroker.rs:44-416).  Input is the stream of
// (x, y, width) points and sub-path closes produced by the device flattening
// pass; output is Line/Close ops that Plotter::stroke fills NonZero
// (src/plotter.rs:356-365).
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/footile_hip.h"

namespace ft {

struct WidePoint {
    float x, y, w;
};

class Stroker {
  public:
    Stroker(ft_join join, float miter_limit, float tol_sq);
    void add_point(WidePoint p);     // stroker.rs:204-216
    void close(bool joined);         // stroker.rs:230-236
    ft_status path_ops(std::vector<ft_path_op> &out);  // stroker.rs:239-247

  private:
    struct Sub {
        uint16_t start, n_points;
        bool joined, done;
    };
    uint16_t sub_len(const Sub &s) const;
    uint16_t sub_next(const Sub &s, uint16_t vid, bool fwd) const;
    uint16_t next(uint16_t vid, bool fwd);
    void stroke_side(std::vector<ft_path_op> &out, size_t i, uint16_t start, bool fwd);
    void join(std::vector<ft_path_op> &out, WidePoint p, float a0x, float a0y, float a1x, float a1y, float b0x,
              float b0y, float b1x, float b1y);
    void arc(std::vector<ft_path_op> &out, WidePoint p, float ax, float ay, float bx, float by, int depth);

    ft_join join_;
    float miter_;
    float tol_sq_;
    std::vector<WidePoint> pts_;
    std::vector<Sub> subs_;
    ft_status err_ = FT_OK;
};

}  // namespace ft
