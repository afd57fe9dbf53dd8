/* diff_oracle.h -- FP64 CPU oracle of the differentiable fill (test
 * infrastructure only: tests/ and bench.py's checks load it; the product never
 * does).
 *
 * footile has no differentiable path; this is the specification the GPU's
 * float32 forward and hand-written adjoint are checked against.  The model:
 *
 *  - scene: images of W x H pixels; paths in composite order, each with an
 *    image index, a start point and n >= 1 cubics (3 control points each);
 *    premultiplied RGBA colour per path.
 *  - flattening: every cubic becomes FD_SEG segments at the dyadic parameters
 *    t = j / FD_SEG (Bernstein form), so the outline is a polygon whose
 *    vertices are fixed linear combinations of the control points; the last
 *    vertex closes back to the start point.
 *  - coverage of pixel square [X, X+1] x [Y, Y+1] by a path: the integral of
 *    its winding number over the square, A = sum over edges of
 *    sign * h * mean(clamp(X + 1 - x(y), 0, 1)) over the edge's part within
 *    the row (h its height there, sign +1 for a descending edge), and
 *    alpha = min(|A|, 1).
 *  - compositing: C_k = c_k * alpha_k + C_{k-1} * (1 - a_k * alpha_k) per
 *    channel from C_0 = 0 (premultiplied SrcOver in path order).
 *  - loss: mean over all channels of (C - T)^2.
 *
 * fd_forward renders and returns the loss; fd_backward returns d loss / d
 * control points and d loss / d colours (analytic adjoint in FP64).
 */
#ifndef DIFF_ORACLE_H
#define DIFF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FD_SEG 16         /* segments per cubic */
#define FD_EPS_W 1e-5     /* below this x-extent an edge piece is treated as vertical */

typedef struct fd_scene {
    uint32_t n_images, width, height;
    uint32_t n_paths;
    const uint32_t *path_image;  /* [n_paths], non-decreasing */
    const uint32_t *path_cubics; /* [n_paths], >= 1 */
} fd_scene;

/* points: per path (1 + 3 n) (x, y) pairs, paths back to back.
 * colours: [n_paths][4]; target: [n_images][H][W][4] or NULL (loss = 0).
 * canvas (out, may be NULL): [n_images][H][W][4]. */
double fd_forward(const fd_scene *s, const double *points, const double *colours, const double *target,
                  double *canvas);

/* VJP of the canvas: grad_canvas [n_images][H][W][4] -> grad_points, grad_colours.
 * If grad_canvas is NULL the loss gradient 2 (C - T) / N against `target` is
 * used.  Returns the loss (or 0). */
double fd_backward(const fd_scene *s, const double *points, const double *colours, const double *target,
                   const double *grad_canvas, double *grad_points, double *grad_colours);

/* Coverage A (signed, before the clamp) of one pixel by one path. */
double fd_pixel_area(const fd_scene *s, const double *points, uint32_t path, uint32_t x, uint32_t y);

/* Flattened polygon of one path: writes up to cap (x, y) vertices, returns the count. */
uint32_t fd_polygon(const fd_scene *s, const double *points, uint32_t path, double *xy, uint32_t cap);

#ifdef __cplusplus
}
#endif

#endif
