/* diff_oracle.c -- FP64 CPU oracle of the differentiable fill (test
 * infrastructure only).  The model is stated in diff_oracle.h.
 *
 * The adjoint is the exact derivative of the forward below, written out by
 * hand: an edge's row piece is carried with its partials with respect to the
 * edge's upper and lower end points (a 4-partial dual number), and the
 * per-pixel mean of clamp(u, 0, 1) is differentiated in closed form.  Kinks
 * (a vertex exactly on a row line, |A| = 1, u at 0 or 1) take the one-sided
 * choices written below; the GPU kernels take the same ones.
 */
#include "diff_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double v, d[4]; /* value, d/d(x_top, y_top, x_bot, y_bot) */
} d4;

static d4 d4_const(double c) {
    d4 r;
    memset(&r, 0, sizeof r);
    r.v = c;
    return r;
}

static d4 d4_var(double v, int i) {
    d4 r = d4_const(v);
    r.d[i] = 1.0;
    return r;
}

static d4 d4_add(d4 a, d4 b) {
    d4 r;
    r.v = a.v + b.v;
    for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}

static d4 d4_sub(d4 a, d4 b) {
    d4 r;
    r.v = a.v - b.v;
    for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}

static d4 d4_mul(d4 a, d4 b) {
    d4 r;
    r.v = a.v * b.v;
    for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}

static d4 d4_div(d4 a, d4 b) {
    d4 r;
    r.v = a.v / b.v;
    for (int i = 0; i < 4; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}

static double clamp01(double u) { return u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u); }

/* Mean of clamp(u, 0, 1) over u moving linearly from ua to ub, and its
 * partials.  Pieces narrower than FD_EPS_W use the midpoint value. */
static double mean_clamp(double ua, double ub, double *d_ua, double *d_ub) {
    const double lo = ua < ub ? ua : ub, hi = ua < ub ? ub : ua;
    double m, dlo = 0.0, dhi = 0.0;
    if (hi <= 0.0) {
        m = 0.0;
    } else if (lo >= 1.0) {
        m = 1.0;
    } else {
        const double w = hi - lo;
        if (w < FD_EPS_W) {
            const double mid = (lo + hi) * 0.5;
            if (mid <= 0.0) {
                m = 0.0;
            } else if (mid >= 1.0) {
                m = 1.0;
            } else {
                m = mid;
                dlo = dhi = 0.5;
            }
        } else {
            const double a = clamp01(lo), b = clamp01(hi);
            const double integral = (b - a) * (a + b) * 0.5 + (hi > 1.0 ? hi - 1.0 : 0.0);
            m = integral / w;
            dhi = (b - m) / w;
            dlo = (m - a) / w;
        }
    }
    if (ua <= ub) {
        *d_ua = dlo;
        *d_ub = dhi;
    } else {
        *d_ua = dhi;
        *d_ub = dlo;
    }
    return m;
}

/* Contribution of edge (x0, y0) -> (x1, y1) to the winding integral over
 * pixel (X, Y), and its partials with respect to (x0, y0, x1, y1). */
static double edge_area(double x0, double y0, double x1, double y1, int X, int Y, double df[4]) {
    df[0] = df[1] = df[2] = df[3] = 0.0;
    if (y0 == y1) return 0.0;
    const int down = y1 > y0;
    const double sign = down ? 1.0 : -1.0;
    const d4 xt = d4_var(down ? x0 : x1, 0), yt = d4_var(down ? y0 : y1, 1);
    const d4 xbt = d4_var(down ? x1 : x0, 2), ybt = d4_var(down ? y1 : y0, 3);
    const d4 ya = yt.v > (double)Y ? yt : d4_const((double)Y);
    const d4 yb = ybt.v < (double)(Y + 1) ? ybt : d4_const((double)(Y + 1));
    if (yb.v <= ya.v) return 0.0;
    const d4 s = d4_div(d4_sub(xbt, xt), d4_sub(ybt, yt));
    const d4 xa = d4_add(xt, d4_mul(d4_sub(ya, yt), s));
    const d4 xb = d4_add(xt, d4_mul(d4_sub(yb, yt), s));
    const d4 h = d4_sub(yb, ya);
    double mua, mub;
    const double M = mean_clamp((double)X + 1.0 - xa.v, (double)X + 1.0 - xb.v, &mua, &mub);
    double dF[4];
    for (int i = 0; i < 4; ++i) dF[i] = sign * (h.d[i] * M - h.v * (mua * xa.d[i] + mub * xb.d[i]));
    if (down) {
        for (int i = 0; i < 4; ++i) df[i] = dF[i];
    } else {
        df[0] = dF[2];
        df[1] = dF[3];
        df[2] = dF[0];
        df[3] = dF[1];
    }
    return sign * h.v * M;
}

/* edge_area without the partials. */
static double edge_area_v(double x0, double y0, double x1, double y1, int X, int Y) {
    if (y0 == y1) return 0.0;
    const int down = y1 > y0;
    const double xt = down ? x0 : x1, yt = down ? y0 : y1, xbt = down ? x1 : x0, ybt = down ? y1 : y0;
    const double ya = yt > (double)Y ? yt : (double)Y;
    const double yb = ybt < (double)(Y + 1) ? ybt : (double)(Y + 1);
    if (yb <= ya) return 0.0;
    const double sl = (xbt - xt) / (ybt - yt);
    const double xa = xt + (ya - yt) * sl, xb = xt + (yb - yt) * sl;
    double mua, mub;
    const double M = mean_clamp((double)X + 1.0 - xa, (double)X + 1.0 - xb, &mua, &mub);
    return (down ? 1.0 : -1.0) * (yb - ya) * M;
}

static void bernstein(int j, double w[4]) {
    const double t = (double)j / FD_SEG, u = 1.0 - t;
    w[0] = u * u * u;
    w[1] = 3.0 * u * u * t;
    w[2] = 3.0 * u * t * t;
    w[3] = t * t * t;
}

static uint64_t *point_offsets(const fd_scene *s) {
    uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (s->n_paths + 1));
    off[0] = 0;
    for (uint32_t k = 0; k < s->n_paths; ++k) off[k + 1] = off[k] + 1 + 3 * (uint64_t)s->path_cubics[k];
    return off;
}

static uint32_t polygon_of(const fd_scene *s, const double *P, uint32_t k, double *xy) {
    const uint32_t n = s->path_cubics[k];
    xy[0] = P[0];
    xy[1] = P[1];
    uint32_t v = 1;
    for (uint32_t c = 0; c < n; ++c) {
        const double *p0 = P + 2 * (3 * c), *p1 = p0 + 2, *p2 = p0 + 4, *p3 = p0 + 6;
        for (int j = 1; j <= FD_SEG; ++j) {
            double w[4];
            bernstein(j, w);
            xy[2 * v] = w[0] * p0[0] + w[1] * p1[0] + w[2] * p2[0] + w[3] * p3[0];
            xy[2 * v + 1] = w[0] * p0[1] + w[1] * p1[1] + w[2] * p2[1] + w[3] * p3[1];
            ++v;
        }
    }
    return v;
}

uint32_t fd_polygon(const fd_scene *s, const double *points, uint32_t path, double *xy, uint32_t cap) {
    uint64_t *off = point_offsets(s);
    const uint32_t V = 1 + FD_SEG * s->path_cubics[path];
    double *tmp = (double *)malloc(sizeof(double) * 2 * V);
    polygon_of(s, points + 2 * off[path], path, tmp);
    const uint32_t n = V < cap ? V : cap;
    memcpy(xy, tmp, sizeof(double) * 2 * n);
    free(tmp);
    free(off);
    return V;
}

/* Per-path polygons and per-(path, row) edge lists (CSR). */
typedef struct {
    uint64_t *vtx_off;  /* [n_paths + 1] */
    double *vtx;        /* (x, y) */
    uint64_t *row_off;  /* [n_paths * H + 1] */
    uint32_t *row_edge; /* local edge index within the path */
} geom;

static void geom_build(const fd_scene *s, const double *points, geom *g) {
    const uint32_t H = s->height;
    uint64_t *poff = point_offsets(s);
    g->vtx_off = (uint64_t *)malloc(sizeof(uint64_t) * (s->n_paths + 1));
    g->vtx_off[0] = 0;
    for (uint32_t k = 0; k < s->n_paths; ++k) g->vtx_off[k + 1] = g->vtx_off[k] + 1 + FD_SEG * (uint64_t)s->path_cubics[k];
    g->vtx = (double *)malloc(sizeof(double) * 2 * (g->vtx_off[s->n_paths] + 1));
    for (uint32_t k = 0; k < s->n_paths; ++k) polygon_of(s, points + 2 * poff[k], k, g->vtx + 2 * g->vtx_off[k]);
    const uint64_t cells = (uint64_t)s->n_paths * H;
    g->row_off = (uint64_t *)calloc(cells + 1, sizeof(uint64_t));
    for (int pass = 0; pass < 2; ++pass) {
        uint64_t *cur = NULL;
        if (pass == 1) {
            for (uint64_t c = 0; c < cells; ++c) g->row_off[c + 1] += g->row_off[c];
            g->row_edge = (uint32_t *)malloc(sizeof(uint32_t) * (g->row_off[cells] + 1));
            cur = (uint64_t *)malloc(sizeof(uint64_t) * (cells + 1));
            memcpy(cur, g->row_off, sizeof(uint64_t) * (cells + 1));
        }
        for (uint32_t k = 0; k < s->n_paths; ++k) {
            const uint64_t V = g->vtx_off[k + 1] - g->vtx_off[k];
            const double *v = g->vtx + 2 * g->vtx_off[k];
            for (uint64_t i = 0; i < V; ++i) {
                const uint64_t i1 = i + 1 < V ? i + 1 : 0;
                const double ya = v[2 * i + 1], yb = v[2 * i1 + 1];
                if (ya == yb) continue;
                const double lo = ya < yb ? ya : yb, hi = ya < yb ? yb : ya;
                double r0 = floor(lo), r1 = ceil(hi) - 1.0;
                if (r0 < 0.0) r0 = 0.0;
                if (r1 > (double)H - 1.0) r1 = (double)H - 1.0;
                for (double r = r0; r <= r1; r += 1.0) {
                    const uint64_t c = (uint64_t)k * H + (uint64_t)r;
                    if (pass == 0)
                        g->row_off[c + 1]++;
                    else
                        g->row_edge[cur[c]++] = (uint32_t)i;
                }
            }
        }
        free(cur);
    }
    free(poff);
}

static void geom_free(geom *g) {
    free(g->vtx_off);
    free(g->vtx);
    free(g->row_off);
    free(g->row_edge);
}

/* A of pixel (X, Y) by path k; with `gv` non-NULL also adds dA * dA/dvertex
 * into gv (the path's vertex gradient, (x, y) pairs). */
static double pixel_area(const fd_scene *s, const geom *g, uint32_t k, int X, int Y, double dA, double *gv) {
    const uint64_t c = (uint64_t)k * s->height + (uint64_t)Y;
    const uint64_t V = g->vtx_off[k + 1] - g->vtx_off[k];
    const double *v = g->vtx + 2 * g->vtx_off[k];
    double A = 0.0;
    for (uint64_t q = g->row_off[c]; q < g->row_off[c + 1]; ++q) {
        const uint64_t i = g->row_edge[q], i1 = i + 1 < V ? i + 1 : 0;
        if (!gv) {
            A += edge_area_v(v[2 * i], v[2 * i + 1], v[2 * i1], v[2 * i1 + 1], X, Y);
            continue;
        }
        double df[4];
        A += edge_area(v[2 * i], v[2 * i + 1], v[2 * i1], v[2 * i1 + 1], X, Y, df);
        {
            gv[2 * i] += dA * df[0];
            gv[2 * i + 1] += dA * df[1];
            gv[2 * i1] += dA * df[2];
            gv[2 * i1 + 1] += dA * df[3];
        }
    }
    return A;
}

double fd_pixel_area(const fd_scene *s, const double *points, uint32_t path, uint32_t x, uint32_t y) {
    geom g;
    geom_build(s, points, &g);
    const double A = pixel_area(s, &g, path, (int)x, (int)y, 0.0, NULL);
    geom_free(&g);
    return A;
}

static double alpha_of(double A, double *dalpha_dA) {
    const double m = fabs(A);
    if (m < 1.0) {
        *dalpha_dA = A > 0.0 ? 1.0 : (A < 0.0 ? -1.0 : 0.0);
        return m;
    }
    *dalpha_dA = 0.0;
    return 1.0;
}

/* First path of each image ([n_images + 1]). */
static uint32_t *image_paths(const fd_scene *s) {
    uint32_t *f = (uint32_t *)malloc(sizeof(uint32_t) * (s->n_images + 1));
    uint32_t k = 0;
    for (uint32_t i = 0; i <= s->n_images; ++i) {
        while (k < s->n_paths && s->path_image[k] < i) ++k;
        f[i] = k;
    }
    return f;
}

double fd_forward(const fd_scene *s, const double *points, const double *colours, const double *target,
                  double *canvas) {
    geom g;
    geom_build(s, points, &g);
    uint32_t *first = image_paths(s);
    const uint32_t W = s->width, H = s->height;
    const double N = 4.0 * s->n_images * (double)W * H;
    double loss = 0.0;
    for (uint32_t im = 0; im < s->n_images; ++im)
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                double C[4] = {0, 0, 0, 0};
                for (uint32_t k = first[im]; k < first[im + 1]; ++k) {
                    double da;
                    const double al = alpha_of(pixel_area(s, &g, k, (int)x, (int)y, 0.0, NULL), &da);
                    const double *c = colours + 4 * (uint64_t)k;
                    const double keep = 1.0 - c[3] * al;
                    for (int ch = 0; ch < 4; ++ch) C[ch] = c[ch] * al + C[ch] * keep;
                }
                const uint64_t px = (((uint64_t)im * H + y) * W + x) * 4;
                for (int ch = 0; ch < 4; ++ch) {
                    if (canvas) canvas[px + ch] = C[ch];
                    if (target) loss += (C[ch] - target[px + ch]) * (C[ch] - target[px + ch]);
                }
            }
    free(first);
    geom_free(&g);
    return target ? loss / N : 0.0;
}

double fd_backward(const fd_scene *s, const double *points, const double *colours, const double *target,
                   const double *grad_canvas, double *grad_points, double *grad_colours) {
    geom g;
    geom_build(s, points, &g);
    uint32_t *first = image_paths(s);
    uint64_t *poff = point_offsets(s);
    const uint32_t W = s->width, H = s->height;
    const double N = 4.0 * s->n_images * (double)W * H;
    double *gv = (double *)calloc(2 * (g.vtx_off[s->n_paths] + 1), sizeof(double));
    memset(grad_colours, 0, sizeof(double) * 4 * s->n_paths);
    uint32_t max_paths = 0;
    for (uint32_t im = 0; im < s->n_images; ++im)
        if (first[im + 1] - first[im] > max_paths) max_paths = first[im + 1] - first[im];
    double *A = (double *)malloc(sizeof(double) * (max_paths + 1));
    double *Cb = (double *)malloc(sizeof(double) * 4 * (max_paths + 1)); /* canvas before each path */
    double loss = 0.0;
    for (uint32_t im = 0; im < s->n_images; ++im)
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                const uint32_t k0 = first[im], np = first[im + 1] - k0;
                double C[4] = {0, 0, 0, 0};
                for (uint32_t j = 0; j < np; ++j) {
                    const uint32_t k = k0 + j;
                    A[j] = pixel_area(s, &g, k, (int)x, (int)y, 0.0, NULL);
                    double da;
                    const double al = alpha_of(A[j], &da);
                    const double *c = colours + 4 * (uint64_t)k;
                    const double keep = 1.0 - c[3] * al;
                    for (int ch = 0; ch < 4; ++ch) {
                        Cb[4 * j + ch] = C[ch];
                        C[ch] = c[ch] * al + C[ch] * keep;
                    }
                }
                const uint64_t px = (((uint64_t)im * H + y) * W + x) * 4;
                double gC[4];
                for (int ch = 0; ch < 4; ++ch) {
                    if (grad_canvas) {
                        gC[ch] = grad_canvas[px + ch];
                    } else {
                        const double r = C[ch] - (target ? target[px + ch] : 0.0);
                        gC[ch] = 2.0 * r / N;
                        loss += r * r;
                    }
                }
                for (uint32_t jj = np; jj-- > 0;) {
                    const uint32_t k = k0 + jj;
                    double da;
                    const double al = alpha_of(A[jj], &da);
                    const double *c = colours + 4 * (uint64_t)k;
                    double *gc = grad_colours + 4 * (uint64_t)k;
                    double dal = 0.0, gdotC = 0.0;
                    for (int ch = 0; ch < 4; ++ch) {
                        dal += gC[ch] * (c[ch] - c[3] * Cb[4 * jj + ch]);
                        gc[ch] += gC[ch] * al;
                        gdotC += gC[ch] * Cb[4 * jj + ch];
                    }
                    gc[3] -= al * gdotC;
                    const double keep = 1.0 - c[3] * al;
                    for (int ch = 0; ch < 4; ++ch) gC[ch] *= keep;
                    const double dA = dal * da;
                    if (dA != 0.0) pixel_area(s, &g, k, (int)x, (int)y, dA, gv + 2 * g.vtx_off[k]);
                }
            }
    /* vertex gradients -> control points */
    memset(grad_points, 0, sizeof(double) * 2 * poff[s->n_paths]);
    for (uint32_t k = 0; k < s->n_paths; ++k) {
        const double *v = gv + 2 * g.vtx_off[k];
        double *gp = grad_points + 2 * poff[k];
        gp[0] += v[0];
        gp[1] += v[1];
        for (uint32_t c = 0; c < s->path_cubics[k]; ++c)
            for (int j = 1; j <= FD_SEG; ++j) {
                double w[4];
                bernstein(j, w);
                const uint64_t vi = 1 + (uint64_t)c * FD_SEG + (uint64_t)(j - 1);
                for (int m = 0; m < 4; ++m) {
                    gp[2 * (3 * c + m)] += w[m] * v[2 * vi];
                    gp[2 * (3 * c + m) + 1] += w[m] * v[2 * vi + 1];
                }
            }
    }
    free(A);
    free(Cb);
    free(gv);
    free(poff);
    free(first);
    geom_free(&g);
    return grad_canvas ? 0.0 : loss / N;
}
