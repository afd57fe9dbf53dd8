// ft_diff.hip -- the differentiable fill on gfx950: float32 analytic-coverage
// forward and a hand-written adjoint.  footile has no counterpart; the model
// (fixed dyadic flattening of cubics, winding-integral pixel coverage, SrcOver
// in path order) is specified by the FP64 oracle the tests compare against.
//
// Per step:
//   k_vertices   control points -> polygon vertices (16 per cubic, Bernstein)
//   k_rows       each edge adds +1/-1 at the ends of its row range per path
//   scan, scan   -> edges per (path, row) cell, CSR offsets
//   k_fill_cells, k_sort_cells   edge lists per cell, in edge order
//   k_forward    one lane per pixel: coverage of every path of the image from
//                its cell's edges, composite, write the canvas
//   k_loss, k_loss_sum   L2 against a target: d loss / d canvas and the loss,
//                summed on the device in a fixed order
//   k_backward   one lane per pixel: replays the forward into canvas
//                checkpoints every KSEG paths (inside ft_diff_step k_forward
//                has written them), walks each segment's paths backwards with
//                their coverage recomputed into LDS (or read from the kept
//                coverage buffer, FT_DIFF_KEEP_COVERAGE), reduces the per-edge
//                and per-colour gradients over the wavefront and gathers them
//                per block in LDS before one atomic add per touched vertex and
//                colour channel
//   k_point_grad vertex gradients -> control point gradients
#include <algorithm>
#include <cstring>
#include <vector>

#include "ft_scan.h"

namespace ft {
namespace diff {

constexpr int DSEG = 16;          // segments per cubic
constexpr float EPS_W = 1e-5f;    // narrower pieces use the midpoint value
constexpr int DBLOCK = 128;       // pixels of one row per block
constexpr int KSEG = 8;           // paths per checkpoint segment
constexpr uint32_t VCACHE = 256;  // vertex gradients a backward block gathers in LDS per path
constexpr uint64_t COV_MAX_BYTES = 16ull << 30;  // largest kept-coverage buffer

struct DScene {
    uint32_t n_images, W, H, n_paths;
    uint32_t n_vtx, n_cells;
    const uint32_t *pt_off;     // [n_paths + 1] first point of each path
    const uint32_t *vtx_off;    // [n_paths + 1] first vertex of each path
    const uint32_t *vtx_path;   // [n_vtx]
    const uint32_t *img_first;  // [n_images + 1] first path of each image
    float2 *vtx;                // [n_vtx]
    uint32_t *cell_cnt;         // [n_cells] edges per (path, row); first the differences
    uint32_t *cell_off;         // [n_cells]
    uint32_t *cell_cur;         // [n_cells]
    uint32_t *cell_edge;        // [cap_pairs] local edge index
    uint32_t cap_pairs;
};

// Bernstein weights at t = j / 16: integers over 4096, exact in f32.
__device__ inline void bern(int j, float w[4]) {
    const float u = (float)(DSEG - j), t = (float)j;
    w[0] = u * u * u / 4096.0f;
    w[1] = 3.0f * u * u * t / 4096.0f;
    w[2] = 3.0f * u * t * t / 4096.0f;
    w[3] = t * t * t / 4096.0f;
}

__device__ inline float clamp01(float u) { return u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u); }

// Mean of clamp(u, 0, 1) for u moving linearly from ua to ub, and its partials.
__device__ inline float mean_clamp(float ua, float ub, float &d_ua, float &d_ub) {
    const float lo = ua < ub ? ua : ub, hi = ua < ub ? ub : ua;
    float m, dlo = 0.0f, dhi = 0.0f;
    if (hi <= 0.0f) {
        m = 0.0f;
    } else if (lo >= 1.0f) {
        m = 1.0f;
    } else {
        const float w = hi - lo;
        if (w < EPS_W) {
            const float mid = (lo + hi) * 0.5f;
            if (mid <= 0.0f) {
                m = 0.0f;
            } else if (mid >= 1.0f) {
                m = 1.0f;
            } else {
                m = mid;
                dlo = dhi = 0.5f;
            }
        } else {
            const float a = clamp01(lo), b = clamp01(hi);
            const float integral = (b - a) * (a + b) * 0.5f + (hi > 1.0f ? hi - 1.0f : 0.0f);
            m = integral / w;
            dhi = (b - m) / w;
            dlo = (m - a) / w;
        }
    }
    d_ua = ua <= ub ? dlo : dhi;
    d_ub = ua <= ub ? dhi : dlo;
    return m;
}

// The row piece of an edge: its part within row Y, oriented top to bottom.
struct Piece {
    bool live;
    bool down;       // the edge descends from its first to its second vertex
    float sign;      // +1 descending, -1 ascending
    float xa, xb;    // x at the piece's top and bottom
    float h;         // piece height
};

__device__ inline Piece piece_of(float2 p, float2 q, int Y) {
    Piece r;
    r.live = false;
    r.down = q.y > p.y;
    r.sign = r.down ? 1.0f : -1.0f;
    r.xa = r.xb = r.h = 0.0f;
    if (p.y == q.y) return r;
    const float2 t = r.down ? p : q, b = r.down ? q : p;
    const float ya = t.y > (float)Y ? t.y : (float)Y;
    const float yb = b.y < (float)(Y + 1) ? b.y : (float)(Y + 1);
    if (!(yb > ya)) return r;
    const float sl = (b.x - t.x) / (b.y - t.y);
    r.xa = t.x + (ya - t.y) * sl;
    r.xb = t.x + (yb - t.y) * sl;
    r.h = yb - ya;
    r.live = true;
    return r;
}

// d(sign * h * M) / d(x_top, y_top, x_bot, y_bot) given S1 = sum dA M,
// S2 = sum dA dM/dua, S3 = sum dA dM/dub over the pixels of the row.
__device__ inline void piece_grad(float2 p, float2 q, int Y, float S1, float S2, float S3, float g[4]) {
    const bool down = q.y > p.y;
    const float2 t = down ? p : q, b = down ? q : p;
    const float sign = down ? 1.0f : -1.0f;
    const bool clip_top = !(t.y > (float)Y), clip_bot = !(b.y < (float)(Y + 1));
    const float ya = clip_top ? (float)Y : t.y, yb = clip_bot ? (float)(Y + 1) : b.y;
    const float dy = b.y - t.y, dx = b.x - t.x;
    const float sl = dx / dy;
    // partials of ya, yb, sl w.r.t. (xt, yt, xb, yb_end)
    const float dya[4] = {0.0f, clip_top ? 0.0f : 1.0f, 0.0f, 0.0f};
    const float dyb[4] = {0.0f, 0.0f, 0.0f, clip_bot ? 0.0f : 1.0f};
    const float dsl[4] = {-1.0f / dy, sl / dy, 1.0f / dy, -sl / dy};
    const float ra = ya - t.y, rb = yb - t.y;
    float gd[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float dt = i == 1 ? 1.0f : 0.0f, dxt = i == 0 ? 1.0f : 0.0f;
        const float dra = dya[i] - dt, drb = dyb[i] - dt;
        const float dxa = dxt + dra * sl + ra * dsl[i];
        const float dxb = dxt + drb * sl + rb * dsl[i];
        const float dh = dyb[i] - dya[i];
        gd[i] = sign * (dh * S1 - (yb - ya) * (dxa * S2 + dxb * S3));
    }
    // back to (x_p, y_p, x_q, y_q)
    g[0] = down ? gd[0] : gd[2];
    g[1] = down ? gd[1] : gd[3];
    g[2] = down ? gd[2] : gd[0];
    g[3] = down ? gd[3] : gd[1];
}

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Coverage A of pixel X (row Y) by path k.
__device__ inline float path_area(const DScene &S, uint32_t k, uint32_t off, uint32_t cnt, int X, int Y,
                                  uint32_t &bad) {
    const uint32_t vb = S.vtx_off[k], V = S.vtx_off[k + 1] - vb;
    float A = 0.0f;
    for (uint32_t q = 0; q < cnt; ++q) {
        const uint32_t e = S.cell_edge[off + q];
        bad |= e < V ? 0u : 1u;
        const uint32_t i = e < V ? e : 0u;
        const uint32_t i1 = i + 1 < V ? i + 1 : 0;
        const Piece r = piece_of(S.vtx[vb + i], S.vtx[vb + i1], Y);
        if (!r.live) continue;
        float da, db;
        const float M = mean_clamp((float)X + 1.0f - r.xa, (float)X + 1.0f - r.xb, da, db);
        A += r.sign * r.h * M;
    }
    return A;
}

__device__ inline float alpha_of(float A, float &dadA) {
    const float m = fabsf(A);
    if (m < 1.0f) {
        dadA = A > 0.0f ? 1.0f : (A < 0.0f ? -1.0f : 0.0f);
        return m;
    }
    dadA = 0.0f;
    return 1.0f;
}

__device__ inline float4 over(float4 c, float al, float4 C) {
    const float keep = 1.0f - c.w * al;
    return make_float4(c.x * al + C.x * keep, c.y * al + C.y * keep, c.z * al + C.z * keep, c.w * al + C.w * keep);
}

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_vertices(DScene S, const float2 *points, DevStatus *st) {
    for (uint32_t v = blockIdx.x * BLOCK + threadIdx.x; v < S.n_vtx; v += gridDim.x * BLOCK) {
        const uint32_t k = S.vtx_path[v];
        if (k >= S.n_paths) {
            raise_err(st, FT_EDEVICE);
            continue;
        }
        const uint32_t local = v - S.vtx_off[k];
        const float2 *P = points + S.pt_off[k];
        float2 r = P[0];
        if (local) {
            const uint32_t c = (local - 1) / DSEG;
            const int j = (int)((local - 1) % DSEG) + 1;
            float w[4];
            bern(j, w);
            const float2 p0 = P[3 * c], p1 = P[3 * c + 1], p2 = P[3 * c + 2], p3 = P[3 * c + 3];
            r.x = w[0] * p0.x + w[1] * p1.x + w[2] * p2.x + w[3] * p3.x;
            r.y = w[0] * p0.y + w[1] * p1.y + w[2] * p2.y + w[3] * p3.y;
        }
        S.vtx[v] = r;
    }
}

// Rows [r0, r1] an edge overlaps within [0, H); false when none.
__device__ inline bool edge_row_range(float2 a, float2 b, uint32_t H, uint32_t &r0, uint32_t &r1) {
    if (!(a.y != b.y)) return false;  // horizontal, or NaN
    float lo = a.y < b.y ? a.y : b.y, hi = a.y < b.y ? b.y : a.y;
    if (!(hi > 0.0f) || !(lo < (float)H)) return false;
    lo = lo < 0.0f ? 0.0f : lo;
    hi = hi > (float)H ? (float)H : hi;
    const float f0 = floorf(lo), f1 = ceilf(hi) - 1.0f;
    if (f1 < f0) return false;
    r0 = (uint32_t)f0;
    r1 = (uint32_t)f1 < H - 1 ? (uint32_t)f1 : H - 1;
    return r0 <= r1;
}

__device__ inline void edge_at(const DScene &S, uint32_t v, uint32_t &k, float2 &a, float2 &b) {
    k = S.vtx_path[v];
    const uint32_t vb = S.vtx_off[k], V = S.vtx_off[k + 1] - vb, local = v - vb;
    a = S.vtx[v];
    b = S.vtx[local + 1 < V ? v + 1 : vb];
}

// Clears the cell counters, the scans' look-back words, the status and (when
// given) the gradients.
__global__ __launch_bounds__(BLOCK) void k_clear_cells(DScene S, float2 *gvtx, float *gcol, unsigned long long *lb,
                                                       uint32_t n_lb, DevStatus *st) {
    uint32_t n = S.n_cells > S.n_vtx ? S.n_cells : S.n_vtx;
    n = n > 4 * S.n_paths ? n : 4 * S.n_paths;
    n = n > n_lb ? n : n_lb;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        if (i < n_lb) lb[i] = 0;
        if (i < S.n_cells) {
            S.cell_cnt[i] = 0;
            S.cell_cur[i] = 0;
        }
        if (gvtx && i < S.n_vtx) gvtx[i] = make_float2(0.0f, 0.0f);
        if (gcol && i < 4 * S.n_paths) gcol[i] = 0.0f;
    }
    if (blockIdx.x == 0 && threadIdx.x < sizeof(DevStatus) / 4) reinterpret_cast<uint32_t *>(st)[threadIdx.x] = 0;
}

__global__ __launch_bounds__(BLOCK) void k_rows(DScene S, DevStatus *st) {
    for (uint32_t v = blockIdx.x * BLOCK + threadIdx.x; v < S.n_vtx; v += gridDim.x * BLOCK) {
        uint32_t k, r0, r1;
        float2 a, b;
        edge_at(S, v, k, a, b);
        if (!edge_row_range(a, b, S.H, r0, r1)) continue;
        // +1 at the first row's cell, -1 at the cell after the last row (the
        // next path's first row when r1 is the last row): the inclusive scan
        // over all cells is then the edge count of every (path, row)
        const uint64_t c = (uint64_t)k * S.H;
        atomicAdd(&S.cell_cnt[c + r0], 1u);
        if (c + r1 + 1 < S.n_cells) atomicAdd(&S.cell_cnt[c + r1 + 1], 0xFFFFFFFFu);
    }
}

__global__ __launch_bounds__(BLOCK) void k_fill_cells(DScene S, DevStatus *st) {
    if (run_failed(st)) return;
    if (st->n_pairs > S.cap_pairs) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(&st->overflow, 2u);
            st->need_pairs = st->n_pairs;
        }
        return;
    }
    for (uint32_t v = blockIdx.x * BLOCK + threadIdx.x; v < S.n_vtx; v += gridDim.x * BLOCK) {
        uint32_t k, r0, r1;
        float2 a, b;
        edge_at(S, v, k, a, b);
        if (!edge_row_range(a, b, S.H, r0, r1)) continue;
        const uint32_t local = v - S.vtx_off[k];
        for (uint32_t r = r0; r <= r1; ++r) {
            const uint64_t c = (uint64_t)k * S.H + r;
            const uint32_t at = S.cell_off[c] + atomicAdd(&S.cell_cur[c], 1u);
            if (at < S.cap_pairs)
                S.cell_edge[at] = local;
            else
                raise_err(st, FT_EDEVICE);
        }
    }
}

// Each cell's edges in ascending order, so that sums run in edge order.
__global__ __launch_bounds__(BLOCK) void k_sort_cells(DScene S, DevStatus *st) {
    if (run_failed(st)) return;
    for (uint32_t c = blockIdx.x * BLOCK + threadIdx.x; c < S.n_cells; c += gridDim.x * BLOCK) {
        const uint32_t off = S.cell_off[c], n = S.cell_cnt[c];
        if ((uint64_t)off + n > S.cap_pairs) {
            raise_err(st, FT_EDEVICE);
            continue;
        }
        uint32_t *e = S.cell_edge + off;
        for (uint32_t i = 1; i < n; ++i) {
            const uint32_t x = e[i];
            uint32_t j = i;
            while (j > 0 && e[j - 1] > x) {
                e[j] = e[j - 1];
                --j;
            }
            e[j] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// Forward, loss, backward
// ---------------------------------------------------------------------------
// With ck non-null it also leaves what k_backward's replay would write: the
// canvas before every KSEG-th path, and each path's coverage when cov is
// non-null (same slots as k_backward).
__global__ __launch_bounds__(DBLOCK) void k_forward(DScene S, const float4 *colours, float4 *canvas, float4 *ck,
                                                    uint64_t ck_stride, float *cov, DevStatus *st) {
    if (run_failed(st)) return;
    const int X = (int)(blockIdx.x * DBLOCK + threadIdx.x), Y = (int)blockIdx.y;
    const uint32_t im = blockIdx.z;
    const uint32_t k0 = S.img_first[im], k1 = S.img_first[im + 1];
    const uint64_t slot = ((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * (gridDim.x * DBLOCK) +
                          (uint64_t)blockIdx.x * DBLOCK + threadIdx.x;
    float4 C = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t bad = 0;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t j = k - k0;
        if (ck && j % KSEG == 0) ck[(uint64_t)(j / KSEG) * ck_stride + slot] = C;
        const uint64_t c = (uint64_t)k * S.H + (uint32_t)Y;
        const uint32_t cnt = S.cell_cnt[c];
        if (!cnt) continue;
        float dadA;
        const float A = path_area(S, k, S.cell_off[c], cnt, X, Y, bad);
        if (ck && cov) cov[(uint64_t)j * ck_stride + slot] = A;  // j < max_paths, slot < ck_stride
        const float al = alpha_of(A, dadA);
        C = over(colours[k], al, C);
    }
    if (bad) raise_err(st, FT_EDEVICE);
    if ((uint32_t)X < S.W) canvas[((uint64_t)im * S.H + (uint32_t)Y) * S.W + (uint32_t)X] = C;
}

// loss partials per block (sum of squared differences); grad = 2 (C - T) / N.
__global__ __launch_bounds__(BLOCK) void k_loss(const float *canvas, const float *target, float *grad, uint64_t n,
                                                double *partial) {
    __shared__ double s_w[BLOCK / 64];
    double acc = 0.0;
    const float scale = (float)(2.0 / (double)n);
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
        const float r = canvas[i] - target[i];
        acc += (double)r * (double)r;
        if (grad) grad[i] = scale * r;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) t += s_w[w];
        partial[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(DBLOCK) void k_backward(DScene S, const float4 *colours, const float4 *grad_canvas,
                                                     float4 *ck, uint64_t ck_stride, float *cov, uint32_t replayed,
                                                     float2 *gvtx, float *gcol,
                                                     DevStatus *st) {
    __shared__ float s_A[KSEG][DBLOCK];  // coverage of the segment's paths; each thread uses its own column
    __shared__ float2 s_gv[VCACHE];      // the current path's vertex gradients from this block's pixels
    __shared__ float s_gc[4];            // and its colour gradient
    if (run_failed(st)) return;
    const int tid = threadIdx.x, lane = tid & 63;
    for (uint32_t v = tid; v < VCACHE; v += DBLOCK) s_gv[v] = make_float2(0.0f, 0.0f);
    if (tid < 4) s_gc[tid] = 0.0f;
    __syncthreads();
    const int X = (int)(blockIdx.x * DBLOCK + tid), Y = (int)blockIdx.y;
    const uint32_t im = blockIdx.z;
    const bool active = (uint32_t)X < S.W;
    const uint64_t slot = ((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * (gridDim.x * DBLOCK) +
                          (uint64_t)blockIdx.x * DBLOCK + tid;
    const uint32_t k0 = S.img_first[im], np = S.img_first[im + 1] - k0;
    // pass 1: forward, checkpointing the canvas before every KSEG-th path
    // (skipped when k_forward of the same inputs has written both: `replayed`)
    float4 C = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t bad = 0;
    for (uint32_t j = 0; !replayed && j < np; ++j) {
        if (j % KSEG == 0) ck[(uint64_t)(j / KSEG) * ck_stride + slot] = C;
        const uint32_t k = k0 + j;
        const uint64_t c = (uint64_t)k * S.H + (uint32_t)Y;
        const uint32_t cnt = S.cell_cnt[c];
        if (!cnt) continue;
        float dadA;
        const float A = path_area(S, k, S.cell_off[c], cnt, X, Y, bad);
        if (cov) cov[(uint64_t)j * ck_stride + slot] = A;  // j < max_paths, slot < ck_stride
        const float al = alpha_of(A, dadA);
        C = over(colours[k], al, C);
    }
    float4 g = active ? grad_canvas[((uint64_t)im * S.H + (uint32_t)Y) * S.W + (uint32_t)X]
                      : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // pass 2: segments last to first
    const uint32_t n_seg = (np + KSEG - 1) / KSEG;
    for (uint32_t seg = n_seg; seg-- > 0;) {
        const uint32_t j0 = seg * KSEG, j1 = j0 + KSEG < np ? j0 + KSEG : np;
        for (uint32_t j = j0; j < j1; ++j) {
            const uint32_t k = k0 + j;
            const uint64_t c = (uint64_t)k * S.H + (uint32_t)Y;
            const uint32_t cnt = S.cell_cnt[c];
            s_A[j - j0][tid] = !cnt ? 0.0f
                               : cov  ? cov[(uint64_t)j * ck_stride + slot]
                                      : path_area(S, k, S.cell_off[c], cnt, X, Y, bad);
        }
        const float4 Cseg = ck[(uint64_t)seg * ck_stride + slot];
        for (uint32_t j = j1; j-- > j0;) {
            const uint32_t k = k0 + j;
            float4 Cb = Cseg;  // canvas before path j
            for (uint32_t i = j0; i < j; ++i) {
                float d;
                Cb = over(colours[k0 + i], alpha_of(s_A[i - j0][tid], d), Cb);
            }
            const float A = s_A[j - j0][tid];
            float dadA;
            const float al = alpha_of(A, dadA);
            const float4 cc = colours[k];
            const float dal = g.x * (cc.x - cc.w * Cb.x) + g.y * (cc.y - cc.w * Cb.y) + g.z * (cc.z - cc.w * Cb.z) +
                              g.w * (cc.w - cc.w * Cb.w);
            const float gdotC = g.x * Cb.x + g.y * Cb.y + g.z * Cb.z + g.w * Cb.w;
            if (__ballot(al != 0.0f)) {
                const float r0 = wave_sum(g.x * al), r1 = wave_sum(g.y * al), r2 = wave_sum(g.z * al);
                const float r3 = wave_sum(g.w * al - al * gdotC);
                if (lane == 0) {  // al != 0 needs edges in this row: flushed below
                    atomicAdd(&s_gc[0], r0);
                    atomicAdd(&s_gc[1], r1);
                    atomicAdd(&s_gc[2], r2);
                    atomicAdd(&s_gc[3], r3);
                }
            }
            const float keep = 1.0f - cc.w * al;
            g = make_float4(g.x * keep, g.y * keep, g.z * keep, g.w * keep);
            const float dA = active ? dal * dadA : 0.0f;
            // k and Y are the same for the whole block, so cnt, V and
            // `gather` are too: the barriers below are reached by every thread
            const uint64_t c = (uint64_t)k * S.H + (uint32_t)Y;
            const uint32_t cnt = S.cell_cnt[c], off = S.cell_off[c];
            const uint32_t vb = S.vtx_off[k], V = S.vtx_off[k + 1] - vb;
            const bool gather = V <= VCACHE;
            if (__ballot(dA != 0.0f)) {
                for (uint32_t q = 0; q < cnt; ++q) {
                    const uint32_t e = S.cell_edge[off + q];
                    bad |= e < V ? 0u : 1u;
                    const uint32_t i = e < V ? e : 0u;
                    const uint32_t i1 = i + 1 < V ? i + 1 : 0;
                    const float2 p = S.vtx[vb + i], pq = S.vtx[vb + i1];
                    const Piece r = piece_of(p, pq, Y);
                    if (!r.live) continue;
                    float mua, mub;
                    const float M = mean_clamp((float)X + 1.0f - r.xa, (float)X + 1.0f - r.xb, mua, mub);
                    const float S1 = wave_sum(dA * M), S2 = wave_sum(dA * mua), S3 = wave_sum(dA * mub);
                    if (lane == 0) {
                        float gr[4];
                        piece_grad(p, pq, Y, S1, S2, S3, gr);
                        if (gather) {  // i, i1 < V <= VCACHE
                            atomicAdd(&s_gv[i].x, gr[0]);
                            atomicAdd(&s_gv[i].y, gr[1]);
                            atomicAdd(&s_gv[i1].x, gr[2]);
                            atomicAdd(&s_gv[i1].y, gr[3]);
                        } else {
                            atomicAdd(&gvtx[vb + i].x, gr[0]);
                            atomicAdd(&gvtx[vb + i].y, gr[1]);
                            atomicAdd(&gvtx[vb + i1].x, gr[2]);
                            atomicAdd(&gvtx[vb + i1].y, gr[3]);
                        }
                    }
                }
            }
            if (cnt != 0) {  // one global atomic per colour channel and touched vertex, per block
                __syncthreads();
                if (tid < 4) {
                    const float a = s_gc[tid];
                    if (a != 0.0f) {
                        atomicAdd(&gcol[4 * (uint64_t)k + tid], a);
                        s_gc[tid] = 0.0f;
                    }
                }
                for (uint32_t v = tid; gather && v < V; v += DBLOCK) {
                    const float2 a = s_gv[v];
                    if ((a.x != 0.0f) | (a.y != 0.0f)) {
                        atomicAdd(&gvtx[vb + v].x, a.x);
                        atomicAdd(&gvtx[vb + v].y, a.y);
                        s_gv[v] = make_float2(0.0f, 0.0f);
                    }
                }
                __syncthreads();
            }
        }
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// The loss partials summed in a fixed order by one block: the same value for
// a step and for a separate ft_diff_loss of the same canvas.
__global__ __launch_bounds__(BLOCK) void k_loss_sum(const double *partial, uint32_t n, double *out) {
    __shared__ double s_w[BLOCK / 64];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) acc += partial[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) t += s_w[w];
        out[0] = t;
    }
}

// Control point gradients from vertex gradients (one thread per point).
__global__ __launch_bounds__(BLOCK) void k_point_grad(DScene S, const float2 *gvtx, float2 *gpts, uint32_t n_pts,
                                                      const uint32_t *pt_path) {
    for (uint32_t m = blockIdx.x * BLOCK + threadIdx.x; m < n_pts; m += gridDim.x * BLOCK) {
        const uint32_t k = pt_path[m];
        const uint32_t local = m - S.pt_off[k];
        const uint32_t n_cub = (S.pt_off[k + 1] - S.pt_off[k] - 1) / 3;
        const uint32_t vb = S.vtx_off[k];
        float2 acc = make_float2(0.0f, 0.0f);
        if (local == 0) acc = gvtx[vb];
        // the point is control r of cubic c (and control 0 of cubic c + 1 when r == 3)
        for (int pass = 0; pass < 2; ++pass) {
            uint32_t c, r;
            if (pass == 0) {
                if (local == 0) continue;
                c = (local - 1) / 3;
                r = (local - 1) % 3 + 1;
            } else {
                if (local % 3 != 0) continue;
                c = local / 3;
                r = 0;
                if (c >= n_cub) continue;
            }
            for (int j = 1; j <= DSEG; ++j) {
                float w[4];
                bern(j, w);
                const float2 gv = gvtx[vb + 1 + c * DSEG + (uint32_t)(j - 1)];
                acc.x += w[r] * gv.x;
                acc.y += w[r] * gv.y;
            }
        }
        gpts[m] = acc;
    }
}

}  // namespace diff
}  // namespace ft

// ---------------------------------------------------------------------------
// Host side and C ABI
// ---------------------------------------------------------------------------
using namespace ft;
using namespace ft::diff;

struct ft_diff {
    hipStream_t stream = nullptr;
    uint32_t n_images = 0, W = 0, H = 0, n_paths = 0, max_paths = 0;
    uint32_t n_pts = 0, n_vtx = 0;
    uint64_t n_cells = 0, canvas_floats = 0;
    uint32_t *d_pt_off = nullptr, *d_vtx_off = nullptr, *d_vtx_path = nullptr, *d_img_first = nullptr,
             *d_pt_path = nullptr;
    float2 *d_vtx = nullptr, *d_gvtx = nullptr;
    uint32_t *d_cell_cnt = nullptr, *d_cell_off = nullptr, *d_cell_cur = nullptr, *d_cell_edge = nullptr;
    uint32_t cap_pairs = 0;
    size_t cap_tiles = 0;
    unsigned long long *d_lb = nullptr;  // scan look-back words
    float4 *d_ck = nullptr;
    uint64_t ck_stride = 0;
    float *d_cov = nullptr;  // [max_paths][ck_stride] kept coverage (FT_DIFF_KEEP_COVERAGE), or null: recomputed
    double *d_partial = nullptr;  // [LOSS_GRID] loss partials, then [LOSS_GRID] their sum
    double *h_loss = nullptr;     // pinned: the summed loss
    DevStatus *d_status = nullptr, *h_status = nullptr;
};

namespace {

constexpr uint32_t LOSS_GRID = 1024;

#define FD_HIP(call)                                 \
    do {                                             \
        if ((call) != hipSuccess) return FT_EDEVICE; \
    } while (0)

template <typename T>
ft_status dalloc(T *&p, size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc(reinterpret_cast<void **>(&p), (n ? n : 1) * sizeof(T)) == hipSuccess ? FT_OK : FT_ENOMEM;
}

DScene scene_of(const ft_diff *d) {
    DScene S;
    S.n_images = d->n_images;
    S.W = d->W;
    S.H = d->H;
    S.n_paths = d->n_paths;
    S.n_vtx = d->n_vtx;
    S.n_cells = (uint32_t)d->n_cells;
    S.pt_off = d->d_pt_off;
    S.vtx_off = d->d_vtx_off;
    S.vtx_path = d->d_vtx_path;
    S.img_first = d->d_img_first;
    S.vtx = d->d_vtx;
    S.cell_cnt = d->d_cell_cnt;
    S.cell_off = d->d_cell_off;
    S.cell_cur = d->d_cell_cur;
    S.cell_edge = d->d_cell_edge;
    S.cap_pairs = d->cap_pairs;
    return S;
}

uint32_t grid_of(uint64_t n) {
    uint64_t g = (n + BLOCK - 1) / BLOCK;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, GRID_MAX));
}

// Geometry for one parameter set: vertices, per-(path, row) edge lists.
ft_status launch_geometry(ft_diff *d, const float *points, bool clear_grads, float *gcol) {
    hipStream_t s = d->stream;
    DScene S = scene_of(d);
    const uint32_t n_lb = (uint32_t)(4 * d->cap_tiles);  // look-back words of the two cell scans
    const uint64_t n_clear = std::max<uint64_t>(std::max<uint64_t>(d->n_cells, d->n_vtx),
                                                std::max<uint64_t>(4ull * d->n_paths, n_lb));
    k_clear_cells<<<grid_of(n_clear), BLOCK, 0, s>>>(S, clear_grads ? d->d_gvtx : nullptr,
                                                     clear_grads ? gcol : nullptr, d->d_lb, n_lb, d->d_status);
    k_vertices<<<grid_of(d->n_vtx), BLOCK, 0, s>>>(S, reinterpret_cast<const float2 *>(points), d->d_status);
    k_rows<<<grid_of(d->n_vtx), BLOCK, 0, s>>>(S, d->d_status);
    const uint32_t nc = (uint32_t)d->n_cells;
    scan_launch<uint32_t, false>(s, scan_resident_blocks(), d->d_cell_cnt, d->d_cell_cnt, nullptr, nc, nc, d->d_lb,
                                 nullptr, d->d_status, false, true);
    scan_launch<uint32_t, false>(s, scan_resident_blocks(), d->d_cell_cnt, d->d_cell_off, nullptr, nc, nc,
                                 d->d_lb + 2 * d->cap_tiles, &d->d_status->n_pairs, d->d_status, false, false);
    k_fill_cells<<<grid_of(d->n_vtx), BLOCK, 0, s>>>(S, d->d_status);
    k_sort_cells<<<grid_of(d->n_cells), BLOCK, 0, s>>>(S, d->d_status);
    FD_HIP(hipGetLastError());
    return FT_OK;
}

// The run's one host wait: status (and, with loss, the device-summed loss).
ft_status finish(ft_diff *d, bool loss = false) {
    FD_HIP(hipMemcpyAsync(d->h_status, d->d_status, sizeof(DevStatus), hipMemcpyDeviceToHost, d->stream));
    if (loss)
        FD_HIP(hipMemcpyAsync(d->h_loss, d->d_partial + LOSS_GRID, sizeof(double), hipMemcpyDeviceToHost, d->stream));
    FD_HIP(hipStreamSynchronize(d->stream));
    if (d->h_status->err) return (ft_status)d->h_status->err;
    if (d->h_status->overflow) return (ft_status)0x100;
    return FT_OK;
}

ft_status grow_pairs(ft_diff *d) {
    const uint64_t need = (uint64_t)d->h_status->need_pairs;
    const uint64_t cap = need + need / 4 + 4096;
    if (cap >= 0xFFFFFFFFull) return FT_ENOMEM;
    ft_status st = dalloc(d->d_cell_edge, cap);
    if (st) return st;
    d->cap_pairs = (uint32_t)cap;
    return FT_OK;
}

dim3 raster_grid(const ft_diff *d) { return dim3((d->W + DBLOCK - 1) / DBLOCK, d->H, d->n_images); }

ft_status run_forward(ft_diff *d, const float *points, const float *colours, float *canvas) {
    for (int attempt = 0; attempt < 3; ++attempt) {
        ft_status st = launch_geometry(d, points, false, nullptr);
        if (st) return st;
        k_forward<<<raster_grid(d), DBLOCK, 0, d->stream>>>(scene_of(d), reinterpret_cast<const float4 *>(colours),
                                                             reinterpret_cast<float4 *>(canvas), nullptr,
                                                             d->ck_stride, nullptr, d->d_status);
        FD_HIP(hipGetLastError());
        st = finish(d);
        if (st != (ft_status)0x100) return st;
        if ((st = grow_pairs(d))) return st;
    }
    return FT_ENOMEM;
}

// Enqueues the loss partials and their fixed-order sum (at d_partial[LOSS_GRID]).
ft_status launch_loss(ft_diff *d, const float *canvas, const float *target, float *grad) {
    k_loss<<<LOSS_GRID, BLOCK, 0, d->stream>>>(canvas, target, grad, d->canvas_floats, d->d_partial);
    k_loss_sum<<<1, BLOCK, 0, d->stream>>>(d->d_partial, LOSS_GRID, d->d_partial + LOSS_GRID);
    FD_HIP(hipGetLastError());
    return FT_OK;
}

ft_status run_loss(ft_diff *d, const float *canvas, const float *target, float *grad, double *loss) {
    ft_status st = launch_loss(d, canvas, target, grad);
    if (!st) {
        FD_HIP(hipMemcpyAsync(d->h_loss, d->d_partial + LOSS_GRID, sizeof(double), hipMemcpyDeviceToHost,
                              d->stream));
        FD_HIP(hipStreamSynchronize(d->stream));
        if (loss) *loss = *d->h_loss / (double)d->canvas_floats;
    }
    return st;
}

// Enqueues the backward kernels.  replayed: k_forward with checkpoints of the
// same points and colours ran just before on the same geometry.
ft_status launch_backward(ft_diff *d, const float *colours, const float *grad_canvas, float *grad_points,
                          float *grad_colours, bool replayed) {
    DScene S = scene_of(d);
    k_backward<<<raster_grid(d), DBLOCK, 0, d->stream>>>(
        S, reinterpret_cast<const float4 *>(colours), reinterpret_cast<const float4 *>(grad_canvas), d->d_ck,
        d->ck_stride, d->d_cov, replayed ? 1u : 0u, d->d_gvtx, grad_colours, d->d_status);
    k_point_grad<<<grid_of(d->n_pts), BLOCK, 0, d->stream>>>(S, d->d_gvtx, reinterpret_cast<float2 *>(grad_points),
                                                             d->n_pts, d->d_pt_path);
    FD_HIP(hipGetLastError());
    return FT_OK;
}

ft_status run_backward(ft_diff *d, const float *points, const float *colours, const float *grad_canvas,
                       float *grad_points, float *grad_colours) {
    for (int attempt = 0; attempt < 3; ++attempt) {
        ft_status st = launch_geometry(d, points, true, grad_colours);
        if (!st) st = launch_backward(d, colours, grad_canvas, grad_points, grad_colours, false);
        if (st) return st;
        st = finish(d);
        if (st != (ft_status)0x100) return st;
        if ((st = grow_pairs(d))) return st;
    }
    return FT_ENOMEM;
}

// forward (leaving checkpoints), loss and backward on one geometry, with one
// host wait; a pair overflow turns the later kernels into no-ops, and the
// step is run again with room for the pairs.
ft_status run_step(ft_diff *d, const float *points, const float *colours, const float *target, float *canvas,
                   float *grad_canvas, float *grad_points, float *grad_colours, double *loss) {
    for (int attempt = 0; attempt < 3; ++attempt) {
        ft_status st = launch_geometry(d, points, true, grad_colours);
        if (st) return st;
        k_forward<<<raster_grid(d), DBLOCK, 0, d->stream>>>(scene_of(d), reinterpret_cast<const float4 *>(colours),
                                                             reinterpret_cast<float4 *>(canvas), d->d_ck,
                                                             d->ck_stride, d->d_cov, d->d_status);
        FD_HIP(hipGetLastError());
        if ((st = launch_loss(d, canvas, target, grad_canvas))) return st;
        if ((st = launch_backward(d, colours, grad_canvas, grad_points, grad_colours, true))) return st;
        st = finish(d, true);
        if (st == FT_OK && loss) *loss = *d->h_loss / (double)d->canvas_floats;
        if (st != (ft_status)0x100) return st;
        if ((st = grow_pairs(d))) return st;
    }
    return FT_ENOMEM;
}

}  // namespace

extern "C" {

ft_status ft_diff_new(const ft_diff_layout *L, ft_diff **out) {
    if (!L || !out || !L->path_image || !L->path_cubics || !L->n_images || !L->width || !L->height ||
        (L->flags & ~FT_DIFF_KEEP_COVERAGE))
        return FT_EINVAL;
    if (L->width > FT_MAX_WIDTH || L->height > FT_MAX_WIDTH) return FT_ENOTSUP;
    std::vector<uint32_t> pt_off(L->n_paths + 1), vtx_off(L->n_paths + 1), img_first(L->n_images + 1);
    uint64_t np = 0, nv = 0;
    for (uint32_t k = 0; k < L->n_paths; ++k) {
        if (L->path_image[k] >= L->n_images || (k && L->path_image[k] < L->path_image[k - 1])) return FT_EINVAL;
        if (L->path_cubics[k] == 0 || L->path_cubics[k] > (1u << 20)) return FT_EINVAL;
        pt_off[k] = (uint32_t)np;
        vtx_off[k] = (uint32_t)nv;
        np += 1 + 3ull * L->path_cubics[k];
        nv += 1 + (uint64_t)DSEG * L->path_cubics[k];
        if (np >= 0x7FFFFFFFull || nv >= 0x7FFFFFFFull) return FT_ENOTSUP;
    }
    pt_off[L->n_paths] = (uint32_t)np;
    vtx_off[L->n_paths] = (uint32_t)nv;
    uint32_t max_paths = 0;
    {
        uint32_t k = 0;
        for (uint32_t i = 0; i <= L->n_images; ++i) {
            while (k < L->n_paths && L->path_image[k] < i) ++k;
            img_first[i] = k;
        }
        for (uint32_t i = 0; i < L->n_images; ++i) max_paths = std::max(max_paths, img_first[i + 1] - img_first[i]);
    }
    const uint64_t cells = (uint64_t)L->n_paths * L->height;
    if (cells >= 0xFFFFFFFFull) return FT_ENOTSUP;
    std::vector<uint32_t> vtx_path(nv), pt_path(np);
    for (uint32_t k = 0; k < L->n_paths; ++k) {
        for (uint32_t v = vtx_off[k]; v < vtx_off[k + 1]; ++v) vtx_path[v] = k;
        for (uint32_t m = pt_off[k]; m < pt_off[k + 1]; ++m) pt_path[m] = k;
    }
    ft_diff *d = new ft_diff();
    d->n_images = L->n_images;
    d->W = L->width;
    d->H = L->height;
    d->n_paths = L->n_paths;
    d->max_paths = max_paths;
    d->n_pts = (uint32_t)np;
    d->n_vtx = (uint32_t)nv;
    d->n_cells = cells;
    d->canvas_floats = 4ull * L->n_images * L->width * L->height;
    const dim3 g = raster_grid(d);
    d->ck_stride = (uint64_t)g.x * g.y * g.z * DBLOCK;
    const uint64_t n_ck = std::max<uint64_t>(1, (max_paths + KSEG - 1) / KSEG);
    d->cap_tiles = (cells + SCAN_TILE_MIN - 1) / SCAN_TILE_MIN + 1;
    d->cap_pairs = (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, 8 * nv + 4096);
    ft_status st;
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) st = FT_EDEVICE;
    else if ((st = dalloc(d->d_pt_off, L->n_paths + 1)) || (st = dalloc(d->d_vtx_off, L->n_paths + 1)) ||
             (st = dalloc(d->d_vtx_path, nv)) || (st = dalloc(d->d_pt_path, np)) ||
             (st = dalloc(d->d_img_first, L->n_images + 1)) || (st = dalloc(d->d_vtx, nv)) ||
             (st = dalloc(d->d_gvtx, nv)) || (st = dalloc(d->d_cell_cnt, cells)) ||
             (st = dalloc(d->d_cell_off, cells)) || (st = dalloc(d->d_cell_cur, cells)) ||
             (st = dalloc(d->d_cell_edge, d->cap_pairs)) || (st = dalloc(d->d_lb, 4 * d->cap_tiles)) ||
             (st = dalloc(d->d_ck, n_ck * d->ck_stride)) || (st = dalloc(d->d_partial, LOSS_GRID + 1)) ||
             (st = dalloc(d->d_status, 1)))
        ;
    else if (hipHostMalloc(reinterpret_cast<void **>(&d->h_loss), sizeof(double), hipHostMallocDefault) !=
                 hipSuccess ||
             hipHostMalloc(reinterpret_cast<void **>(&d->h_status), sizeof(DevStatus), hipHostMallocDefault) !=
                 hipSuccess)
        st = FT_ENOMEM;
    else if (hipMemcpy(d->d_pt_off, pt_off.data(), pt_off.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d->d_vtx_off, vtx_off.data(), vtx_off.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d->d_vtx_path, vtx_path.data(), nv * 4, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d->d_pt_path, pt_path.data(), np * 4, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d->d_img_first, img_first.data(), img_first.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        st = FT_EDEVICE;
    else
        st = FT_OK;
    // By default the backward recomputes every path's coverage.  On request
    // (FT_DIFF_KEEP_COVERAGE) the forward keeps it, 4 B per path and pixel,
    // which spares the second evaluation; without room for the buffer the
    // handle recomputes anyway.
    const uint64_t cov_elems = (L->flags & FT_DIFF_KEEP_COVERAGE) ? (uint64_t)max_paths * d->ck_stride : 0;
    if (!st && cov_elems && cov_elems <= COV_MAX_BYTES / sizeof(float) && dalloc(d->d_cov, cov_elems)) {
        d->d_cov = nullptr;
        (void)hipGetLastError();  // the failed optional allocation is not an error of a later call
    }
    if (st) {
        ft_diff_free(d);
        return st;
    }
    *out = d;
    return FT_OK;
}

void ft_diff_free(ft_diff *d) {
    if (!d) return;
    void *bufs[] = {d->d_pt_off,   d->d_vtx_off,  d->d_vtx_path, d->d_pt_path,   d->d_img_first, d->d_vtx,
                    d->d_gvtx,     d->d_cell_cnt, d->d_cell_off, d->d_cell_cur,  d->d_cell_edge, d->d_lb,
                    d->d_ck,       d->d_cov,       d->d_partial,   d->d_status};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (d->h_loss) (void)hipHostFree(d->h_loss);
    if (d->h_status) (void)hipHostFree(d->h_status);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

ft_status ft_diff_info(const ft_diff *d, uint64_t *n_points, uint64_t *n_paths, uint64_t *canvas_floats) {
    if (!d) return FT_EINVAL;
    if (n_points) *n_points = d->n_pts;
    if (n_paths) *n_paths = d->n_paths;
    if (canvas_floats) *canvas_floats = d->canvas_floats;
    return FT_OK;
}

ft_status ft_diff_forward(ft_diff *d, const float *points, const float *colours, float *canvas) {
    if (!d || !points || !colours || !canvas) return FT_EINVAL;
    return run_forward(d, points, colours, canvas);
}

ft_status ft_diff_loss(ft_diff *d, const float *canvas, const float *target, float *grad_canvas, double *loss) {
    if (!d || !canvas || !target) return FT_EINVAL;
    return run_loss(d, canvas, target, grad_canvas, loss);
}

ft_status ft_diff_backward(ft_diff *d, const float *points, const float *colours, const float *grad_canvas,
                           float *grad_points, float *grad_colours) {
    if (!d || !points || !colours || !grad_canvas || !grad_points || !grad_colours) return FT_EINVAL;
    return run_backward(d, points, colours, grad_canvas, grad_points, grad_colours);
}

ft_status ft_diff_step(ft_diff *d, const float *points, const float *colours, const float *target, float *canvas,
                       float *grad_canvas, float *grad_points, float *grad_colours, double *loss) {
    if (!d || !points || !colours || !target || !canvas || !grad_canvas || !grad_points || !grad_colours)
        return FT_EINVAL;
    return run_step(d, points, colours, target, canvas, grad_canvas, grad_points, grad_colours, loss);
}

}  // extern "C"
