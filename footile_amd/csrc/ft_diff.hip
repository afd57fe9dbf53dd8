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
constexpr int DBLOCK = 256;       // pixels of one row per block (four wavefronts)
constexpr int KSEG = 8;           // paths per checkpoint segment
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

// One DPP step of a wavefront float scan: v + (v of the source lane, or 0).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float v) {
    const int src = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xf, false);
    return v + __builtin_bit_cast(float, src);
}

// Sum over the 64 lanes (all active), returned to every lane: an inclusive
// scan with DPP (row_shr 1, 2, 4, 8 within rows of 16 lanes, then
// row_bcast:15 and row_bcast:31), read from lane 63.  The same lanes add in
// the same order on every call.
__device__ inline float wave_sum(float v) {
    v = dpp_add<0x111, 0xf>(v);
    v = dpp_add<0x112, 0xf>(v);
    v = dpp_add<0x114, 0xf>(v);
    v = dpp_add<0x118, 0xf>(v);
    v = dpp_add<0x142, 0xa>(v);
    v = dpp_add<0x143, 0xc>(v);
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// Three sums at once (their DPP steps interleave).
__device__ inline void wave_sum3(float &a, float &b, float &c) {
#define FT_DPP3(CTRL, RM) a = dpp_add<CTRL, RM>(a); b = dpp_add<CTRL, RM>(b); c = dpp_add<CTRL, RM>(c);
    FT_DPP3(0x111, 0xf)
    FT_DPP3(0x112, 0xf)
    FT_DPP3(0x114, 0xf)
    FT_DPP3(0x118, 0xf)
    FT_DPP3(0x142, 0xa)
    FT_DPP3(0x143, 0xc)
#undef FT_DPP3
    a = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a), 63));
    b = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, b), 63));
    c = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c), 63));
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

// Edge lists per (path, row) in edge order, one block per path, without
// atomics or a sort: the block stages its edges' row ranges in LDS; each
// wavefront takes every fourth row of the path's span and walks the edges in
// chunks of 64, giving each crossing edge its slot by a ballot count of the
// crossing edges before it.  Paths of more than CELL_EDGES edges take the
// atomic fill and the sort below instead.
constexpr uint32_t CELL_EDGES = 2048;

__global__ __launch_bounds__(BLOCK) void k_cells(DScene S, DevStatus *st) {
    __shared__ uint2 s_rows[CELL_EDGES];
    __shared__ uint32_t s_lo[BLOCK / 64], s_hi[BLOCK / 64];
    if (run_failed(st)) return;
    if (st->n_pairs > S.cap_pairs) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(&st->overflow, 2u);
            st->need_pairs = st->n_pairs;
        }
        return;
    }
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t k = blockIdx.x; k < S.n_paths; k += gridDim.x) {
        const uint32_t vb = S.vtx_off[k], V = S.vtx_off[k + 1] - vb;
        if (V > CELL_EDGES) continue;  // uniform: left to the atomic fill
        uint32_t lo = 0xFFFFFFFFu, hi = 0;
        for (uint32_t e = threadIdx.x; e < V; e += BLOCK) {
            uint32_t r0 = 1, r1 = 0;
            const float2 a = S.vtx[vb + e], b = S.vtx[vb + (e + 1 < V ? e + 1 : 0u)];
            if (!edge_row_range(a, b, S.H, r0, r1)) {
                r0 = 1;
                r1 = 0;
            }
            s_rows[e] = make_uint2(r0, r1);
            if (r0 <= r1) {
                lo = r0 < lo ? r0 : lo;
                hi = r1 > hi ? r1 : hi;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
            lo = l2 < lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        }
        if (lane == 0) {
            s_lo[wid] = lo;
            s_hi[wid] = hi;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) {
            lo = s_lo[w] < lo ? s_lo[w] : lo;
            hi = s_hi[w] > hi ? s_hi[w] : hi;
        }
        const uint64_t c0 = (uint64_t)k * S.H;
        uint32_t bad = 0;
        for (uint32_t rb = lo + wid * 64; lo <= hi && rb <= hi; rb += BLOCK) {
            // the list offsets of this wavefront's next 64 rows rb, rb + 1, ...
            // (rows interleave over the wavefronts in runs of 64)
            const uint32_t my_row = rb + lane;
            const uint32_t my_off = my_row <= hi ? S.cell_off[c0 + my_row] : 0u;
            const uint32_t n_r = hi - rb + 1 < 64u ? hi - rb + 1 : 64u;
            for (uint32_t i = 0; i < n_r; ++i) {
                const uint32_t r = rb + i;
                const uint32_t base = (uint32_t)__shfl((int)my_off, (int)i, 64);
                uint32_t run = 0;
                for (uint32_t e0 = 0; e0 < V; e0 += 64) {
                    const uint32_t e = e0 + lane;
                    const uint2 rr = e < V ? s_rows[e] : make_uint2(1u, 0u);
                    const bool x = (rr.x <= r) & (r <= rr.y);
                    const unsigned long long m = __ballot(x);
                    if (x) {
                        const uint32_t at = base + run +
                                            __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        if (at < S.cap_pairs)
                            S.cell_edge[at] = e;
                        else
                            bad = 1;
                    }
                    run += (uint32_t)__popcll(m);
                }
            }
        }
        if (bad) raise_err(st, FT_EDEVICE);
        __syncthreads();  // s_rows and s_lo / s_hi are the next path's
    }
}

// FILL_LANES threads per edge, each taking every FILL_LANES-th run of
// FILL_ROWS rows, with the run's FILL_ROWS slot reservations (returning
// atomics on the cells' cursors) in flight together: a tall edge does not
// wait on one atomic per row in turn.
constexpr uint32_t FILL_LANES = 4, FILL_ROWS = 8;

__global__ __launch_bounds__(BLOCK) void k_fill_cells(DScene S, DevStatus *st) {
    if (run_failed(st)) return;
    if (st->n_pairs > S.cap_pairs) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(&st->overflow, 2u);
            st->need_pairs = st->n_pairs;
        }
        return;
    }
    const uint64_t n_items = (uint64_t)S.n_vtx * FILL_LANES;
    for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t < n_items; t += (uint64_t)gridDim.x * BLOCK) {
        const uint32_t v = (uint32_t)(t / FILL_LANES), q = (uint32_t)(t % FILL_LANES);
        uint32_t k, r0, r1;
        float2 a, b;
        edge_at(S, v, k, a, b);
        if (S.vtx_off[k + 1] - S.vtx_off[k] <= CELL_EDGES) continue;  // k_cells has the path
        if (!edge_row_range(a, b, S.H, r0, r1)) continue;
        const uint32_t local = v - S.vtx_off[k];
        const uint64_t c0 = (uint64_t)k * S.H;
        for (uint32_t rb = r0 + q * FILL_ROWS; rb <= r1; rb += FILL_LANES * FILL_ROWS) {
            uint32_t at[FILL_ROWS];
#pragma unroll
            for (uint32_t u = 0; u < FILL_ROWS; ++u)
                at[u] = rb + u <= r1 ? S.cell_off[c0 + rb + u] + atomicAdd(&S.cell_cur[c0 + rb + u], 1u) : 0u;
#pragma unroll
            for (uint32_t u = 0; u < FILL_ROWS; ++u) {
                if (rb + u > r1) break;
                if (at[u] < S.cap_pairs)
                    S.cell_edge[at[u]] = local;
                else
                    raise_err(st, FT_EDEVICE);
            }
        }
    }
}

// Each cell's edges in ascending order, so that sums run in edge order.  A
// cell of 2 to 16 edges is loaded into registers (all loads in flight),
// sorted by an odd-even transposition network of 8 or 16 places and stored
// back; longer cells are sorted in place.
template <uint32_t N>
__device__ inline void sort_cell_regs(uint32_t *e, uint32_t n) {
    uint32_t x[N];
#pragma unroll
    for (uint32_t i = 0; i < N; ++i) x[i] = i < n ? e[i] : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t pass = 0; pass < N; ++pass) {
#pragma unroll
        for (uint32_t i = pass & 1u; i + 1 < N; i += 2) {
            const uint32_t lo = x[i] < x[i + 1] ? x[i] : x[i + 1];
            const uint32_t hi = x[i] < x[i + 1] ? x[i + 1] : x[i];
            x[i] = lo;
            x[i + 1] = hi;
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < N; ++i)
        if (i < n) e[i] = x[i];
}

__global__ __launch_bounds__(BLOCK) void k_sort_cells(DScene S, DevStatus *st) {
    if (run_failed(st)) return;
    for (uint32_t c = blockIdx.x * BLOCK + threadIdx.x; c < S.n_cells; c += gridDim.x * BLOCK) {
        const uint32_t off = S.cell_off[c], n = S.cell_cnt[c];
        if ((uint64_t)off + n > S.cap_pairs) {
            raise_err(st, FT_EDEVICE);
            continue;
        }
        if (n < 2 || S.vtx_off[c / S.H + 1] - S.vtx_off[c / S.H] <= CELL_EDGES) continue;  // k_cells: in order
        uint32_t *e = S.cell_edge + off;
        if (n <= 8) {
            sort_cell_regs<8>(e, n);
        } else if (n <= 16) {
            sort_cell_regs<16>(e, n);
        } else {
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
}

// ---------------------------------------------------------------------------
// Forward, loss, backward
// ---------------------------------------------------------------------------
// Row staging.  A block owns DBLOCK pixels of one row of one image.  It
// stages the (cell count, list offset, colour) of a run of the image's paths
// for its row (stage_paths) and then their row pieces in windows of ECAP
// (one thread per piece: edge index, two vertices, piece_of), so that the
// per-pixel loops read every piece and colour from LDS with one broadcast
// load instead of a chain of dependent global loads per edge and path.
constexpr uint32_t PCHUNK = DBLOCK;  // paths staged per round (one per thread)
constexpr uint32_t ECAP = 512;       // pieces staged per window
constexpr uint32_t LAST_BIT = 0x80000000u, LIVE_BIT = 0x40000000u, J_MASK = 0x3FFFFFFFu;

struct Stage {
    uint32_t *cnt;  // [PCHUNK] pieces of each staged path in the row
    uint32_t *off;  // [PCHUNK] its cell's list offset
    uint32_t *pb;   // [PCHUNK] exclusive scan of cnt
    float4 *col;    // [PCHUNK] its colour
    float4 *pc;     // [ECAP] piece: x at top, x at bottom, sign * height, 0
    uint32_t *ev;   // [ECAP] staged path index | LAST_BIT | LIVE_BIT
    uint2 *vi;      // [ECAP] global indices of the edge's two vertices (or null: not staged)
    float4 *pq;     // [ECAP] their coordinates (or null)
    uint32_t *w;    // [DBLOCK / 64] scan scratch
};

// Counts, offsets and colours of row Y for the n <= PCHUNK paths from global
// path kbase; returns their total number of pieces.  Every thread calls it.
__device__ uint32_t stage_paths(const DScene &S, const float4 *colours, uint32_t kbase, uint32_t n, int Y,
                                const Stage &sg) {
    const uint32_t t = threadIdx.x;
    const bool in = t < n;
    const uint64_t cell = (uint64_t)(kbase + (in ? t : 0u)) * S.H + (uint32_t)Y;
    const uint32_t c = in ? S.cell_cnt[cell] : 0u;
    const uint32_t o = in ? S.cell_off[cell] : 0u;
    const float4 col = colours[kbase + (in ? t : 0u)];
    const uint32_t incl = wave_incl_dpp<false>(c);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) sg.w[wid] = incl;
    __syncthreads();
    uint32_t pre = incl - c, total = 0;
#pragma unroll
    for (int w = 0; w < DBLOCK / 64; ++w) {
        pre += w < wid ? sg.w[w] : 0u;
        total += sg.w[w];
    }
    if (in) {
        sg.cnt[t] = c;
        sg.off[t] = o;
        sg.pb[t] = pre;
        sg.col[t] = col;
    }
    __syncthreads();
    return total;
}

// Pieces [e0, e0 + n) (n <= ECAP) of the np staged paths into the piece
// arrays at [0, n).  Piece e belongs to the last staged path t with
// pb[t] <= e.  The caller synchronises before and after.
__device__ void stage_pieces(const DScene &S, uint32_t kbase, uint32_t np, int Y, uint32_t e0, uint32_t n,
                             const Stage &sg, uint32_t &bad) {
    for (uint32_t i = threadIdx.x; i < n; i += DBLOCK) {
        const uint32_t e = e0 + i;
        uint32_t lo = 0, hi = np;  // pb[lo] <= e < pb[hi] (pb[np] taken as the total)
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sg.pb[mid] <= e)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t q = e - sg.pb[lo];
        const uint32_t k = kbase + lo;
        const uint32_t vb = S.vtx_off[k], V = S.vtx_off[k + 1] - vb;
        const uint32_t edge = S.cell_edge[sg.off[lo] + q];
        bad |= edge < V ? 0u : 1u;
        const uint32_t i0 = edge < V ? edge : 0u;
        const uint32_t i1 = i0 + 1 < V ? i0 + 1 : 0u;
        const float2 a = S.vtx[vb + i0], b = S.vtx[vb + i1];
        const Piece r = piece_of(a, b, Y);
        sg.pc[i] = make_float4(r.xa, r.xb, r.live ? r.sign * r.h : 0.0f, 0.0f);
        sg.ev[i] = lo | (q + 1 == sg.cnt[lo] ? LAST_BIT : 0u) | (r.live ? LIVE_BIT : 0u);
        if (sg.vi) {
            sg.vi[i] = make_uint2(vb + i0, vb + i1);
            sg.pq[i] = make_float4(a.x, a.y, b.x, b.y);
        }
    }
}

// sign * h * M of a staged piece at pixel X, M = mean_clamp(X + 1 - xa,
// X + 1 - xb).  A wavefront whose pixels [xw0, xw0 + 63] all lie left of the
// piece (X + 1 <= min x: M = 0) or all right of it (X >= max x: M = 1) takes
// the answer without evaluating M; these are exactly the values mean_clamp
// returns there.
__device__ inline float piece_cover(float4 pc, float Xf, float xw0) {
    const float lo = fminf(pc.x, pc.y), hi = fmaxf(pc.x, pc.y);
    if (xw0 + 64.0f <= lo) return 0.0f;
    if (xw0 >= hi) return pc.z;
    float da, db;
    return pc.z * mean_clamp(Xf + 1.0f - pc.x, Xf + 1.0f - pc.y, da, db);
}

// The image's np paths (from global path k0) composited over a transparent
// pixel (X, Y).  With ck it also writes the canvas before every KSEG-th path
// (the backward's checkpoints) and, with cov, each path's coverage (paths
// with no piece in the row are left out, as their coverage is 0).  Every
// thread of the block calls it.
__device__ float4 forward_row(const DScene &S, const float4 *colours, uint32_t k0, uint32_t np, int X, int Y,
                              uint64_t slot, float4 *ck, uint64_t ck_stride, float *cov, const Stage &sg,
                              uint32_t &bad) {
    float4 C = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float A = 0.0f;
    uint32_t next_ck = 0;
    const float Xf = (float)X;
    const float xw0 = (float)(X - (int)(threadIdx.x & 63));
    for (uint32_t ja = 0; ja < np; ja += PCHUNK) {
        const uint32_t n = np - ja < PCHUNK ? np - ja : PCHUNK;
        __syncthreads();  // the previous round's staging is no longer read
        const uint32_t total = stage_paths(S, colours, k0 + ja, n, Y, sg);
        for (uint32_t e0 = 0; e0 < total; e0 += ECAP) {
            const uint32_t ne = total - e0 < ECAP ? total - e0 : ECAP;
            __syncthreads();
            stage_pieces(S, k0 + ja, n, Y, e0, ne, sg, bad);
            __syncthreads();
            auto close_path = [&](uint32_t ev) {  // the path's last piece: composite it
                const uint32_t t = ev & J_MASK, j = ja + t;
                if (ck)
                    for (; next_ck * KSEG <= j; ++next_ck) ck[(uint64_t)next_ck * ck_stride + slot] = C;
                if (ck && cov) cov[(uint64_t)j * ck_stride + slot] = A;  // j < max_paths, slot < ck_stride
                float dadA;
                C = over(sg.col[t], alpha_of(A, dadA), C);
                A = 0.0f;
            };
            // four pieces' loads at a time, their covers added in order
            uint32_t i = 0;
            for (; i + 4 <= ne; i += 4) {
                const float4 p0 = sg.pc[i], p1 = sg.pc[i + 1], p2 = sg.pc[i + 2], p3 = sg.pc[i + 3];
                const uint4 ev = *reinterpret_cast<const uint4 *>(sg.ev + i);
                const float c0 = piece_cover(p0, Xf, xw0), c1 = piece_cover(p1, Xf, xw0);
                const float c2 = piece_cover(p2, Xf, xw0), c3 = piece_cover(p3, Xf, xw0);
                A += c0;
                if (ev.x & LAST_BIT) close_path(ev.x);
                A += c1;
                if (ev.y & LAST_BIT) close_path(ev.y);
                A += c2;
                if (ev.z & LAST_BIT) close_path(ev.z);
                A += c3;
                if (ev.w & LAST_BIT) close_path(ev.w);
            }
            for (; i < ne; ++i) {
                A += piece_cover(sg.pc[i], Xf, xw0);
                const uint32_t ev = sg.ev[i];
                if (ev & LAST_BIT) close_path(ev);
            }
        }
    }
    if (ck)
        for (; next_ck * KSEG < np; ++next_ck) ck[(uint64_t)next_ck * ck_stride + slot] = C;
    return C;
}

#define DIFF_STAGE_LDS(VI, PQ)                                          \
    __shared__ uint32_t s_cnt[PCHUNK], s_off[PCHUNK], s_pb[PCHUNK];    \
    __shared__ float4 s_col[PCHUNK];                                    \
    __shared__ float4 s_pc[ECAP];                                       \
    __shared__ __align__(16) uint32_t s_ev[ECAP];                       \
    __shared__ uint32_t s_w[DBLOCK / 64];                               \
    const Stage sg{s_cnt, s_off, s_pb, s_col, s_pc, s_ev, VI, PQ, s_w}

// With ck non-null it also leaves what k_backward's replay would write: the
// canvas before every KSEG-th path, and each path's coverage when cov is
// non-null (same slots as k_backward).
__global__ __launch_bounds__(DBLOCK) void k_forward(DScene S, const float4 *colours, float4 *canvas, float4 *ck,
                                                    uint64_t ck_stride, float *cov, DevStatus *st) {
    DIFF_STAGE_LDS(nullptr, nullptr);
    if (run_failed(st)) return;
    const int X = (int)(blockIdx.x * DBLOCK + threadIdx.x), Y = (int)blockIdx.y;
    const uint32_t im = blockIdx.z;
    const uint32_t k0 = S.img_first[im], k1 = S.img_first[im + 1];
    const uint64_t slot = ((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * (gridDim.x * DBLOCK) +
                          (uint64_t)blockIdx.x * DBLOCK + threadIdx.x;
    uint32_t bad = 0;
    const float4 C = forward_row(S, colours, k0, k1 - k0, X, Y, slot, ck, ck_stride, ck ? cov : nullptr, sg, bad);
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

// One lane per pixel.  Pass 1 replays the forward into the checkpoints
// (skipped when k_forward of the same inputs has written them: `replayed`).
// Pass 2 walks the checkpoint segments last to first.  Per segment: the
// segment's paths (counts, offsets, colours) and their coverage per pixel
// into LDS (recomputed from the staged pieces, or read from the kept
// coverage buffer); then its paths last to first, with no block barrier: the
// canvas under path j re-composited from the segment's checkpoint, the colour
// gradient (wave sums into LDS), and per piece of path j the sums S1 = sum
// dA M, S2 = sum dA dM/dua, S3 = sum dA dM/dub over the block's pixels (a
// wavefront wholly right of the piece adds its dA total to S1 alone, one
// wholly left adds nothing), accumulated in LDS.  When the walk leaves the
// staged window of pieces (and at the segment's end) one thread per piece
// forms its vertex gradients from S1..S3 with global atomics; the colour
// gradients are flushed once per segment.
__global__ __launch_bounds__(DBLOCK) void k_backward(DScene S, const float4 *colours, const float4 *grad_canvas,
                                                     float4 *ck, uint64_t ck_stride, float *cov, uint32_t replayed,
                                                     float2 *gvtx, float *gcol, DevStatus *st) {
    __shared__ uint2 s_vi[ECAP];
    __shared__ float4 s_pq[ECAP];
    DIFF_STAGE_LDS(s_vi, s_pq);
    __shared__ float s_A[KSEG][DBLOCK];  // coverage of the segment's paths; each thread uses its own column
    __shared__ float4 s_S[ECAP];         // per staged piece: S1, S2, S3
    __shared__ float s_gc[KSEG][4];      // colour gradients of the segment's paths
    // coverage recompute, per staged window: each wavefront's sum of the
    // pieces wholly left of it per path, and its list of the pieces that
    // touch its columns
    constexpr int NW = DBLOCK / 64;
    __shared__ float s_Lw[KSEG][NW];
    __shared__ uint16_t s_tl[NW][ECAP];
    __shared__ uint32_t s_tn[NW];
    if (run_failed(st)) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int X = (int)(blockIdx.x * DBLOCK + tid), Y = (int)blockIdx.y;
    const float Xf = (float)X, xw0 = (float)(X - lane);
    const uint32_t im = blockIdx.z;
    const bool active = (uint32_t)X < S.W;
    const uint64_t slot = ((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * (gridDim.x * DBLOCK) +
                          (uint64_t)blockIdx.x * DBLOCK + tid;
    const uint32_t k0 = S.img_first[im], np = S.img_first[im + 1] - k0;
    uint32_t bad = 0;
    if (!replayed) (void)forward_row(S, colours, k0, np, X, Y, slot, ck, ck_stride, cov, sg, bad);
    float4 g = active ? grad_canvas[((uint64_t)im * S.H + (uint32_t)Y) * S.W + (uint32_t)X]
                      : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // the vertex gradients of the staged window's pieces [0, n) from their
    // sums (then cleared); every thread calls it
    // (every thread owns staged pieces tid and tid + DBLOCK; both are loaded
    // before either's gradient is formed)
    static_assert(ECAP == 2 * DBLOCK, "two staged pieces per thread");
    auto flush_pieces = [&](uint32_t n) {
        __syncthreads();
        const uint32_t ia = tid, ib = tid + DBLOCK;
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 Sa = ia < n ? s_S[ia] : zero, Sb = ib < n ? s_S[ib] : zero;
        const uint32_t ea = s_ev[ia], eb = s_ev[ib];
        const float4 pqa = s_pq[ia], pqb = s_pq[ib];
        const uint2 via = s_vi[ia], vib = s_vi[ib];
        auto grad_of = [&](float4 Sm, uint32_t ev, float4 pq, uint2 vi) {
            if (((Sm.x != 0.0f) | (Sm.y != 0.0f) | (Sm.z != 0.0f)) && (ev & LIVE_BIT)) {
                float gr[4];
                piece_grad(make_float2(pq.x, pq.y), make_float2(pq.z, pq.w), Y, Sm.x, Sm.y, Sm.z, gr);
                atomicAdd(&gvtx[vi.x].x, gr[0]);
                atomicAdd(&gvtx[vi.x].y, gr[1]);
                atomicAdd(&gvtx[vi.y].x, gr[2]);
                atomicAdd(&gvtx[vi.y].y, gr[3]);
            }
        };
        grad_of(Sa, ea, pqa, via);
        grad_of(Sb, eb, pqb, vib);
        if (ia < n) s_S[ia] = zero;
        if (ib < n) s_S[ib] = zero;
        __syncthreads();
    };
    for (uint32_t i = tid; i < ECAP; i += DBLOCK) s_S[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t n_seg = (np + KSEG - 1) / KSEG;
    for (uint32_t seg = n_seg; seg-- > 0;) {
        const uint32_t j0 = seg * KSEG, j1 = j0 + KSEG < np ? j0 + KSEG : np;
        const uint32_t ns = j1 - j0;
        __syncthreads();  // the previous segment's staging is no longer read
        const uint32_t total = stage_paths(S, colours, k0 + j0, ns, Y, sg);
        if (tid < KSEG * 4) s_gc[tid >> 2][tid & 3] = 0.0f;
        // the segment's coverage per path
        for (uint32_t j = 0; j < ns; ++j) s_A[j][tid] = 0.0f;
        uint32_t w_lo = 0, w_n = 0;  // staged window: pieces [w_lo, w_lo + w_n) of the segment
        if (cov) {
            for (uint32_t j = 0; j < ns; ++j)
                if (s_cnt[j]) s_A[j][tid] = cov[(uint64_t)(j0 + j) * ck_stride + slot];
        } else {
            const int wid = tid >> 6;
            const float xb0 = (float)(blockIdx.x * DBLOCK);
            for (uint32_t e0 = 0; e0 < total; e0 += ECAP) {
                const uint32_t ne = total - e0 < ECAP ? total - e0 : ECAP;
                __syncthreads();
                stage_pieces(S, k0 + j0, ns, Y, e0, ne, sg, bad);
                if (tid < KSEG * NW) s_Lw[tid / NW][tid % NW] = 0.0f;
                if (tid < NW) s_tn[tid] = 0;
                __syncthreads();
                w_lo = e0;
                w_n = ne;
                // each staged piece against each wavefront's 64 columns (the
                // tests of piece_cover): wholly right adds nothing, wholly
                // left its full sign * height, else it goes on the list
                for (uint32_t i = tid; i < ne; i += DBLOCK) {
                    const float4 pc = s_pc[i];
                    const uint32_t ev = s_ev[i];
                    if (!(ev & LIVE_BIT)) continue;  // no coverage
                    const float lo = fminf(pc.x, pc.y), hi = fmaxf(pc.x, pc.y);
                    const uint32_t j = ev & J_MASK;
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        const float x0 = xb0 + 64.0f * (float)w;
                        if (x0 + 64.0f <= lo) continue;
                        if (x0 >= hi) {
                            atomicAdd(&s_Lw[j][w], pc.z);
                        } else {
                            const uint32_t at = atomicAdd(&s_tn[w], 1u);
                            s_tl[w][at] = (uint16_t)i;
                        }
                    }
                }
                __syncthreads();
                // the window's coverage of each path at this lane's pixel
                for (uint32_t j = 0; j < ns; ++j) s_A[j][tid] += s_Lw[j][wid];
                const uint32_t nt = s_tn[wid];
                uint32_t k = 0;
                for (; k + 2 <= nt; k += 2) {
                    const uint32_t ia = s_tl[wid][k], ib = s_tl[wid][k + 1];
                    const float4 pa = s_pc[ia], pb = s_pc[ib];
                    const uint32_t ja = s_ev[ia] & J_MASK, jb = s_ev[ib] & J_MASK;
                    float da, db;
                    const float ca = pa.z * mean_clamp(Xf + 1.0f - pa.x, Xf + 1.0f - pa.y, da, db);
                    const float cb = pb.z * mean_clamp(Xf + 1.0f - pb.x, Xf + 1.0f - pb.y, da, db);
                    s_A[ja][tid] += ca;
                    s_A[jb][tid] += cb;
                }
                if (k < nt) {
                    const uint32_t i = s_tl[wid][k];
                    const float4 pc = s_pc[i];
                    float da, db;
                    s_A[s_ev[i] & J_MASK][tid] += pc.z * mean_clamp(Xf + 1.0f - pc.x, Xf + 1.0f - pc.y, da, db);
                }
            }
        }
        __syncthreads();  // s_gc cleared before the walk's atomics
        // the canvas under each of the segment's paths, composited once from
        // the checkpoint (the same operations, in the same order, as a
        // composite from the checkpoint for each path)
        float4 Cbs[KSEG];
        {
            float4 C = ck[(uint64_t)seg * ck_stride + slot];
#pragma unroll
            for (uint32_t i = 0; i < (uint32_t)KSEG; ++i) {
                Cbs[i] = C;
                if (i + 1 < ns) {
                    float d;
                    C = over(s_col[i], alpha_of(s_A[i][tid], d), C);
                }
            }
        }
        for (uint32_t jl = ns; jl-- > 0;) {
            float4 Cb = Cbs[0];  // canvas before path j
#pragma unroll
            for (uint32_t k = 1; k < (uint32_t)KSEG; ++k)
                if (k == jl) Cb = Cbs[k];
            const float A = s_A[jl][tid];
            float dadA;
            const float al = alpha_of(A, dadA);
            const float4 cc = s_col[jl];
            const float dal = g.x * (cc.x - cc.w * Cb.x) + g.y * (cc.y - cc.w * Cb.y) + g.z * (cc.z - cc.w * Cb.z) +
                              g.w * (cc.w - cc.w * Cb.w);
            const float gdotC = g.x * Cb.x + g.y * Cb.y + g.z * Cb.z + g.w * Cb.w;
            if (__ballot(al != 0.0f)) {
                float r0 = g.x * al, r1 = g.y * al, r2 = g.z * al;
                float r3 = g.w * al - al * gdotC;
                wave_sum3(r0, r1, r2);
                r3 = wave_sum(r3);
                if (lane == 0) {
                    atomicAdd(&s_gc[jl][0], r0);
                    atomicAdd(&s_gc[jl][1], r1);
                    atomicAdd(&s_gc[jl][2], r2);
                    atomicAdd(&s_gc[jl][3], r3);
                }
            }
            const float keep = 1.0f - cc.w * al;
            g = make_float4(g.x * keep, g.y * keep, g.z * keep, g.w * keep);
            const float dA = active ? dal * dadA : 0.0f;
            // s_cnt and s_pb are the same for the whole block, so the window
            // decisions (and the barriers in them) are too
            const uint32_t cnt = s_cnt[jl], pb = s_pb[jl];
            if (!cnt) continue;
            const bool wave_on = __ballot(dA != 0.0f) != 0;
            const float Tw = wave_sum(dA);  // the wavefront's dA total
            // the path's pieces [pb, pb + cnt), in windows from the top
            for (uint32_t e_hi = pb + cnt; e_hi > pb;) {
                const uint32_t e_lo = e_hi - pb > ECAP ? e_hi - ECAP : pb;
                if (!(e_lo >= w_lo && e_hi <= w_lo + w_n)) {
                    // the walk leaves the staged window: its pieces' gradients, then
                    // a window ending at e_hi (pieces of this path and the ones before)
                    flush_pieces(w_n);
                    const uint32_t lo = e_hi > ECAP ? e_hi - ECAP : 0u;
                    stage_pieces(S, k0 + j0, ns, Y, lo, e_hi - lo, sg, bad);
                    __syncthreads();
                    w_lo = lo;
                    w_n = e_hi - lo;
                }
                // a piece's sums: S1 = Σ dA·M, S2 = Σ dA·∂M/∂u_a, S3 = Σ dA·∂M/∂u_b
                // over the wavefront
                auto piece_sums = [&](uint32_t i, float4 pc, uint32_t ev) {
                    if (!(ev & LIVE_BIT)) return;
                    const float lo = fminf(pc.x, pc.y), hi = fmaxf(pc.x, pc.y);
                    if (xw0 + 64.0f <= lo) return;  // M = 0 on every pixel of the wavefront
                    if (xw0 >= hi) {               // M = 1, dM = 0
                        if (lane == 0) atomicAdd(&s_S[i].x, Tw);
                        return;
                    }
                    float mua, mub;
                    const float M = mean_clamp(Xf + 1.0f - pc.x, Xf + 1.0f - pc.y, mua, mub);
                    float S1 = dA * M, S2 = dA * mua, S3 = dA * mub;
                    wave_sum3(S1, S2, S3);
                    if (lane == 0) {
                        atomicAdd(&s_S[i].x, S1);
                        atomicAdd(&s_S[i].y, S2);
                        atomicAdd(&s_S[i].z, S3);
                    }
                };
                // four pieces' loads at a time
                uint32_t e = e_lo;
                for (; wave_on && e + 4 <= e_hi; e += 4) {
                    const uint32_t i = e - w_lo;
                    const float4 p0 = s_pc[i], p1 = s_pc[i + 1], p2 = s_pc[i + 2], p3 = s_pc[i + 3];
                    const uint32_t v0 = s_ev[i], v1 = s_ev[i + 1], v2 = s_ev[i + 2], v3 = s_ev[i + 3];
                    piece_sums(i, p0, v0);
                    piece_sums(i + 1, p1, v1);
                    piece_sums(i + 2, p2, v2);
                    piece_sums(i + 3, p3, v3);
                }
                for (; wave_on && e < e_hi; ++e) piece_sums(e - w_lo, s_pc[e - w_lo], s_ev[e - w_lo]);
                e_hi = e_lo;
            }
        }
        flush_pieces(w_n);  // the segment's last window (and its barrier before the colours)
        if (tid < (int)ns * 4) {
            const float a = s_gc[tid >> 2][tid & 3];
            if (a != 0.0f) atomicAdd(&gcol[4 * (uint64_t)(k0 + j0 + (tid >> 2)) + (tid & 3)], a);
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
    uint32_t max_path_vtx = 0;  // most vertices (edges) of one path
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
    k_cells<<<(uint32_t)std::min<size_t>(std::max<size_t>(d->n_paths, 1), 65535), BLOCK, 0, s>>>(S, d->d_status);
    if (d->max_path_vtx > CELL_EDGES) {
        k_fill_cells<<<grid_of((size_t)d->n_vtx * FILL_LANES), BLOCK, 0, s>>>(S, d->d_status);
        k_sort_cells<<<grid_of(d->n_cells), BLOCK, 0, s>>>(S, d->d_status);
    }
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
    for (uint32_t k = 0; k < L->n_paths; ++k) d->max_path_vtx = std::max(d->max_path_vtx, vtx_off[k + 1] - vtx_off[k]);
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
