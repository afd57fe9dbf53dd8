// ft_pipeline.hip -- gfx950 kernels and launch sequence of the fill pipeline.
//
// The reference (footile 0.7.0) rasterises one figure on one CPU thread:
// recursive De Casteljau flattening (src/plotter.rs:226-332), a (y, x) vertex
// sort with an active-edge table scanned row by row (src/fig.rs:480-617), and
// a per-row cumulative sum of i16 signed-area cells (src/imgbuf.rs,
// src/fig.rs:619-682).  Here every stage is data-parallel:
//
//  * each curve is flattened by one lane with an iterative, stack-free
//    subdivision walk that re-derives a node's control points from the root,
//    so the emitted points (and their order) are those of the recursion;
//  * Fig::add_point / Fig::close sequencing (coincident-point drop, closing
//    pop, the 65,535-point cap) is restated as flags + prefix sums;
//  * the active-edge scan becomes independent (edge, row) items: an edge's
//    x at row r is x_bot0 + (r - r0) * inv_slope (Fixed adds are integer
//    adds), so every row's cell deltas are computed in closed form;
//  * cell deltas are accumulated in LDS with i32 atomics (integer adds commute,
//    so the sums are exact; i16 wrap is applied at the row scan), then a
//    wavefront scan per row maps coverage and writes the raster once.
#include "ft_device.h"
#include "ft_pipeline.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace ft {

constexpr int BLOCK = 256;
constexpr uint32_t GRID_MAX = 2048;     // grid-stride kernels: 8 blocks per CU
constexpr uint32_t LDS_CELLS = 8192;    // i32 signed-area cells per raster block
constexpr uint32_t BAND_ROWS_MAX = 64;
constexpr uint32_t CELL_ALIGN = 16;     // row stride of cells in LDS
constexpr uint32_t CLOSE_BIT = 0x80000000u;
constexpr uint32_t NO_IDX = 0xFFFFFFFFu;
constexpr unsigned long long NO_KEY = ~0ull;
constexpr uint32_t SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = BLOCK * SCAN_ITEMS;
constexpr uint32_t SCAN_SLOTS = 4;      // look-back state instances per run
constexpr uint32_t SPIN_LIMIT = 1u << 24;

enum : uint32_t { TAG_CLOSE = 0, TAG_MOVE, TAG_LINE, TAG_QUAD, TAG_CUBIC, TAG_PEN_WIDTH };
enum : uint32_t { FMT_MATTE8 = 0, FMT_RGBA8P = 1, FMT_GRAYA8P = 2 };

__device__ inline void raise_err(DevStatus *st, uint32_t code) { atomicCAS(&st->err, 0u, code); }
__device__ inline bool run_failed(const DevStatus *st) {
    return __hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
           __hip_atomic_load(&st->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

// ---------------------------------------------------------------------------
// Flattening (src/plotter.rs:163-332, src/geom.rs:9-35)
// ---------------------------------------------------------------------------
struct WP {
    float x, y, w;
};

// WidePt::midpoint: position midpoint (a+b)/2 per axis, width (w0+w1)/2.
__device__ inline WP wmid(WP a, WP b) {
    WP r;
    r.x = (a.x + b.x) / 2.0f;
    r.y = (a.y + b.y) / 2.0f;
    r.w = (a.w + b.w) / 2.0f;
    return r;
}

// Plotter::is_within_tolerance2: distance_sq <= tol_sq (plotter.rs:273-276)
__device__ inline bool within_tol(WP a, WP b, float tol_sq) {
    float dx = a.x - b.x;
    float dy = a.y - b.y;
    return dx * dx + dy * dy <= tol_sq;
}

// Plotter::transform_point (plotter.rs:169-172): width is not transformed.
__device__ inline WP xform(const PathDev &P, float x, float y, float w) {
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
__device__ inline bool quad_split(const Quad &q, float tol_sq, Quad &L, Quad &R) {
    WP ab = wmid(q.a, q.b);
    WP bc = wmid(q.b, q.c);
    WP ab_bc = wmid(ab, bc);
    WP ac = wmid(q.a, q.c);
    L.a = q.a; L.b = ab; L.c = ab_bc;
    R.a = ab_bc; R.b = bc; R.c = q.c;
    return within_tol(ab_bc, ac, tol_sq);
}

// One subdivision step of cubic_to_tran (plotter.rs:319-326).
__device__ inline bool cubic_split(const Cubic &q, float tol_sq, Cubic &L, Cubic &R) {
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

__device__ inline WP end_of(const Quad &q) { return q.c; }
__device__ inline WP end_of(const Cubic &q) { return q.d; }

// Depth-first walk of the subdivision tree in the recursion's order (left
// subtree, then right).  `bits` holds the left/right choices from the root;
// moving to a right sibling replays those choices from the root, so every
// node is computed by exactly the float operations the recursion performs.
// Returns 0, or FT_ERANGE when a node at FT_FLATTEN_MAX_DEPTH is still out of
// tolerance or the curve yields more than FT_FLATTEN_MAX_LEAVES points.
template <class Curve, class Emit>
__device__ uint32_t flatten_walk(const Curve &root, float tol_sq, Emit &&emit) {
    Curve cur = root;
    uint32_t depth = 0, bits = 0, leaves = 0;
    for (;;) {
        Curve L, R;
        bool leaf;
        if constexpr (std::is_same_v<Curve, Quad>)
            leaf = quad_split(cur, tol_sq, L, R);
        else
            leaf = cubic_split(cur, tol_sq, L, R);
        if (leaf) {
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
                if constexpr (std::is_same_v<Curve, Quad>)
                    quad_split(cur, tol_sq, l2, r2);
                else
                    cubic_split(cur, tol_sq, l2, r2);
                cur = ((bits >> lv) & 1u) ? r2 : l2;
            }
        } else {
            if (depth >= FT_FLATTEN_MAX_DEPTH) return FT_ERANGE;
            cur = L;
            bits <<= 1;
            ++depth;
        }
    }
}

// Index of the op that last set the pen before op i (plotter.rs:128-130,
// 164-166): PenWidth ops do not move the pen.  Returns < begin at path start.
__device__ inline int64_t pen_event_before(const ft_path_op *ops, uint32_t i, uint32_t begin) {
    int64_t j = (int64_t)i - 1;
    while (j >= (int64_t)begin && ops[j].tag == TAG_PEN_WIDTH) --j;
    return j;
}

// s_width in effect just before op k: the last PenWidth op of the path before
// k, else the width the fill started with (plotter.rs:53,151-153).
__device__ inline float width_before(const ft_path_op *ops, const uint32_t *lastpw, int64_t k,
                                     const PathDev &P) {
    if (!lastpw || k < (int64_t)P.op_begin) return P.s_width0;
    uint32_t l = lastpw[k];
    if (l == 0 || l - 1 < P.op_begin) return P.s_width0;
    return ops[l - 1].v[0];
}

// Pen (untransformed WidePt) before op i.
__device__ inline WP pen_before(const ft_path_op *ops, const uint32_t *lastpw, uint32_t i,
                                const PathDev &P) {
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

__device__ inline Quad quad_of(const ft_path_op &o, WP pen, float sw, const PathDev &P) {
    Quad q;  // plotter.rs:233-240
    q.a = xform(P, pen.x, pen.y, pen.w);
    q.b = xform(P, o.v[0], o.v[1], (pen.w + sw) / 2.0f);
    q.c = xform(P, o.v[2], o.v[3], sw);
    return q;
}

__device__ inline Cubic cubic_of(const ft_path_op &o, WP pen, float sw, const PathDev &P) {
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
// Stage 0: clear per-run state.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_clear(PathState *ps, uint32_t n_paths, uint32_t *band_cnt,
                                                 uint32_t *band_cur, uint32_t n_bands, uint32_t *flags,
                                                 uint32_t n_flags, DevStatus *st) {
    uint32_t n = n_paths > n_bands ? n_paths : n_bands;
    n = n > n_flags ? n : n_flags;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        if (i < n_paths) {
            PathState s;
            s.topkey = NO_KEY;
            s.topidx = NO_IDX;
            s.offered = 0;
            s.repop = 0;
            s.pt_base = 0;
            s.n_edges = 0;
            s.stored = 0;
            ps[i] = s;
        }
        if (i < n_bands) {
            band_cnt[i] = 0;
            band_cur[i] = 0;
        }
        if (i < n_flags) flags[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < sizeof(DevStatus) / 4)
        reinterpret_cast<uint32_t *>(st)[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Single-pass prefix scan with decoupled look-back.  Tiles are taken in
// ticket order, so a tile only waits on tiles already owned by running blocks,
// each of which publishes its aggregate before it waits on anything.
// ---------------------------------------------------------------------------
template <bool IS_MAX>
__device__ inline unsigned long long scan_op(unsigned long long a, unsigned long long b) {
    if constexpr (IS_MAX)
        return a > b ? a : b;
    else
        return a + b;
}

template <bool IS_MAX>
__device__ inline unsigned long long wave_incl(unsigned long long v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned long long t = __shfl_up(v, o, 64);
        if (lane >= o) v = scan_op<IS_MAX>(v, t);
    }
    return v;
}

template <typename T, bool IS_MAX>
__global__ __launch_bounds__(BLOCK) void k_scan(const T *in, T *out, const uint32_t *n_dev, uint32_t n_host,
                                                uint32_t n_cap, uint32_t *flags, unsigned long long *agg,
                                                unsigned long long *incl, uint32_t *ticket, T *total,
                                                DevStatus *st, bool check_failed) {
    __shared__ uint32_t s_tile;
    __shared__ unsigned long long s_wave[BLOCK / 64];
    __shared__ unsigned long long s_prefix;
    if (check_failed && run_failed(st)) return;
    uint32_t n = n_dev ? *n_dev : n_host;
    if (n > n_cap) n = n_cap;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        const uint64_t base = (uint64_t)tile * SCAN_TILE;
        if (base >= n) break;
        unsigned long long v[SCAN_ITEMS];
        unsigned long long run = 0;
        const uint64_t mine = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
        for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
            unsigned long long x = (mine + j < n) ? (unsigned long long)in[mine + j] : 0ull;
            v[j] = run;  // exclusive within the thread
            run = scan_op<IS_MAX>(run, x);
        }
        unsigned long long winc = wave_incl<IS_MAX>(run);
        unsigned long long wexc = __shfl_up(winc, 1, 64);
        if (lane == 0) wexc = 0;
        if (lane == 63) s_wave[wid] = winc;
        __syncthreads();
        unsigned long long wpre = 0, btot = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) {
            if (w < wid) wpre = scan_op<IS_MAX>(wpre, s_wave[w]);
            btot = scan_op<IS_MAX>(btot, s_wave[w]);
        }
        if (threadIdx.x == 0) {
            unsigned long long prefix = 0;
            if (tile == 0) {
                __hip_atomic_store(&incl[0], btot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&flags[0], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_store(&agg[tile], btot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&flags[tile], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                int64_t j = (int64_t)tile - 1;
                uint32_t spins = 0;
                while (j >= 0) {
                    uint32_t f = __hip_atomic_load(&flags[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                    if (f == 0u) {
                        if (++spins > SPIN_LIMIT) {  // never expected: report, do not hang
                            atomicOr(&st->spin_fail, 1u);
                            raise_err(st, FT_EDEVICE);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    if (f == 2u) {
                        prefix = scan_op<IS_MAX>(
                            __hip_atomic_load(&incl[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), prefix);
                        break;
                    }
                    prefix = scan_op<IS_MAX>(
                        __hip_atomic_load(&agg[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), prefix);
                    --j;
                }
                __hip_atomic_store(&incl[tile], scan_op<IS_MAX>(prefix, btot), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&flags[tile], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            s_prefix = prefix;
            if (base + SCAN_TILE >= n && total) *total = (T)scan_op<IS_MAX>(prefix, btot);
        }
        __syncthreads();
        const unsigned long long pre = scan_op<IS_MAX>(s_prefix, scan_op<IS_MAX>(wpre, wexc));
#pragma unroll
        for (uint32_t j = 0; j < SCAN_ITEMS; ++j)
            if (mine + j < n) out[mine + j] = (T)scan_op<IS_MAX>(pre, v[j]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Stage 1: per-op event counts (plotter.rs:175-224 + the final fig.close()).
// An op yields: Close 1 close; Move 1 close + 1 point; Line 1 point; Quad and
// Cubic one point per subdivision leaf; PenWidth nothing.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_pen_width_mark(const ft_path_op *ops, uint32_t n_ops, uint32_t *op_pw) {
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n_ops; i += gridDim.x * BLOCK)
        op_pw[i] = ops[i].tag == TAG_PEN_WIDTH ? i + 1 : 0u;
}

struct CountLeaf {
    __device__ void operator()(WP) const {}
};

__global__ __launch_bounds__(BLOCK) void k_count(const ft_path_op *ops, const uint32_t *op_path,
                                                 const PathDev *paths, uint32_t n_ops, uint32_t mode,
                                                 const uint32_t *lastpw, unsigned long long *op_cnt,
                                                 DevStatus *st) {
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n_ops; i += gridDim.x * BLOCK) {
        uint32_t ne = 0, nc = 0;
        const uint32_t p = op_path[i];
        if (p != NO_IDX) {
            const PathDev &P = paths[p];
            const ft_path_op o = ops[i];
            switch (o.tag) {
            case TAG_CLOSE: ne = 1; nc = 1; break;
            case TAG_MOVE: ne = 2; nc = 1; break;
            case TAG_LINE: ne = 1; break;
            case TAG_QUAD:
            case TAG_CUBIC: {
                WP pen = pen_before(ops, lastpw, i, P);
                float sw = width_before(ops, lastpw, i, P);
                uint32_t leaves = 0;
                auto cnt = [&](WP) { ++leaves; };
                uint32_t rc = o.tag == TAG_QUAD ? flatten_walk(quad_of(o, pen, sw, P), P.tol_sq, cnt)
                                                : flatten_walk(cubic_of(o, pen, sw, P), P.tol_sq, cnt);
                if (rc) raise_err(st, rc);
                else ne = leaves;
                break;
            }
            case TAG_PEN_WIDTH: break;
            default: raise_err(st, FT_EINVAL); break;
            }
            if (mode == MODE_FILL && i + 1 == P.op_begin + P.op_count) {
                ++ne;  // Plotter::fill closes the figure (plotter.rs:347)
                ++nc;
            }
        }
        op_cnt[i] = ((unsigned long long)ne << 32) | nc;
    }
}

// ---------------------------------------------------------------------------
// Stage 2: emit events.  Fill mode writes Fixed points (f32 -> Fixed as
// Fig::add_point does, fig.rs:435-436) and the run (sub-figure) id of every
// event; stroke mode writes (x, y, width, kind) for the host stroker.
// ---------------------------------------------------------------------------
struct Emitter {
    uint32_t mode;
    uint64_t e;
    uint32_t run, p;
    int2 *ev_xy;
    float4 *ev_f;
    uint32_t *ev_meta;
    int32_t *run_first, *run_last;
    uint32_t *run_path, *run_start, *run_end;
    __device__ void point(WP q) {
        if (mode == MODE_FILL) {
            ev_xy[e] = make_int2(fx_from_f32(q.x), fx_from_f32(q.y));
            ev_meta[e] = run;
        } else {
            ev_f[e] = make_float4(q.x, q.y, q.w, 0.0f);
        }
        ++e;
    }
    __device__ void close(bool joined) {
        if (mode == MODE_FILL) {
            ev_meta[e] = run | CLOSE_BIT;
            run_first[run] = -1;
            run_last[run] = -1;
            run_path[run] = p;
            run_start[run] = NO_IDX;
            run_end[run] = 0;
        } else {
            ev_f[e] = make_float4(__builtin_nanf(""), __builtin_nanf(""), joined ? 1.0f : 0.0f,
                                  joined ? 2.0f : 1.0f);
        }
        ++run;
        ++e;
    }
};

__global__ __launch_bounds__(BLOCK) void k_emit(const ft_path_op *ops, const uint32_t *op_path,
                                                const PathDev *paths, uint32_t n_ops, uint32_t mode,
                                                const uint32_t *lastpw, const unsigned long long *op_off,
                                                uint64_t cap_events, Emitter em_base, DevStatus *st) {
    if (run_failed(st)) return;
    const unsigned long long tot = st->op_total;
    const uint32_t n_events = (uint32_t)(tot >> 32);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->n_events = n_events;
        st->n_closes = (uint32_t)tot;
        if (n_events > cap_events) {
            atomicOr(&st->overflow, 1u);
            st->need_events = n_events;
        }
    }
    if (n_events > cap_events) return;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n_ops; i += gridDim.x * BLOCK) {
        const uint32_t p = op_path[i];
        if (p == NO_IDX) continue;
        const PathDev &P = paths[p];
        const ft_path_op o = ops[i];
        Emitter em = em_base;
        em.e = op_off[i] >> 32;
        em.run = (uint32_t)op_off[i];
        em.p = p;
        switch (o.tag) {
        case TAG_CLOSE: em.close(true); break;  // plotter.rs:200-203
        case TAG_MOVE: {                         // plotter.rs:208-214
            em.close(false);
            em.point(xform(P, o.v[0], o.v[1], width_before(ops, lastpw, i, P)));
            break;
        }
        case TAG_LINE:  // plotter.rs:219-224
            em.point(xform(P, o.v[0], o.v[1], width_before(ops, lastpw, i, P)));
            break;
        case TAG_QUAD:
        case TAG_CUBIC: {
            WP pen = pen_before(ops, lastpw, i, P);
            float sw = width_before(ops, lastpw, i, P);
            auto put = [&](WP q) { em.point(q); };
            if (o.tag == TAG_QUAD)
                flatten_walk(quad_of(o, pen, sw, P), P.tol_sq, put);
            else
                flatten_walk(cubic_of(o, pen, sw, P), P.tol_sq, put);
            break;
        }
        default: break;
        }
        if (mode == MODE_FILL && i + 1 == P.op_begin + P.op_count) em.close(true);
    }
}

// ---------------------------------------------------------------------------
// Stage 3: Fig::add_point / Fig::close semantics (fig.rs:373-383,428-451).
// Within a run (events between closes) a point is pushed if it is the run's
// first or differs in Fixed from the previous point event; closing pops the
// last pushed point when it equals the run's first point.
// ---------------------------------------------------------------------------
__device__ inline bool int2_eq(int2 a, int2 b) { return a.x == b.x && a.y == b.y; }

__device__ inline bool is_pushed(const uint32_t *meta, const int2 *xy, uint32_t e) {
    if (e == 0 || (meta[e - 1] & CLOSE_BIT)) return true;
    return !int2_eq(xy[e], xy[e - 1]);
}

__global__ __launch_bounds__(BLOCK) void k_ingest(const uint32_t *ev_meta, const int2 *ev_xy, int32_t *run_first,
                                                  int32_t *run_last, const uint32_t *run_path, PathState *ps,
                                                  DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n = st->n_events;
    const int lane = threadIdx.x & 63;
    for (uint32_t b = blockIdx.x * BLOCK; b < n; b += gridDim.x * BLOCK) {
        const uint32_t e = b + threadIdx.x;
        bool is_pt = false;
        uint32_t p = NO_IDX;
        if (e < n) {
            const uint32_t m = ev_meta[e];
            if (!(m & CLOSE_BIT)) {
                is_pt = true;
                const bool first = e == 0 || (ev_meta[e - 1] & CLOSE_BIT);
                if (first) run_first[m] = (int32_t)e;
                if (first || !int2_eq(ev_xy[e], ev_xy[e - 1])) atomicMax(&run_last[m], (int32_t)e);
                p = run_path[m];
            }
        }
        // offered points per path, aggregated per wavefront
        const unsigned long long hm = __ballot(is_pt);
        if (hm) {
            const int fl = __ffsll((long long)hm) - 1;
            const uint32_t lead = __shfl(p, fl, 64);
            const unsigned long long same = __ballot(is_pt && p == lead);
            if (lane == fl) atomicAdd(&ps[lead].offered, (uint32_t)__popcll(same));
            if (is_pt && p != lead) atomicAdd(&ps[p].offered, 1u);
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_kept(const uint32_t *ev_meta, const int2 *ev_xy, const int32_t *run_first,
                                                const int32_t *run_last, uint32_t *ev_kept, DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n = st->n_events;
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < n; e += gridDim.x * BLOCK) {
        const uint32_t m = ev_meta[e];
        uint32_t kept = 0;
        if (!(m & CLOSE_BIT) && is_pushed(ev_meta, ev_xy, e)) {
            const int32_t rf = run_first[m], rl = run_last[m];
            const bool pop = int2_eq(ev_xy[rf], ev_xy[rl]);
            kept = (pop && (int32_t)e == rl) ? 0u : 1u;
        }
        ev_kept[e] = kept;
    }
}

// Local point index of kept event e in its path, or NO_IDX when the 65,535
// cap drops it.  `repop` tells whether the cap's re-pop removed local 65534.
__device__ inline uint32_t capped_local(uint32_t idx, uint32_t base, uint32_t repop) {
    const uint32_t local = idx - base;
    if (local >= FT_MAX_POINTS) return NO_IDX;
    if (local == FT_MAX_POINTS - 1 && repop) return NO_IDX;
    return local;
}

__device__ inline uint32_t path_event_begin(const PathDev &P, const unsigned long long *op_off) {
    return (uint32_t)(op_off[P.op_begin] >> 32);
}

__device__ inline uint32_t path_event_end(const PathDev &P, const unsigned long long *op_off, uint32_t n_ops,
                                          uint32_t n_events) {
    const uint32_t k = P.op_begin + P.op_count;
    return k < n_ops ? (uint32_t)(op_off[k] >> 32) : n_events;
}

__device__ inline unsigned long long point_key(int2 q) {
    return ((unsigned long long)((uint32_t)q.y ^ 0x80000000u) << 32) | (unsigned long long)((uint32_t)q.x ^ 0x80000000u);
}

// Stage 4: compaction into the per-path point arrays (the Fig's `points`),
// with the 65,535-point cap (vid.rs:17, fig.rs:429-431).  Once a path's
// stored count reaches the cap every later point is ignored; the run holding
// the cap point then closes against its truncated last point, which may pop
// local point 65534.  Every later run only adds a point and pops it again.
__global__ __launch_bounds__(BLOCK) void k_compact(const uint32_t *ev_meta, const int2 *ev_xy, const uint32_t *ev_kept,
                                                   const uint32_t *ev_idx, const int32_t *run_first,
                                                   const uint32_t *run_path, uint32_t *run_start,
                                                   uint32_t *run_end, const PathDev *paths,
                                                   const unsigned long long *op_off, uint32_t n_ops, int2 *pts,
                                                   uint32_t *pt_run, PathState *ps, DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n = st->n_events;
    const int lane = threadIdx.x & 63;
    for (uint32_t b = blockIdx.x * BLOCK; b < n; b += gridDim.x * BLOCK) {
        const uint32_t e = b + threadIdx.x;
        bool has = false;
        uint32_t p = NO_IDX;
        unsigned long long key = NO_KEY;
        if (e < n) {
            const uint32_t m = ev_meta[e];
            const uint32_t run = m & ~CLOSE_BIT;
            p = run_path[run];
            const PathDev &P = paths[p];
            const uint32_t eb = path_event_begin(P, op_off);
            const uint32_t base = ev_idx[eb];
            if (e == eb) ps[p].pt_base = base;
            if (ev_kept[e]) {
                const uint32_t idx = ev_idx[e];
                uint32_t local = idx - base;
                bool keep = local < FT_MAX_POINTS;
                if (local == FT_MAX_POINTS - 1) {
                    // Find the next pushed point of the path: if it lies in the
                    // same run, the cap hit this run and its close compares the
                    // run's first point with this one (fig.rs:373-383).
                    const uint32_t ee = path_event_end(P, op_off, n_ops, n);
                    for (uint32_t e2 = e + 1; e2 < ee; ++e2) {
                        if (ev_meta[e2] & CLOSE_BIT) continue;
                        if (!is_pushed(ev_meta, ev_xy, e2)) continue;
                        if (ev_meta[e2] == run && int2_eq(ev_xy[run_first[run]], ev_xy[e])) {
                            keep = false;
                            ps[p].repop = 1;
                        }
                        break;
                    }
                }
                if (keep) {
                    const int2 q = ev_xy[e];
                    pts[idx] = q;
                    pt_run[idx] = run;
                    if ((int32_t)e == run_first[run]) run_start[run] = idx;
                    atomicMax(&run_end[run], idx + 1);
                    has = true;
                    key = point_key(q);
                }
            }
        }
        // top-left point per path: min over (y, x), reduced per wavefront
        const unsigned long long hm = __ballot(has);
        if (hm) {
            const int fl = __ffsll((long long)hm) - 1;
            const uint32_t lead = __shfl(p, fl, 64);
            unsigned long long k = (has && p == lead) ? key : NO_KEY;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                unsigned long long t = __shfl_xor(k, o, 64);
                k = t < k ? t : k;
            }
            if (lane == fl) atomicMin(&ps[lead].topkey, k);
            if (has && p != lead) atomicMin(&ps[p].topkey, key);
        }
    }
}

// Scan rows of an edge mapped to raster rows (fig.rs:496-500,536-553): scan
// row y_row = raster row + min(top_row, 0); rows run from max(top_row, 0) to
// the raster bottom.  Returns false when the edge touches no raster row.
__device__ inline bool edge_rows(const EdgeRec &E, int32_t top_row, uint32_t H, int32_t &rr0, int32_t &rr1) {
    const int64_t first = top_row > 0 ? top_row : 0;
    if (first >= (int64_t)H) return false;
    const int64_t shift = top_row < 0 ? top_row : 0;
    int64_t a = (int64_t)fx_int(E.y_upper) - shift;
    int64_t b = (int64_t)fx_int(E.y_lower) - shift;
    if (a < first) a = first;
    if (b > (int64_t)H - 1) b = (int64_t)H - 1;
    if (a > b) return false;
    rr0 = (int32_t)a;
    rr1 = (int32_t)b;
    return true;
}

__device__ inline int32_t top_row_of(unsigned long long key) {
    return ((int32_t)((uint32_t)(key >> 32) ^ 0x80000000u)) >> 16;
}

// Add `v` to a per-path counter for every lane with `has`, one atomic per
// distinct path in the wavefront for the common single-path case.
__device__ inline void wave_count(uint32_t *base, size_t stride_words, bool has, uint32_t p, uint32_t v) {
    const int lane = threadIdx.x & 63;
    const unsigned long long hm = __ballot(has);
    if (!hm) return;
    const int fl = __ffsll((long long)hm) - 1;
    const uint32_t lead = __shfl(p, fl, 64);
    const bool same = has && p == lead;
    uint32_t t = same ? v : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == fl && t) atomicAdd(base + (size_t)lead * stride_words, t);
    if (has && !same && v) atomicAdd(base + (size_t)p * stride_words, v);
}

constexpr size_t PS_WORDS = sizeof(PathState) / 4;

// Stage 5: one edge per stored point: the segment to its successor within the
// sub-figure (SubFig::next, fig.rs:143-163).  Edge::new (fig.rs:179-210) runs
// once per edge; the edge's rows are binned into row bands.

__global__ __launch_bounds__(BLOCK) void k_edges(const uint32_t *ev_meta, const uint32_t *ev_kept,
                                                 const uint32_t *ev_idx, const uint32_t *run_path,
                                                 const uint32_t *run_start, const uint32_t *run_end,
                                                 const PathDev *paths, const int2 *pts, EdgeRec *edges,
                                                 uint32_t *band_cnt, PathState *ps, DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n = st->n_events;
    for (uint32_t b = blockIdx.x * BLOCK; b < n; b += gridDim.x * BLOCK) {
        const uint32_t e = b + threadIdx.x;
        bool valid = false, has_edge = false;
        uint32_t p = NO_IDX;
        if (e < n && ev_kept[e]) {
            const uint32_t run = ev_meta[e];
            p = run_path[run];
            const uint32_t i = ev_idx[e];
            const uint32_t local = capped_local(i, ps[p].pt_base, ps[p].repop);
            if (local != NO_IDX) {
                valid = true;
                const int2 q = pts[i];
                if (point_key(q) == ps[p].topkey) atomicMin(&ps[p].topidx, local);
                const uint32_t nx = (i + 1 < run_end[run]) ? i + 1 : run_start[run];
                EdgeRec E;
                E.flags = 0;
                E.y_upper = E.y_lower = E.x_bot0 = E.inv = E.step = 0;
                if (nx != i) {
                    const int2 q1 = pts[nx];
                    if (q1.y != q.y) {
                        const bool fwd = q1.y > q.y;  // Forward: descends in vertex order
                        const int2 up = fwd ? q : q1, lo = fwd ? q1 : q;
                        const int32_t dx = wsub(lo.x, up.x), dy = wsub(lo.y, up.y);
                        int32_t step = 0;
                        if (dx != 0) {  // Edge::calculate_step
                            step = fx_abs(fx_div(dy, dx));
                            if (step > FX_ONE) step = FX_ONE;
                        }
                        E.inv = fx_div(dx, dy);
                        E.step = step;
                        E.y_upper = up.y;
                        E.y_lower = lo.y;
                        E.x_bot0 = wadd(up.x, fx_mul(E.inv, wsub(fx_floor(wadd(up.y, FX_ONE)), up.y)));
                        E.flags = 1u | (fwd ? 2u : 0u);
                        has_edge = true;
                        const PathDev &P = paths[p];
                        int32_t r0, r1;
                        if (edge_rows(E, top_row_of(ps[p].topkey), P.height, r0, r1)) {
                            const uint32_t b0 = (uint32_t)r0 / P.band_rows, b1 = (uint32_t)r1 / P.band_rows;
                            for (uint32_t bb = b0; bb <= b1; ++bb) atomicAdd(&band_cnt[P.band_base + bb], 1u);
                        }
                    }
                }
                edges[i] = E;
            }
        }
        wave_count(&ps[0].stored, PS_WORDS, valid, p, 1u);
        wave_count(&ps[0].n_edges, PS_WORDS, has_edge, p, 1u);
    }
}

__global__ __launch_bounds__(BLOCK) void k_bin(const uint32_t *ev_meta, const uint32_t *ev_kept,
                                               const uint32_t *ev_idx, const uint32_t *run_path,
                                               const PathDev *paths, const EdgeRec *edges,
                                               const uint32_t *band_off, uint32_t *band_cur, uint32_t *band_list,
                                               uint64_t cap_pairs, const PathState *ps, DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n_pairs = st->n_pairs;
    if (n_pairs > cap_pairs) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(&st->overflow, 2u);
            st->need_pairs = n_pairs;
        }
        return;
    }
    const uint32_t n = st->n_events;
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < n; e += gridDim.x * BLOCK) {
        if (!ev_kept[e]) continue;
        const uint32_t run = ev_meta[e];
        const uint32_t p = run_path[run];
        const uint32_t i = ev_idx[e];
        if (capped_local(i, ps[p].pt_base, ps[p].repop) == NO_IDX) continue;
        const EdgeRec E = edges[i];
        if (!(E.flags & 1u)) continue;
        const PathDev &P = paths[p];
        int32_t r0, r1;
        if (!edge_rows(E, top_row_of(ps[p].topkey), P.height, r0, r1)) continue;
        const uint32_t b0 = (uint32_t)r0 / P.band_rows, b1 = (uint32_t)r1 / P.band_rows;
        for (uint32_t bb = b0; bb <= b1; ++bb) {
            const uint32_t gb = P.band_base + bb;
            band_list[band_off[gb] + atomicAdd(&band_cur[gb], 1u)] = i;
        }
    }
}

// ---------------------------------------------------------------------------
// Stage 6: raster.  One block per row band.
// ---------------------------------------------------------------------------

// Signed-area cell deltas of one edge on one scan row: the closed form of
// Edge::{starting_cov, continuing_cov, calculate_x_limits_*, scan_area,
// first_cov, step_cov} (fig.rs:238-321).  Cells left of x = 0 land in cell 0;
// cells at x >= W are dropped (fig.rs:292-295).
__device__ inline void edge_row_cells(const EdgeRec &E, int32_t sr, int32_t ed, int32_t W, int32_t *cells) {
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
    const int32_t full = cov << 8;  // Fixed::from(cov as f32 / 256.0), exact
    const int32_t r = (min_pix == max_pix) ? fx_mul(wsub(FX_ONE, fx_fract(fx_avg(max_x, min_x))), full)
                                           : fx_mul(wsub(FX_ONE, fx_fract(min_x)), FX_HALF);
    const int64_t xc0 = E.step > 0 ? fx_mul(r, E.step) : r;
    const int64_t st = E.step > 0 ? E.step : FX_ONE;
    // x_cov after k steps: min(xc0 + k * st, ONE); x_pix_k = min(pixel_cov, cov)
    auto xp_at = [&](int64_t k) -> int32_t {
        int64_t xc = xc0 + k * st;
        if (xc > FX_ONE) xc = FX_ONE;
        int32_t xp = pixel_cov((int32_t)xc);
        return xp < cov ? xp : cov;
    };
    int32_t prev = 0;
    int64_t k = 0;
    int32_t x = min_pix;
    if (min_pix < 0) {
        // every step at x <= 0 adds to cell 0: their sum is x_pix at x = 0
        const int64_t k0 = -(int64_t)min_pix;
        const int32_t xp = xp_at(k0 < 65537 ? k0 : 65537);
        if (xp) atomicAdd(&cells[0], xp * ed);
        if (xp >= cov) return;
        prev = xp;
        k = k0 + 1;
        x = 1;
    }
    for (; x < W; ++x, ++k) {
        const int32_t xp = xp_at(k);
        const int32_t d = xp - prev;
        if (d) atomicAdd(&cells[x], d * ed);
        if (xp >= cov) break;
        prev = xp;
    }
}

__device__ inline uint32_t wave_incl_u32(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// SrcOver with coverage alpha a for one premultiplied pixel of N channels
// (pix Pixel::composite_channels_alpha; alpha is the last channel).
template <int N>
__device__ inline void composite_px(uint8_t *d, const uint32_t *sw, uint32_t a) {
    const uint32_t aw = ch8_widen(a);
    const uint32_t sa = ch8_mul_w(sw[N - 1], aw);
    const uint32_t invw = ch8_widen(255u - sa);
#pragma unroll
    for (int c = 0; c < N; ++c) d[c] = (uint8_t)sat_add_u8(ch8_mul_w(sw[c], aw), ch8_mul_w(ch8_widen(d[c]), invw));
}

template <int N>
__device__ inline void composite_px_fresh(uint8_t *d, const uint32_t *sw, uint32_t a) {
    const uint32_t aw = ch8_widen(a);
#pragma unroll
    for (int c = 0; c < N; ++c) d[c] = (uint8_t)ch8_mul_w(sw[c], aw);  // over zero: d*x = 0
}

// Prefix-sum one raster row of cells (one wavefront), map coverage with the
// fill rule and store / composite CPT pixels per lane.
template <int FMT, int CPT>
__device__ void row_pass(const int32_t *cells, uint32_t W, uint8_t *row, uint32_t rule, uint32_t clr,
                         bool fresh) {
    constexpr int BPP = FMT == FMT_RGBA8P ? 4 : (FMT == FMT_GRAYA8P ? 2 : 1);
    const int lane = threadIdx.x & 63;
    uint32_t sw[4];
    if (FMT == FMT_RGBA8P) {
        for (int c = 0; c < 4; ++c) sw[c] = ch8_widen((clr >> (8 * c)) & 0xFFu);
    } else if (FMT == FMT_GRAYA8P) {
        sw[0] = ch8_widen(clr & 0xFFu);
        sw[1] = ch8_widen((clr >> 8) & 0xFFu);
    }
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < W; c0 += 64u * CPT) {
        const uint32_t x = c0 + (uint32_t)lane * CPT;
        uint32_t v[CPT];
        uint32_t run = 0;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            run += (x + j < W) ? (uint32_t)cells[x + j] : 0u;
            v[j] = run;
        }
        const uint32_t incl = wave_incl_u32(run);
        const uint32_t total = __shfl(incl, 63, 64);
        const uint32_t base = carry + incl - run;
        carry += total;
        if (x >= W) continue;
        uint8_t a[CPT];
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int32_t s = (int32_t)(base + v[j]);
            a[j] = (uint8_t)(rule ? coverage_even_odd(s) : coverage_non_zero(s));
        }
        uint8_t *dst = row + (size_t)x * BPP;
        const bool full = x + CPT <= W && (((uintptr_t)dst) & (CPT * BPP >= 16 ? 15 : CPT * BPP - 1)) == 0;
        if (FMT == FMT_MATTE8) {
            if constexpr (CPT == 16) {
                if (full) {
                    uint4 pk;
                    pk.x = a[0] | (a[1] << 8) | (a[2] << 16) | ((uint32_t)a[3] << 24);
                    pk.y = a[4] | (a[5] << 8) | (a[6] << 16) | ((uint32_t)a[7] << 24);
                    pk.z = a[8] | (a[9] << 8) | (a[10] << 16) | ((uint32_t)a[11] << 24);
                    pk.w = a[12] | (a[13] << 8) | (a[14] << 16) | ((uint32_t)a[15] << 24);
                    *reinterpret_cast<uint4 *>(dst) = pk;
                    continue;
                }
            } else if constexpr (CPT == 4) {
                if (full) {
                    *reinterpret_cast<uint32_t *>(dst) = a[0] | (a[1] << 8) | (a[2] << 16) | ((uint32_t)a[3] << 24);
                    continue;
                }
            }
            for (int j = 0; j < CPT; ++j)
                if (x + j < W) dst[j] = a[j];
        } else {
            if (full && CPT * BPP == 16) {
                uint4 px = fresh ? make_uint4(0, 0, 0, 0) : *reinterpret_cast<const uint4 *>(dst);
                uint8_t *pb = reinterpret_cast<uint8_t *>(&px);
#pragma unroll
                for (int j = 0; j < CPT; ++j) {
                    if (fresh)
                        composite_px_fresh<BPP>(pb + j * BPP, sw, a[j]);
                    else
                        composite_px<BPP>(pb + j * BPP, sw, a[j]);
                }
                *reinterpret_cast<uint4 *>(dst) = px;
            } else {
                for (int j = 0; j < CPT; ++j) {
                    if (x + j >= W) break;
                    if (fresh)
                        composite_px_fresh<BPP>(dst + j * BPP, sw, a[j]);
                    else
                        composite_px<BPP>(dst + j * BPP, sw, a[j]);
                }
            }
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_raster(const PathDev *paths, const uint32_t *band_path,
                                                  const uint32_t *band_off, const uint32_t *band_cnt,
                                                  const uint32_t *band_list, const EdgeRec *edges,
                                                  const int2 *pts, const uint32_t *pt_run, const uint32_t *run_start,
                                                  const uint32_t *run_end, const PathState *ps, DevStatus *st) {
    extern __shared__ int32_t s_cells[];
    __shared__ int s_fwd;
    if (run_failed(st)) return;
    const uint32_t gb = blockIdx.x;
    const uint32_t p = band_path[gb];
    const PathDev &P = paths[p];
    const uint32_t R = P.band_rows, H = P.height, W = P.width;
    const uint32_t y0 = (gb - P.band_base) * R;
    const uint32_t y1 = y0 + R < H ? y0 + R : H;
    const unsigned long long topkey = ps[p].topkey;
    const bool has_pts = topkey != NO_KEY && ps[p].topidx != NO_IDX;
    const int32_t top_row = has_pts ? top_row_of(topkey) : 0;
    const int64_t first = !has_pts ? (int64_t)H : (top_row > 0 ? top_row : 0);
    const uint32_t bpp = P.fmt == FMT_RGBA8P ? 4 : (P.fmt == FMT_GRAYA8P ? 2 : 1);
    uint8_t *const base = reinterpret_cast<uint8_t *>(P.dst);
    if (P.flags & PATH_FRESH) {  // rows above the figure stay as a fresh raster: zero
        const uint32_t ze = first < (int64_t)y1 ? (uint32_t)first : y1;
        for (uint32_t r = y0; r < ze; ++r) {
            uint8_t *row = base + (size_t)r * P.pitch;
            for (uint32_t x = threadIdx.x; x < W * bpp; x += BLOCK) row[x] = 0;
        }
    }
    const uint32_t ry0 = first > (int64_t)y0 ? (uint32_t)first : y0;
    if ((int64_t)ry0 >= (int64_t)y1) return;
    const uint32_t Wp = (W + CELL_ALIGN - 1) / CELL_ALIGN * CELL_ALIGN;
    const uint32_t rows = y1 - ry0;
    if (threadIdx.x == 0) {
        // get_dir at the top-left vertex (fig.rs:402-411,495)
        const uint32_t i = ps[p].pt_base + ps[p].topidx;
        const uint32_t rn = pt_run[i], rs = run_start[rn], re = run_end[rn];
        const uint32_t nf = (i + 1 < re) ? i + 1 : rs;
        const uint32_t nr = (i > rs) ? i - 1 : re - 1;
        const int2 q = pts[i], qf = pts[nf], qr = pts[nr];
        s_fwd = widdershins(wsub(qr.x, q.x), wsub(qr.y, q.y), wsub(qf.x, q.x), wsub(qf.y, q.y)) ? 1 : 0;
    }
    for (uint32_t c = threadIdx.x; c < rows * Wp; c += BLOCK) s_cells[c] = 0;
    __syncthreads();
    const int fwd = s_fwd;
    const int32_t shift = top_row < 0 ? top_row : 0;
    const uint32_t off = band_off[gb], cnt = band_cnt[gb];
    for (uint32_t k = threadIdx.x; k < cnt; k += BLOCK) {
        const EdgeRec E = edges[band_list[off + k]];
        const int32_t ed = ((int)((E.flags >> 1) & 1u) == fwd) ? 1 : -1;
        int64_t a = (int64_t)fx_int(E.y_upper) - shift, b = (int64_t)fx_int(E.y_lower) - shift;
        if (a < ry0) a = ry0;
        if (b > (int64_t)y1 - 1) b = (int64_t)y1 - 1;
        for (int64_t r = a; r <= b; ++r)
            edge_row_cells(E, (int32_t)(r + shift), ed, (int32_t)W, s_cells + (size_t)(r - ry0) * Wp);
    }
    __syncthreads();
    const int wid = threadIdx.x >> 6;
    const bool fresh = (P.flags & PATH_FRESH) != 0;
    for (uint32_t r = wid; r < rows; r += BLOCK / 64) {
        const int32_t *cells = s_cells + (size_t)r * Wp;
        uint8_t *row = base + (size_t)(ry0 + r) * P.pitch;
        switch (P.fmt) {
        case FMT_MATTE8:
            if (W >= 1024)
                row_pass<FMT_MATTE8, 16>(cells, W, row, P.rule, P.clr, fresh);
            else
                row_pass<FMT_MATTE8, 4>(cells, W, row, P.rule, P.clr, fresh);
            break;
        case FMT_RGBA8P: row_pass<FMT_RGBA8P, 4>(cells, W, row, P.rule, P.clr, fresh); break;
        default: row_pass<FMT_GRAYA8P, 8>(cells, W, row, P.rule, P.clr, fresh); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define FT_HIP(call)                                   \
    do {                                               \
        if ((call) != hipSuccess) return FT_EDEVICE;   \
    } while (0)

template <typename T>
static ft_status dev_grow(T *&ptr, size_t &cap, size_t need) {
    if (need <= cap && ptr) return FT_OK;
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    size_t n = need < 1 ? 1 : need;
    if (hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)) != hipSuccess) {
        cap = 0;
        return FT_ENOMEM;
    }
    cap = n;
    return FT_OK;
}

template <typename T>
static ft_status dev_alloc(T *&ptr, size_t n) {
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    if (hipMalloc(reinterpret_cast<void **>(&ptr), (n ? n : 1) * sizeof(T)) != hipSuccess) return FT_ENOMEM;
    return FT_OK;
}

void band_geometry(uint32_t width, uint32_t height, uint32_t *rows, uint32_t *n_bands, uint32_t *cells) {
    const uint32_t wp = (width + CELL_ALIGN - 1) / CELL_ALIGN * CELL_ALIGN;
    uint32_t r = wp ? LDS_CELLS / wp : BAND_ROWS_MAX;
    if (r > BAND_ROWS_MAX) r = BAND_ROWS_MAX;
    if (r < 1) r = 1;
    if (height && r > height) r = height;
    *rows = r;
    *n_bands = height ? (height + r - 1) / r : 0;
    *cells = r * wp;
}

ft_status ws_init(Workspace &ws) {
    FT_HIP(hipGetDevice(&ws.device));
    FT_HIP(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking));
    FT_HIP(hipMalloc(reinterpret_cast<void **>(&ws.d_status), sizeof(DevStatus)));
    FT_HIP(hipHostMalloc(reinterpret_cast<void **>(&ws.h_status), sizeof(DevStatus), hipHostMallocDefault));
    memset(ws.h_status, 0, sizeof(DevStatus));
    return FT_OK;
}

void ws_free(Workspace &ws) {
    void *bufs[] = {ws.d_ops, ws.d_op_path, ws.d_paths, ws.d_band_path, ws.d_op_cnt, ws.d_op_off, ws.d_op_pw,
                    ws.d_op_lastpw, ws.d_ev_xy, ws.d_ev_f, ws.d_ev_meta, ws.d_ev_kept, ws.d_ev_idx,
                    ws.d_run_first, ws.d_run_last, ws.d_run_path, ws.d_run_start, ws.d_run_end, ws.d_pts,
                    ws.d_pt_run, ws.d_edges, ws.d_band_cnt, ws.d_band_off, ws.d_band_cur, ws.d_band_list,
                    ws.d_scan_flags, ws.d_scan_agg, ws.d_scan_incl, ws.d_pstate, ws.d_status};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (ws.h_status) (void)hipHostFree(ws.h_status);
    if (ws.ev) {
        for (int i = 0; i < 4 * Workspace::TIMING_RING; ++i)
            if (ws.ev[i]) (void)hipEventDestroy(ws.ev[i]);
        delete[] ws.ev;
    }
    if (ws.stream) (void)hipStreamDestroy(ws.stream);
    ws = Workspace();
}

static ft_status ensure_event_capacity(Workspace &ws, size_t need) {
    if (need <= ws.cap_events && ws.d_ev_meta) return FT_OK;
    size_t cap = need < 4096 ? 4096 : need;
    ft_status s;
    if ((s = dev_alloc(ws.d_ev_xy, cap)) || (s = dev_alloc(ws.d_ev_f, cap)) || (s = dev_alloc(ws.d_ev_meta, cap)) ||
        (s = dev_alloc(ws.d_ev_kept, cap)) || (s = dev_alloc(ws.d_ev_idx, cap)) ||
        (s = dev_alloc(ws.d_run_first, cap)) || (s = dev_alloc(ws.d_run_last, cap)) ||
        (s = dev_alloc(ws.d_run_path, cap)) || (s = dev_alloc(ws.d_run_start, cap)) ||
        (s = dev_alloc(ws.d_run_end, cap)) || (s = dev_alloc(ws.d_pts, cap)) || (s = dev_alloc(ws.d_pt_run, cap)) ||
        (s = dev_alloc(ws.d_edges, cap))) {
        ws.cap_events = 0;
        return s;
    }
    ws.cap_events = cap;
    return FT_OK;
}

static ft_status ensure_scan_capacity(Workspace &ws) {
    size_t n = ws.cap_events;
    if (ws.n_ops > n) n = ws.n_ops;
    if (ws.n_bands > n) n = ws.n_bands;
    size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE + 1;
    if (tiles <= ws.cap_tiles && ws.d_scan_flags) return FT_OK;
    ft_status s;
    if ((s = dev_alloc(ws.d_scan_flags, tiles * SCAN_SLOTS)) || (s = dev_alloc(ws.d_scan_agg, tiles * SCAN_SLOTS)) ||
        (s = dev_alloc(ws.d_scan_incl, tiles * SCAN_SLOTS)))
        return s;
    ws.cap_tiles = tiles;
    return FT_OK;
}

ft_status ws_upload_ops(Workspace &ws, const ft_path_op *ops, size_t n_ops) {
    if (n_ops != ws.n_ops) return FT_EINVAL;
    if (n_ops) FT_HIP(hipMemcpyAsync(ws.d_ops, ops, n_ops * sizeof(ft_path_op), hipMemcpyHostToDevice, ws.stream));
    return FT_OK;
}

ft_status ws_set_batch(Workspace &ws, const ft_path_op *ops, size_t n_ops, const ft_fill_desc *paths,
                       size_t n_paths, uint32_t extra_flags) {
    if (n_ops >= 0x7FFFFFFFu || n_paths >= 0x7FFFFFFFu) return FT_EINVAL;
    std::vector<uint32_t> op_path(n_ops, NO_IDX);
    std::vector<PathDev> pd(n_paths);
    std::vector<uint32_t> band_path;
    uint32_t lds = 0;
    for (size_t k = 0; k < n_paths; ++k) {
        const ft_fill_desc &d = paths[k];
        if (d.op_begin > n_ops || d.op_count > n_ops - d.op_begin) return FT_EINVAL;
        if (d.fmt > FT_GRAYA8P || d.rule > FT_EVEN_ODD) return FT_EINVAL;
        if (d.width > FT_MAX_WIDTH) return FT_ENOTSUP;
        const uint32_t bpp = d.fmt == FT_RGBA8P ? 4 : (d.fmt == FT_GRAYA8P ? 2 : 1);
        if ((uint64_t)d.width * d.height && (!d.dst || d.pitch < (uint64_t)d.width * bpp)) return FT_EINVAL;
        for (uint64_t i = d.op_begin; i < d.op_begin + d.op_count; ++i) {
            if (op_path[i] != NO_IDX) return FT_EINVAL;  // overlapping op ranges
            if (ops[i].tag > FT_OP_PEN_WIDTH) return FT_EINVAL;
            op_path[i] = (uint32_t)k;
        }
        PathDev &P = pd[k];
        P.op_begin = (uint32_t)d.op_begin;
        P.op_count = d.op_count;
        P.rule = d.rule;
        P.fmt = d.fmt;
        P.dst = d.dst;
        P.width = d.width;
        P.height = d.height;
        P.pitch = d.pitch;
        P.clr = (uint32_t)d.clr[0] | ((uint32_t)d.clr[1] << 8) | ((uint32_t)d.clr[2] << 16) |
                ((uint32_t)d.clr[3] << 24);
        memcpy(P.tr, d.transform, sizeof P.tr);
        float tol = d.tolerance > 0.01f ? d.tolerance : 0.01f;  // set_tolerance, plotter.rs:133-137
        P.tol_sq = tol * tol;
        P.s_width0 = d.pen_width;
        uint32_t rows, nb, cells;
        band_geometry(d.width, d.height, &rows, &nb, &cells);
        if (!d.width) nb = 0;
        P.band_rows = rows;
        P.band_base = (uint32_t)band_path.size();
        P.n_bands = nb;
        P.flags = (d.fresh ? PATH_FRESH : 0u) | extra_flags;
        if (nb && cells > lds) lds = cells;
        band_path.insert(band_path.end(), nb, (uint32_t)k);
    }
    ft_status s;
    const size_t old_ops = ws.cap_ops;
    if ((s = dev_grow(ws.d_ops, ws.cap_ops, n_ops))) return s;
    if (ws.cap_ops != old_ops || !ws.d_op_cnt) {  // per-op scratch follows the op capacity
        if ((s = dev_alloc(ws.d_op_path, ws.cap_ops)) || (s = dev_alloc(ws.d_op_cnt, ws.cap_ops)) ||
            (s = dev_alloc(ws.d_op_off, ws.cap_ops)) || (s = dev_alloc(ws.d_op_pw, ws.cap_ops)) ||
            (s = dev_alloc(ws.d_op_lastpw, ws.cap_ops)))
            return s;
    }
    const size_t old_paths = ws.cap_paths;
    if ((s = dev_grow(ws.d_paths, ws.cap_paths, n_paths))) return s;
    if (ws.cap_paths != old_paths || !ws.d_pstate)
        if ((s = dev_alloc(ws.d_pstate, ws.cap_paths))) return s;
    size_t nb_total = band_path.size();
    if (nb_total > ws.cap_bands || !ws.d_band_path) {
        size_t cap = nb_total ? nb_total : 1;
        if ((s = dev_alloc(ws.d_band_path, cap)) || (s = dev_alloc(ws.d_band_cnt, cap)) ||
            (s = dev_alloc(ws.d_band_off, cap)) || (s = dev_alloc(ws.d_band_cur, cap)))
            return s;
        ws.cap_bands = cap;
    }
    ws.n_ops = n_ops;
    ws.n_paths = n_paths;
    ws.n_bands = nb_total;
    ws.lds_cells = lds;
    ws.h_pstate.resize(n_paths);
    if (!ws.d_ev_meta) {
        if ((s = ensure_event_capacity(ws, 16 * n_ops + 1024))) return s;
    }
    if (!ws.d_band_list) {
        size_t cap = 4 * (16 * n_ops + 1024);
        if ((s = dev_alloc(ws.d_band_list, cap))) return s;
        ws.cap_pairs = cap;
    }
    if ((s = ensure_scan_capacity(ws))) return s;
    if (n_ops) {
        FT_HIP(hipMemcpyAsync(ws.d_ops, ops, n_ops * sizeof(ft_path_op), hipMemcpyHostToDevice, ws.stream));
        FT_HIP(hipMemcpyAsync(ws.d_op_path, op_path.data(), n_ops * 4, hipMemcpyHostToDevice, ws.stream));
    }
    if (n_paths)
        FT_HIP(hipMemcpyAsync(ws.d_paths, pd.data(), n_paths * sizeof(PathDev), hipMemcpyHostToDevice, ws.stream));
    if (nb_total)
        FT_HIP(hipMemcpyAsync(ws.d_band_path, band_path.data(), nb_total * 4, hipMemcpyHostToDevice, ws.stream));
    FT_HIP(hipStreamSynchronize(ws.stream));  // host vectors go out of scope
    return FT_OK;
}

static inline uint32_t grid_for(size_t n) {
    size_t g = (n + BLOCK - 1) / BLOCK;
    if (g < 1) g = 1;
    return (uint32_t)(g > GRID_MAX ? GRID_MAX : g);
}

static inline uint32_t scan_grid(size_t n) {
    size_t g = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (g < 1) g = 1;
    return (uint32_t)(g > 1024 ? 1024 : g);
}

ft_status ws_launch(Workspace &ws, uint32_t mode) {
    hipStream_t s = ws.stream;
    const uint32_t n_ops = (uint32_t)ws.n_ops;
    const size_t tiles = ws.cap_tiles;
    uint32_t *fl = ws.d_scan_flags;
    unsigned long long *ag = ws.d_scan_agg, *in = ws.d_scan_incl;
    DevStatus *st = ws.d_status;
    const uint32_t n_flags = (uint32_t)(tiles * SCAN_SLOTS);
    hipEvent_t *ev = nullptr;
    if (ws.timing && mode == MODE_FILL && ws.n_bands && ws.n_pending < Workspace::TIMING_RING) {
        if (!ws.ev) {
            ws.ev = new hipEvent_t[4 * Workspace::TIMING_RING]();
            for (int i = 0; i < 4 * Workspace::TIMING_RING; ++i) FT_HIP(hipEventCreate(&ws.ev[i]));
        }
        ev = ws.ev + 4 * ws.n_pending++;
    }
    if (ev) FT_HIP(hipEventRecord(ev[0], s));
    k_clear<<<grid_for(std::max(std::max(ws.n_paths, ws.n_bands), (size_t)n_flags)), BLOCK, 0, s>>>(
        ws.d_pstate, (uint32_t)ws.n_paths, ws.d_band_cnt, ws.d_band_cur, (uint32_t)ws.n_bands, fl, n_flags, st);
    const uint32_t *lastpw = nullptr;
    if (mode != MODE_FILL && n_ops) {
        k_pen_width_mark<<<grid_for(n_ops), BLOCK, 0, s>>>(ws.d_ops, n_ops, ws.d_op_pw);
        k_scan<uint32_t, true><<<scan_grid(n_ops), BLOCK, 0, s>>>(ws.d_op_pw, ws.d_op_lastpw, nullptr, n_ops, n_ops,
                                                                  fl + 3 * tiles, ag + 3 * tiles, in + 3 * tiles,
                                                                  &st->tickets[3], nullptr, st, false);
        lastpw = ws.d_op_lastpw;
    }
    if (n_ops) {
        k_count<<<grid_for(n_ops), BLOCK, 0, s>>>(ws.d_ops, ws.d_op_path, ws.d_paths, n_ops, mode, lastpw, ws.d_op_cnt,
                                                  st);
        k_scan<unsigned long long, false><<<scan_grid(n_ops), BLOCK, 0, s>>>(
            ws.d_op_cnt, ws.d_op_off, nullptr, n_ops, n_ops, fl, ag, in, &st->tickets[0], &st->op_total, st, false);
        Emitter em{};
        em.mode = mode;
        em.ev_xy = ws.d_ev_xy;
        em.ev_f = ws.d_ev_f;
        em.ev_meta = ws.d_ev_meta;
        em.run_first = ws.d_run_first;
        em.run_last = ws.d_run_last;
        em.run_path = ws.d_run_path;
        em.run_start = ws.d_run_start;
        em.run_end = ws.d_run_end;
        k_emit<<<grid_for(n_ops), BLOCK, 0, s>>>(ws.d_ops, ws.d_op_path, ws.d_paths, n_ops, mode, lastpw, ws.d_op_off,
                                                 ws.cap_events, em, st);
    }
    if (mode == MODE_FILL) {
        const uint32_t ge = grid_for(ws.cap_events);
        k_ingest<<<ge, BLOCK, 0, s>>>(ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_last, ws.d_run_path,
                                      ws.d_pstate, st);
        k_kept<<<ge, BLOCK, 0, s>>>(ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_last, ws.d_ev_kept, st);
        k_scan<uint32_t, false><<<scan_grid(ws.cap_events), BLOCK, 0, s>>>(
            ws.d_ev_kept, ws.d_ev_idx, &st->n_events, 0, (uint32_t)ws.cap_events, fl + tiles, ag + tiles,
            in + tiles, &st->tickets[1], &st->n_kept, st, true);
        k_compact<<<ge, BLOCK, 0, s>>>(ws.d_ev_meta, ws.d_ev_xy, ws.d_ev_kept, ws.d_ev_idx, ws.d_run_first,
                                       ws.d_run_path, ws.d_run_start, ws.d_run_end, ws.d_paths, ws.d_op_off, n_ops,
                                       ws.d_pts, ws.d_pt_run, ws.d_pstate, st);
        k_edges<<<ge, BLOCK, 0, s>>>(ws.d_ev_meta, ws.d_ev_kept, ws.d_ev_idx, ws.d_run_path, ws.d_run_start,
                                     ws.d_run_end, ws.d_paths, ws.d_pts, ws.d_edges, ws.d_band_cnt, ws.d_pstate, st);
        if (ws.n_bands) {
            k_scan<uint32_t, false><<<scan_grid(ws.n_bands), BLOCK, 0, s>>>(
                ws.d_band_cnt, ws.d_band_off, nullptr, (uint32_t)ws.n_bands, (uint32_t)ws.n_bands, fl + 2 * tiles,
                ag + 2 * tiles, in + 2 * tiles, &st->tickets[2], &st->n_pairs, st, true);
            k_bin<<<ge, BLOCK, 0, s>>>(ws.d_ev_meta, ws.d_ev_kept, ws.d_ev_idx, ws.d_run_path, ws.d_paths, ws.d_edges,
                                       ws.d_band_off, ws.d_band_cur, ws.d_band_list, ws.cap_pairs, ws.d_pstate, st);
            if (ev) FT_HIP(hipEventRecord(ev[1], s));
            k_raster<<<(uint32_t)ws.n_bands, BLOCK, ws.lds_cells * sizeof(int32_t), s>>>(
                ws.d_paths, ws.d_band_path, ws.d_band_off, ws.d_band_cnt, ws.d_band_list, ws.d_edges, ws.d_pts,
                ws.d_pt_run, ws.d_run_start, ws.d_run_end, ws.d_pstate, st);
            if (ev) FT_HIP(hipEventRecord(ev[2], s));
        }
    }
    if (ev) FT_HIP(hipEventRecord(ev[3], s));
    FT_HIP(hipGetLastError());
    FT_HIP(hipMemcpyAsync(ws.h_status, ws.d_status, sizeof(DevStatus), hipMemcpyDeviceToHost, s));
    return FT_OK;
}

ft_status ws_finish(Workspace &ws, bool read_pstate) {
    if (read_pstate && ws.n_paths)
        FT_HIP(hipMemcpyAsync(ws.h_pstate.data(), ws.d_pstate, ws.n_paths * sizeof(PathState),
                              hipMemcpyDeviceToHost, ws.stream));
    FT_HIP(hipStreamSynchronize(ws.stream));
    for (int k = 0; k < ws.n_pending; ++k) {
        float total = 0.0f, raster = 0.0f;
        hipEvent_t *ev = ws.ev + 4 * k;
        FT_HIP(hipEventElapsedTime(&total, ev[0], ev[3]));
        FT_HIP(hipEventElapsedTime(&raster, ev[1], ev[2]));
        ws.t_total_ms += total;
        ws.t_raster_ms += raster;
        ws.n_timed += 1;
    }
    ws.n_pending = 0;
    const DevStatus &h = *ws.h_status;
    if (h.err) return (ft_status)h.err;
    if (h.overflow) return (ft_status)0x100;  // internal: capacity exceeded
    return FT_OK;
}

ft_status ws_run(Workspace &ws, uint32_t mode, bool read_pstate) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        ft_status s = ws_launch(ws, mode);
        if (s) return s;
        s = ws_finish(ws, read_pstate);
        if (s != (ft_status)0x100) return s;
        const DevStatus &h = *ws.h_status;
        if (h.overflow & 1u) {
            if ((s = ensure_event_capacity(ws, (size_t)h.need_events + h.need_events / 4 + 1024))) return s;
            if ((s = ensure_scan_capacity(ws))) return s;
        }
        if (h.overflow & 2u) {
            size_t cap = (size_t)h.need_pairs + h.need_pairs / 4 + 1024;
            if ((s = dev_alloc(ws.d_band_list, cap))) return s;
            ws.cap_pairs = cap;
        }
    }
    return FT_ENOMEM;
}

}  // namespace ft
