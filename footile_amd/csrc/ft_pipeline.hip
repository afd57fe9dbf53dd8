// ft_pipeline.hip -- gfx950 kernels and launch sequence of the fill pipeline.
//
// The reference (footile 0.7.0) rasterises one figure on one CPU thread:
// recursive De Casteljau flattening (src/plotter.rs:226-332), a (y, x) vertex
// sort with an active-edge table scanned row by row (src/fig.rs:480-617), and
// a per-row cumulative sum of i16 signed-area cells (src/imgbuf.rs,
// src/fig.rs:619-682).  Here every stage is data-parallel:
//
//  * each curve is flattened by OP_LANES lanes, one per subtree a few levels
//    below the root, with an iterative, stack-free subdivision walk that
//    re-derives a node's control points from the root, so the emitted points
//    (and their order) are those of the recursion;
//  * Fig::add_point / Fig::close sequencing (coincident-point drop, closing
//    pop, the 65,535-point cap) is restated as flags + prefix sums;
//  * the active-edge scan becomes independent (edge, row) items: an edge's
//    x at row r is x_bot0 + (r - r0) * inv_slope (Fixed adds are integer
//    adds), so every (edge, row) item is set up independently;
//  * cell deltas are accumulated in LDS with i32 atomics (integer adds commute,
//    so the sums are exact; i16 wrap is applied at the row scan), then a
//    wavefront scan per row maps coverage and writes the raster once.
#include "ft_pipeline.h"
#include "ft_scan.h"
#include "ft_stages.h"
#include "ft_stroke_dev.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstring>
#include <mutex>

namespace ft {

// Block-wide keyed reduction with one atomic (flush) per distinct key per
// block for the keys of the first two leaders of each wavefront; any other
// lane flushes alone.  Keys are path indices of consecutive items, so a block
// holds few.  Every thread of the block must call it.
template <class T, class Comb, class Flush>
__device__ inline void block_reduce_keyed(bool has, uint32_t key, T v, T identity, Comb comb, Flush flush) {
    constexpr int NW = BLOCK / 64;
    __shared__ uint32_t s_key[2 * NW];
    __shared__ T s_val[2 * NW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const unsigned long long hm = __ballot(has);
        const int fl = hm ? __ffsll((long long)hm) - 1 : 0;
        const uint32_t lead = __shfl(key, fl, 64);
        const bool same = has && key == lead;
        T m = same ? v : identity;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = comb(m, (T)__shfl_xor(m, o, 64));
        if (lane == 0) {
            s_key[2 * wid + it] = hm ? lead : NO_IDX;
            s_val[2 * wid + it] = m;
        }
        has = has && !same;
    }
    if (has) flush(key, v);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ck = NO_IDX;
        T cv = identity;
        for (int k = 0; k < 2 * NW; ++k) {
            const uint32_t kk = s_key[k];
            if (kk == NO_IDX) continue;
            if (kk == ck) {
                cv = comb(cv, s_val[k]);
            } else {
                if (ck != NO_IDX) flush(ck, cv);
                ck = kk;
                cv = s_val[k];
            }
        }
        if (ck != NO_IDX) flush(ck, cv);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Stage 0: clear per-run state (a kernel, not hipMemsetAsync: graph-safe).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_clear(PathState *ps, unsigned long long *top_rep, uint32_t n_paths,
                                                 uint32_t *band_cnt, uint32_t *band_cur, uint32_t n_slots,
                                                 unsigned long long *lb, uint32_t n_lb, DevStatus *st,
                                                 uint32_t keep_failed) {
    uint32_t n = n_paths * TOP_REPS > n_slots ? n_paths * TOP_REPS : n_slots;
    n = n > n_lb ? n : n_lb;
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
            s.fwd = 0;
            ps[i] = s;
        }
        if (i < n_paths * TOP_REPS) top_rep[i] = NO_KEY;
        if (i < n_slots) {
            band_cnt[i] = 0;
            band_cur[i] = 0;
        }
        if (i < n_lb) lb[i] = 0;
    }
    // The status (one wavefront's words): with keep_failed, a failed earlier
    // launch whose status the host has not read yet stays failed, so this
    // launch's kernels skip their work and the failure reaches the next read.
    static_assert(sizeof(DevStatus) / 4 <= 64, "the status is one wavefront's words");
    if (blockIdx.x == 0 && threadIdx.x < sizeof(DevStatus) / 4) {
        const bool failed = (keep_failed != 0u) & ((st->err | st->overflow) != 0u);
        if (!failed) reinterpret_cast<uint32_t *>(st)[threadIdx.x] = 0;
    }
}

// ---------------------------------------------------------------------------
// Stages 1-2: ops -> events (count, scan, emit).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_pen_width_mark(const ft_path_op *ops, uint32_t n_ops, uint32_t *op_pw) {
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n_ops; i += gridDim.x * BLOCK)
        op_pw[i] = ops[i].tag == TAG_PEN_WIDTH ? i + 1 : 0u;
}

constexpr unsigned long long SUB_CACHED = 1ull << 30, SUB_LAST_CLOSE = 1ull << 31;

// One thread per (op, lane); an op's OP_LANES lanes are adjacent lanes of one
// wavefront.  Writes the op's event/close count and each lane's offset within
// the op (a prefix over the op's lanes).  In fill runs a curve lane also
// keeps its first LEAF_CACHE points in leaf[k * stride + t] (coalesced
// across the lanes).  The offset's low word carries the lane's facts for
// k_emit above the op's close count (at most 2, bits 0-15): its leaf count
// (bits 16-29), SUB_CACHED (its points are all in the cache) and
// SUB_LAST_CLOSE (it emits the path's final close).
__global__ __launch_bounds__(BLOCK) void k_count(const ft_path_op *ops, const uint32_t *op_path,
                                                 const PathDev *paths, uint32_t n_ops, uint32_t mode,
                                                 const uint32_t *lastpw, const uint32_t *order,
                                                 unsigned long long *op_cnt, unsigned long long *sub_off, I2 *leaf,
                                                 size_t stride, DevStatus *st) {
    const uint64_t n_items = (uint64_t)n_ops * OP_LANES;
    for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t < n_items; t += (uint64_t)gridDim.x * BLOCK) {
        const uint32_t oi = (uint32_t)(t >> FLAT_SPLIT), s = (uint32_t)t & (OP_LANES - 1);
        const uint32_t i = order ? order[oi] : oi;  // item t is lane s of op i
        if (i >= n_ops) {  // a corrupt order table
            raise_err(st, FT_EDEVICE);
            continue;
        }
        const OpCount c = count_op(ops, op_path, paths, i, s, mode, lastpw, leaf ? leaf + t : nullptr, stride);
        const unsigned long long v = pack_count(c);
        unsigned long long inc = v;
        uint32_t leaves = c.leaves;
#pragma unroll
        for (uint32_t o = 1; o < OP_LANES; o <<= 1) {
            const unsigned long long u = __shfl_up(inc, o, OP_LANES);
            if (s >= o) inc += u;
            leaves += __shfl_xor(leaves, o, OP_LANES);
        }
        const unsigned long long lc = c.leaves < 0x3FFFu ? c.leaves : 0x3FFFu;
        const bool cached = leaf && (c.flags & OPC_CURVE) && !c.err && c.leaves <= LEAF_CACHE;
        sub_off[t] = (inc - v) | (lc << 16) | (cached ? SUB_CACHED : 0ull) |
                     ((c.flags & OPC_LAST_CLOSE) ? SUB_LAST_CLOSE : 0ull);
        if (s == OP_LANES - 1) op_cnt[i] = inc;
        if (c.err) raise_err(st, c.err);
        if (s == 0 && leaves > FT_FLATTEN_MAX_LEAVES) raise_err(st, FT_ERANGE);
    }
}

__global__ __launch_bounds__(BLOCK) void k_emit(const ft_path_op *ops, const uint32_t *op_path,
                                                const PathDev *paths, uint32_t n_ops, uint32_t mode,
                                                const uint32_t *lastpw, const uint32_t *order,
                                                const unsigned long long *op_off, const unsigned long long *sub_off,
                                                const I2 *leaf, size_t stride, EventSink sink, DevStatus *st) {
    if (run_failed(st)) return;
    const unsigned long long tot = st->op_total;
    const uint32_t n_events = (uint32_t)(tot >> 32);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->n_events = n_events;
        st->n_closes = (uint32_t)tot;
        if (n_events > sink.cap_events) {
            atomicOr(&st->overflow, 1u);
            st->need_events = n_events;
        }
    }
    if (n_events > sink.cap_events) return;
    uint32_t bad = 0;
    const uint64_t n_items = (uint64_t)n_ops * OP_LANES;
    for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t < n_items; t += (uint64_t)gridDim.x * BLOCK) {
        const uint32_t oi = (uint32_t)(t >> FLAT_SPLIT), s = (uint32_t)t & (OP_LANES - 1);
        const uint32_t i = order ? order[oi] : oi;
        if (i >= n_ops) {
            bad = 1;
            continue;
        }
        const unsigned long long so = sub_off[t];
        const uint32_t cached = (uint32_t)(so >> 16) & 0x3FFFu;
        const unsigned long long off = op_off[i] + (so & ~(0xFFFFull << 16));
        if (so & SUB_CACHED) {  // a curve lane whose points are all cached: no op or path loads
            EventSink em = sink;
            em.e = off >> 32;
            em.run = (uint32_t)off;
            I2 q[LEAF_CACHE];
#pragma unroll
            for (uint32_t k = 0; k < LEAF_CACHE; ++k)
                if (k < cached) q[k] = leaf[k * stride + t];
#pragma unroll
            for (uint32_t k = 0; k < LEAF_CACHE; ++k)
                if (k < cached) em.point_fx(q[k]);
            if (so & SUB_LAST_CLOSE) {
                em.p = op_path[i];
                em.close(true);
            }
            bad |= em.bad;
            continue;
        }
        bad |= emit_op(ops, op_path, paths, i, s, mode, lastpw, off, sink, leaf ? leaf + t : nullptr, cached,
                       stride);
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// Flattening in one pass, for fill runs whose (op, lane) items fit one
// resident grid at one item per thread and are taken in op order: a block
// takes the items of its look-back ticket, counts them (keeping each curve
// lane's first LEAF_CACHE points in LDS), takes its prefix by look-back over
// the tickets before it, and emits them (a copy from LDS unless a lane has
// more leaves).  Writes each op's offset
// (op_off, as the op scan does) and the run's event and close totals (as
// k_emit does).
__global__ __launch_bounds__(BLOCK) void k_flatten(const ft_path_op *ops, const uint32_t *op_path,
                                                   const PathDev *paths, uint32_t n_ops,
                                                   unsigned long long *op_off, unsigned long long *lb,
                                                   EventSink sink, DevStatus *st, bool tickets) {
    __shared__ unsigned long long s_wave[BLOCK / 64];
    __shared__ unsigned long long s_prefix;
    __shared__ uint32_t s_tk;
    __shared__ I2 s_leaf[LEAF_CACHE * BLOCK];  // point k of thread j at s_leaf[k * BLOCK + j]
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t n_items = (uint64_t)n_ops * OP_LANES;
    // the block's items are those of its look-back ticket (or of its index)
    if (tickets) {
        if (threadIdx.x == 0) s_tk = lookback_ticket(lb);
        __syncthreads();
    }
    const uint32_t b = tickets ? s_tk : blockIdx.x;
    const uint64_t t = (uint64_t)b * BLOCK + threadIdx.x;
    const bool live = t < n_items;
    const uint32_t i = live ? (uint32_t)(t >> FLAT_SPLIT) : 0u, s = (uint32_t)t & (OP_LANES - 1);
    OpCount c = {0, 0, 0, 0};
    if (live) c = count_op(ops, op_path, paths, i, s, MODE_FILL, nullptr, s_leaf + threadIdx.x, BLOCK);
    const unsigned long long v = live ? pack_count(c) : 0ull;
    uint32_t leaves = c.leaves;
#pragma unroll
    for (uint32_t o = 1; o < OP_LANES; o <<= 1) leaves += __shfl_xor(leaves, o, OP_LANES);
    if (live & (c.err != 0u)) raise_err(st, c.err);
    if (live & (s == 0) & (leaves > FT_FLATTEN_MAX_LEAVES)) raise_err(st, FT_ERANGE);
    // the block's exclusive prefix of the packed (events, closes) counts
    const unsigned long long winc = wave_incl_t<unsigned long long, false>(v);
    if (lane == 63) s_wave[wid] = winc;
    __syncthreads();
    unsigned long long wpre = 0, btot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        wpre += w < wid ? s_wave[w] : 0ull;
        btot += s_wave[w];
    }
    if (wid == 0) {
        const unsigned long long prefix = lookback_prefix<unsigned long long, false>(lb, b, btot, st);
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    const unsigned long long off = s_prefix + wpre + winc - v;
    if (live & (s == 0)) op_off[i] = off;
    const uint64_t last_t = n_items - 1;
    if ((uint64_t)b == last_t / BLOCK && threadIdx.x == 0) {  // the run's totals
        const unsigned long long tot = s_prefix + btot;
        const uint32_t n_events = (uint32_t)(tot >> 32);
        st->op_total = tot;
        st->n_events = n_events;
        st->n_closes = (uint32_t)tot;
        if (n_events > sink.cap_events) {
            atomicOr(&st->overflow, 1u);
            st->need_events = n_events;
        }
    }
    // a lane whose events would pass the capacity writes nothing: the host
    // grows the buffers and runs again
    if (!live || (off >> 32) + c.ne > sink.cap_events || (uint32_t)off + c.nc > sink.cap_runs) return;
    if (emit_op(ops, op_path, paths, i, s, MODE_FILL, nullptr, off, sink, s_leaf + threadIdx.x, c.leaves, BLOCK))
        raise_err(st, FT_EDEVICE);
}

// Limits of the current run, from the counts the earlier stages produced.
__device__ inline Limits run_limits(const DevStatus *st, uint32_t n_ops, uint32_t n_paths, uint32_t n_bands,
                                    uint32_t cap_points) {
    Limits L;
    L.n_ops = n_ops;
    L.n_paths = n_paths;
    L.n_bands = n_bands;
    L.n_events = st->n_events;
    L.n_runs = st->n_closes;
    L.cap_points = cap_points;
    L.n_pairs = st->n_pairs;
    return L;
}

// ---------------------------------------------------------------------------
// Stages 3-4: Fig ingest (runs, closing pop, cap, compaction, top point).
// ---------------------------------------------------------------------------
// Ingest as scan sources: the first pass of a scan computes its items from
// the events (and stores the per-event side outputs), so no kernel writes the
// items for the scan to read back.
//
// IngestSrc: ingest_event's pushed-point marks, whose running maximum is the
// last pushed point so far; pass 1 also records each run's first point.
struct IngestSrc {
    static constexpr uint32_t ITEMS = 4;
    Limits L;  // n_ops, n_paths, n_bands, cap_points set by the host; the rest by bind()
    const uint32_t *meta;
    const I2 *xy;
    int32_t *run_first;
    const uint32_t *run_path;
    uint32_t bad;
    __device__ void bind(const DevStatus *st) { L = run_limits(st, L.n_ops, L.n_paths, L.n_bands, L.cap_points); }
    __device__ uint32_t value(uint64_t i, bool first) {
        uint32_t fr;
        const uint32_t v = ingest_mark((uint32_t)i, L, meta, xy, run_path, bad, fr);
        if (first & (fr != NO_IDX)) run_first[fr] = (int32_t)i;
        return v;
    }
};

// KeptSrc: kept_event's flags (the point survives the closing pop), whose
// exclusive sum is the kept point's index; pass 1 also stores the flags.
struct KeptSrc {
    static constexpr uint32_t ITEMS = 4;
    Limits L;
    const uint32_t *meta;
    const I2 *xy;
    const int32_t *run_first;
    const uint32_t *run_close;
    const uint32_t *last_scan;
    uint32_t *kept;
    uint32_t bad;
    __device__ void bind(const DevStatus *st) { L = run_limits(st, L.n_ops, L.n_paths, L.n_bands, L.cap_points); }
    __device__ uint32_t value(uint64_t i, bool first) {
        const uint32_t v = kept_event((uint32_t)i, L, meta, xy, run_first, run_close, last_scan, bad);
        if (first) kept[i] = v;
        return v;
    }
};

// KeptLocalSrc: kept_event_local's flags, with no ingest pass before them;
// pass 1 also stores the flags and each run's first point, and raises
// OV_LONG_WALK when a closing walk is too long (the host then launches the
// run again with IngestSrc and KeptSrc).
struct KeptLocalSrc {
    static constexpr uint32_t ITEMS = 4;
    Limits L;
    const uint32_t *meta;
    const I2 *xy;
    int32_t *run_first;
    const uint32_t *run_close;
    uint32_t *kept;
    uint32_t bad;
    DevStatus *st;
    __device__ void bind(DevStatus *s) {
        st = s;
        L = run_limits(s, L.n_ops, L.n_paths, L.n_bands, L.cap_points);
    }
    __device__ uint32_t value(uint64_t i, bool first) {
        uint32_t lw = 0, fr;
        const uint32_t v = kept_event_local((uint32_t)i, L, meta, xy, run_close, KEPT_WALK, bad, lw, fr);
        if (first) {
            kept[i] = v;
            if (fr != NO_IDX) run_first[fr] = (int32_t)i;
            if (lw) atomicOr(&st->overflow, OV_LONG_WALK);
        }
        return v;
    }
};

static_assert(IngestSrc::ITEMS == KeptSrc::ITEMS, "one tile size decides both ingest scans");
static_assert(KeptLocalSrc::ITEMS == KeptSrc::ITEMS, "one tile size decides the kept scans");
constexpr size_t TOP_FEW_PATHS = 64;  // batches below this many paths keep TOP_REPS partial top keys each
static_assert((TOP_REPS & (TOP_REPS - 1)) == 0, "block residues by mask");

// Plain passes for the pushed-point marks and the kept flags, for runs whose
// scans have several tiles per block (where IngestSrc / KeptSrc would
// evaluate every event twice).
__global__ __launch_bounds__(BLOCK) void k_ingest(Limits L0, const uint32_t *ev_meta, const I2 *ev_xy,
                                                  int32_t *run_first, const uint32_t *run_path, uint32_t *ev_last,
                                                  DevStatus *st) {
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    uint32_t bad = 0;
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < L.n_events; e += gridDim.x * BLOCK)
        ingest_event(e, L, ev_meta, ev_xy, run_first, run_path, ev_last, bad);
    if (bad) raise_err(st, FT_EDEVICE);
}

__global__ __launch_bounds__(BLOCK) void k_kept(Limits L0, const uint32_t *ev_meta, const I2 *ev_xy,
                                                const int32_t *run_first, const uint32_t *run_close,
                                                const uint32_t *ev_last, uint32_t *ev_kept, DevStatus *st) {
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    uint32_t bad = 0;
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < L.n_events; e += gridDim.x * BLOCK)
        ev_kept[e] = kept_event(e, L, ev_meta, ev_xy, run_first, run_close, ev_last, bad);
    if (bad) raise_err(st, FT_EDEVICE);
}

// The kept flags' local form as a pass of its own (runs whose scans have
// several tiles per block).
__global__ __launch_bounds__(BLOCK) void k_kept_local(Limits L0, const uint32_t *ev_meta, const I2 *ev_xy,
                                                      int32_t *run_first, const uint32_t *run_close,
                                                      uint32_t *ev_kept, DevStatus *st) {
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    uint32_t bad = 0, lw = 0;
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < L.n_events; e += gridDim.x * BLOCK) {
        uint32_t fr;
        ev_kept[e] = kept_event_local(e, L, ev_meta, ev_xy, run_close, KEPT_WALK, bad, lw, fr);
        if (fr != NO_IDX) run_first[fr] = (int32_t)e;
    }
    if (lw) atomicOr(&st->overflow, OV_LONG_WALK);
    if (bad) raise_err(st, FT_EDEVICE);
}

// One wavefront per path; the cap point is found by a 64-way search (each
// round probes 64 events at once), a few round trips for any path length.
__global__ __launch_bounds__(BLOCK) void k_cap(Limits L0, const PathDev *paths, const unsigned long long *op_off,
                                               const uint32_t *ev_meta, const I2 *ev_xy, const uint32_t *ev_kept,
                                               const uint32_t *ev_idx, const int32_t *run_first, PathState *ps,
                                               DevStatus *st) {
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    uint32_t bad = 0;
    for (uint32_t p = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); p < L.n_paths; p += waves) {
        const CapSpan c = cap_span(p, L, paths, op_off, ev_kept, ev_idx, bad);
        uint32_t repop = 0;
        if (c.over) {
            // cap_after is false below lo and true from hi on (hi = ee: none)
            uint32_t lo = c.eb, hi = c.ee;
            while (hi - lo > 64) {
                const uint32_t step = (hi - lo + 63) / 64;
                uint32_t m = lo + (lane + 1) * step - 1;
                m = m < hi ? m : hi - 1;
                const unsigned long long t = __ballot(cap_after(m, c, ev_kept, ev_idx));
                if (t) {
                    const int f = __ffsll((long long)t) - 1;
                    const uint32_t mf = __shfl(m, f, 64), mp = __shfl(m, f > 0 ? f - 1 : 0, 64);
                    hi = mf;
                    lo = f > 0 ? mp + 1 : lo;
                } else {
                    lo = hi;
                }
            }
            const uint32_t m = lo + lane;
            const unsigned long long t = __ballot((m < hi) && cap_after(m < hi ? m : lo, c, ev_kept, ev_idx));
            const uint32_t ea = t ? lo + (uint32_t)(__ffsll((long long)t) - 1) : hi;
            bad |= ea < c.ee ? 0u : 1u;
            if (lane == 0 && ea < c.ee) repop = cap_repop(ea, c, L, ev_meta, ev_xy, run_first);
        }
        if (lane == 0) cap_store(p, c, repop, ps);
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// Compaction of the stored points, and each path's minimum (y, x) key over
// them: a block merges its points' keys per path and flushes one atomic per
// path it holds, into PathState::topkey (top_rep null) or into one of the
// path's TOP_REPS partial minima (block b into b mod TOP_REPS), so that a
// batch of few paths does not queue every block's atomic on one address;
// k_edges then folds the partials into PathState::topkey.
__global__ __launch_bounds__(BLOCK) void k_compact(Limits L0, const uint32_t *ev_meta, const I2 *ev_xy,
                                                   const uint32_t *ev_kept, const uint32_t *ev_idx,
                                                   const int32_t *run_first, const uint32_t *run_path,
                                                   uint32_t *run_start, uint32_t *run_end, I2 *pts, uint32_t *pt_run,
                                                   PathState *ps, unsigned long long *top_rep, DevStatus *st) {
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    uint32_t bad = 0;
    for (uint32_t base = blockIdx.x * BLOCK; base < L.n_events; base += gridDim.x * BLOCK) {
        const uint32_t e = base + threadIdx.x;
        TopCand c;
        c.path = NO_IDX;
        c.key = NO_KEY;
        if (e < L.n_events)
            compact_event(e, L, ev_meta, ev_xy, ev_kept, ev_idx, run_first, run_path, run_start, run_end, pts,
                          pt_run, ps, bad, &c);
        block_reduce_keyed<unsigned long long>(
            c.path != NO_IDX, c.path, c.key, NO_KEY,
            [](unsigned long long a, unsigned long long b) { return a < b ? a : b; },
            [ps, top_rep, &L](uint32_t k, unsigned long long v) {
                if (k >= L.n_paths) return;
                unsigned long long *const at =
                    top_rep ? &top_rep[(size_t)k * TOP_REPS + (blockIdx.x & (TOP_REPS - 1))] : &ps[k].topkey;
                atomicMin(at, v);
            });
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// ---------------------------------------------------------------------------
// Stage 5: edges and row-band binning.
// ---------------------------------------------------------------------------
// Band binning.  The (edge, band) pairs of a chunk of consecutive stored
// points are counted in an LDS histogram over the window of
// BIN_WINDOW bands starting at the chunk's lowest band, and each nonzero bin
// goes to global memory with one atomic.  The global counters are replicated:
// band g, replica r sits at slot (g << band_shift) + r, and chunk c uses
// replica c mod 2^band_shift, so few bands do not put every chunk's atomics
// on the same cache lines.  An exclusive scan over the slots gives each
// (band, replica) its list segment; a band's list is its replicas' segments
// back to back, starting at slot g << band_shift.
constexpr size_t OP_ORDER_MIN = 1024;  // batches from this many ops flatten in tag order
constexpr uint32_t BIN_WINDOW = 1024;
// Points per thread of a binning chunk (a template parameter of k_edges and
// k_bin, the same for both): 4 for batches of at least BIN_WIDE_MIN points,
// whose kernels run many chunks per block; 1 below, where one chunk per block
// keeps each thread's load chain short.
constexpr size_t BIN_WIDE_MIN = (size_t)1 << 21;
constexpr size_t BAND_SLOTS_MIN = 16384;
constexpr uint32_t BAND_SHIFT_MAX = 5;
static_assert(BIN_WINDOW == 4 * BLOCK, "the window scan gives each thread four bins");

__device__ inline uint32_t block_min_u32(uint32_t v, uint32_t *s_w) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t m = s_w[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) m = s_w[w] < m ? s_w[w] : m;
    __syncthreads();
    return m;
}

// Pairs per band of the chunk's window [gmin, gmin + BIN_WINDOW) into s_h,
// from a difference per edge and an inclusive scan over the window.  Returns
// gmin (NO_IDX: no banded edge in the chunk).  Every thread of the block calls
// it; the caller synchronises before s_h is written again.
template <int NI>
__device__ inline uint32_t bin_window_counts(const bool *on, const uint32_t *g0, const uint32_t *g1, uint32_t *s_h,
                                             uint32_t *s_w) {
    uint32_t m = NO_IDX;
#pragma unroll
    for (int j = 0; j < NI; ++j) m = (on[j] & (g0[j] < m)) ? g0[j] : m;
    const uint32_t gmin = block_min_u32(m, s_w);
    if (gmin == NO_IDX) return NO_IDX;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    reinterpret_cast<uint4 *>(s_h)[t] = make_uint4(0, 0, 0, 0);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        if (on[j] & (g0[j] - gmin < BIN_WINDOW)) {
            const uint32_t hi = g1[j] - gmin < BIN_WINDOW - 1 ? g1[j] - gmin : BIN_WINDOW - 1;
            atomicAdd(&s_h[g0[j] - gmin], 1u);
            if (hi + 1 < BIN_WINDOW) atomicAdd(&s_h[hi + 1], 0xFFFFFFFFu);
        }
    }
    __syncthreads();
    const uint4 q = reinterpret_cast<const uint4 *>(s_h)[t];
    const uint32_t a0 = q.x, a1 = a0 + q.y, a2 = a1 + q.z, a3 = a2 + q.w;
    const uint32_t x = wave_incl_dpp<false>(a3);
    if (lane == 63) s_w[wid] = x;
    __syncthreads();
    uint32_t pre = x - a3;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) pre += w < wid ? s_w[w] : 0u;
    reinterpret_cast<uint4 *>(s_h)[t] = make_uint4(pre + a0, pre + a1, pre + a2, pre + a3);
    __syncthreads();
    return gmin;
}

// One thread per stored point [0, n_kept): Edge::new, the point's band
// record for k_bin, and the pairs per (band, replica).
template <int BIN_ITEMS>
__global__ __launch_bounds__(BLOCK) void k_edges(Limits L0, uint32_t band_shift, const uint32_t *pt_run,
                                                 const uint32_t *run_path, const uint32_t *run_start,
                                                 const uint32_t *run_end, const PathDev *paths, const I2 *pts,
                                                 EdgeRec *edges, U2 *pt_band, uint32_t *band_cnt, PathState *ps,
                                                 const unsigned long long *top_rep, DevStatus *st) {
    __shared__ __align__(16) uint32_t s_h[BIN_WINDOW];
    __shared__ uint32_t s_w[BLOCK / 64];
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    const uint32_t n = st->n_kept < L.cap_points ? st->n_kept : L.cap_points;
    const uint32_t rep_mask = (1u << band_shift) - 1u;
    constexpr uint32_t BIN_CHUNK = BIN_ITEMS * BLOCK;  // points per chunk (one LDS window)
    uint32_t bad = 0;
    for (uint32_t b = blockIdx.x * BIN_CHUNK; b < n; b += gridDim.x * BIN_CHUNK) {
        bool on[BIN_ITEMS];
        uint32_t g0[BIN_ITEMS], g1[BIN_ITEMS];
#pragma unroll
        for (int j = 0; j < (int)BIN_ITEMS; ++j) {
            const uint32_t i = b + (uint32_t)j * BLOCK + threadIdx.x;
            U2 rec;
            rec.x = 1;
            rec.y = 0;
            if (i < n) {
                const EdgeOut r = edge_point<DevAtom>(i, L, pt_run, run_path, run_start, run_end, paths, pts, edges,
                                                      ps, bad, top_rep, TOP_REPS);
                rec = band_record(r);
                pt_band[i] = rec;
            }
            on[j] = (rec.y >= rec.x) & (rec.y < L.n_bands);
            g0[j] = rec.x;
            g1[j] = rec.y;
        }
        const uint32_t rep = (b / BIN_CHUNK) & rep_mask;
        const uint32_t gmin = bin_window_counts<BIN_ITEMS>(on, g0, g1, s_h, s_w);
        if (gmin == NO_IDX) continue;
        const uint4 q = reinterpret_cast<const uint4 *>(s_h)[threadIdx.x];
        const uint32_t c[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t g = gmin + 4 * threadIdx.x + j;
            if (c[j] && g < L.n_bands) atomicAdd(&band_cnt[(g << band_shift) + rep], c[j]);
        }
        // bands past the window, one atomic each (edges far below the chunk's top band)
#pragma unroll
        for (int j = 0; j < (int)BIN_ITEMS; ++j)
            if (on[j])
                for (uint32_t g = gmin + BIN_WINDOW > g0[j] ? gmin + BIN_WINDOW : g0[j]; g <= g1[j]; ++g)
                    atomicAdd(&band_cnt[(g << band_shift) + rep], 1u);
        __syncthreads();
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// Scatters each point's pairs into the band lists: one atomic per nonzero
// (chunk, band) reserves the chunk's range in the replica's segment, and LDS
// cursors hand out slots within it.  Order within a list is not fixed; the
// raster's integer sums do not depend on it.
template <int BIN_ITEMS>
__global__ __launch_bounds__(BLOCK) void k_bin(Limits L0, uint32_t band_shift, const U2 *pt_band,
                                               const uint32_t *band_off, uint32_t *band_cur, uint32_t *band_list,
                                               uint64_t cap_pairs, const I2 *pts, const uint32_t *pt_run,
                                               const uint32_t *run_start, const uint32_t *run_end, PathState *ps,
                                               DevStatus *st) {
    __shared__ __align__(16) uint32_t s_h[BIN_WINDOW];
    __shared__ __align__(16) uint32_t s_base[BIN_WINDOW];
    __shared__ uint32_t s_w[BLOCK / 64];
    if (run_failed(st)) return;
    const uint32_t n_pairs = st->n_pairs;
    if (n_pairs > cap_pairs) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            atomicOr(&st->overflow, 2u);
            st->need_pairs = n_pairs;
        }
        return;
    }
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    const uint32_t n = st->n_kept < L.cap_points ? st->n_kept : L.cap_points;
    const uint32_t rep_mask = (1u << band_shift) - 1u;
    const uint32_t n_slots = L.n_bands << band_shift;
    constexpr uint32_t BIN_CHUNK = BIN_ITEMS * BLOCK;  // points per chunk (one LDS window)
    uint32_t bad = 0;
    // winding direction per path (its top point is final after k_edges)
    for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < L.n_paths; p += gridDim.x * BLOCK) {
        const PathState S = ps[p];
        if ((S.topkey != NO_KEY) & (S.topidx != NO_IDX))
            ps[p].fwd = path_forward(S, L, pts, pt_run, run_start, run_end, bad) ? 1u : 0u;
    }
    // list segment [begin, end) of a slot
    auto seg_end = [&](uint32_t slot) { return slot + 1 < n_slots ? band_off[slot + 1] : n_pairs; };
    for (uint32_t b = blockIdx.x * BIN_CHUNK; b < n; b += gridDim.x * BIN_CHUNK) {
        bool on[BIN_ITEMS];
        uint32_t g0[BIN_ITEMS], g1[BIN_ITEMS];
#pragma unroll
        for (int j = 0; j < (int)BIN_ITEMS; ++j) {
            const uint32_t i = b + (uint32_t)j * BLOCK + threadIdx.x;
            const BinItem it = bin_item(i < n ? i : 0u, pt_band);
            on[j] = it.banded & (i < n) & (it.g1 < L.n_bands);
            g0[j] = it.g0;
            g1[j] = it.g1;
        }
        const uint32_t rep = (b / BIN_CHUNK) & rep_mask;
        const uint32_t gmin = bin_window_counts<BIN_ITEMS>(on, g0, g1, s_h, s_w);
        if (gmin == NO_IDX) continue;
        {
            const uint4 q = reinterpret_cast<const uint4 *>(s_h)[threadIdx.x];
            const uint32_t c[4] = {q.x, q.y, q.z, q.w};
            uint32_t base[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t g = gmin + 4 * threadIdx.x + j;
                base[j] = NO_IDX;
                if (c[j] && g < L.n_bands) {
                    const uint32_t slot = (g << band_shift) + rep;
                    const uint32_t at = band_off[slot] + atomicAdd(&band_cur[slot], c[j]);
                    const bool fits = (uint64_t)at + c[j] <= seg_end(slot);
                    bad |= fits ? 0u : 1u;
                    base[j] = fits ? at : NO_IDX;
                }
            }
            reinterpret_cast<uint4 *>(s_base)[threadIdx.x] = make_uint4(base[0], base[1], base[2], base[3]);
            reinterpret_cast<uint4 *>(s_h)[threadIdx.x] = make_uint4(0, 0, 0, 0);  // cursors
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < (int)BIN_ITEMS; ++jj) {
            if (!on[jj]) continue;
            const uint32_t i = b + (uint32_t)jj * BLOCK + threadIdx.x;
            const uint32_t hi = g1[jj] - gmin < BIN_WINDOW - 1 ? g1[jj] : gmin + BIN_WINDOW - 1;
            for (uint32_t g = g0[jj]; g <= hi; ++g) {
                const uint32_t k = g - gmin;
                const uint32_t slot = atomicAdd(&s_h[k], 1u);
                const uint32_t at = s_base[k];
                if (at != NO_IDX) band_list[at + slot] = i;
            }
            for (uint32_t g = gmin + BIN_WINDOW > g0[jj] ? gmin + BIN_WINDOW : g0[jj]; g <= g1[jj]; ++g) {
                const uint32_t slot = (g << band_shift) + rep;
                const uint32_t at = band_off[slot] + atomicAdd(&band_cur[slot], 1u);
                if (at < seg_end(slot))
                    band_list[at] = i;
                else
                    bad = 1;
            }
        }
        __syncthreads();
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// Edges per path over the stored points (reported on request, not needed by
// the raster).
__global__ __launch_bounds__(BLOCK) void k_edge_count(Limits L0, const EdgeRec *edges, const uint32_t *pt_run,
                                                      const uint32_t *run_path, PathState *ps, DevStatus *st) {
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    const uint32_t n = st->n_kept < L.cap_points ? st->n_kept : L.cap_points;
    for (uint32_t b = blockIdx.x * BLOCK; b < n; b += gridDim.x * BLOCK) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t ii = i < n ? i : 0;
        const uint32_t run = pt_run[ii];
        const uint32_t p = run_path[run < L.n_runs ? run : 0];
        const bool has = (i < n) & (run < L.n_runs) & (p < L.n_paths) & ((edges[ii].flags & 1u) != 0);
        block_reduce_keyed<uint32_t>(
            has, p, 1u, 0u, [](uint32_t a, uint32_t c) { return a + c; },
            [ps, &L](uint32_t k, uint32_t v) {
                if (k < L.n_paths) atomicAdd(&ps[k].n_edges, v);
            });
    }
}

// ---------------------------------------------------------------------------
// Stroker stages (MODE_STROKE): flattened stroke events -> outline events.
// ---------------------------------------------------------------------------
// atan2 and sin evaluated in double and rounded once (they only choose joins)
struct DevMath {
    __device__ static float atan2(float y, float x) { return (float)::atan2((double)y, (double)x); }
    __device__ static float sin(float t) { return (float)::sin((double)t); }
};

__global__ __launch_bounds__(BLOCK) void k_st_flags(const F4 *ev, unsigned long long *flags, uint32_t *last,
                                                    DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n = st->n_events;
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < n; e += gridDim.x * BLOCK) {
        const StrokeFlags f = stroke_flags(e, ev);
        flags[e] = ((unsigned long long)(f.pushed ? 1u : 0u) << 32) | (f.first ? 1u : 0u);
        last[e] = f.is_pt ? 0u : e + 1;
    }
}

__global__ __launch_bounds__(BLOCK) void k_st_compact(const F4 *ev, const unsigned long long *off, F4 *spts,
                                                      uint32_t *sub_start, uint32_t *sub_event, uint32_t cap,
                                                      DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n = st->n_events;
    uint32_t bad = 0;
    for (uint32_t e = blockIdx.x * BLOCK + threadIdx.x; e < n; e += gridDim.x * BLOCK) {
        const StrokeFlags f = stroke_flags(e, ev);
        const unsigned long long o = off[e];
        const uint32_t sidx = (uint32_t)(o >> 32), sub = (uint32_t)o;
        if (f.pushed & (sidx < FT_MAX_POINTS) & (sidx < cap)) spts[sidx] = ev[e];
        if (f.first & (sub < cap)) {
            sub_start[sub] = sidx;
            sub_event[sub] = e;
        }
        bad |= (f.first & (sub >= cap)) ? 1u : 0u;
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// One thread per sub-stroke: point count, joined flag, and the number of
// sub-strokes below the cap; then the outline's item count.
__global__ __launch_bounds__(BLOCK) void k_st_subs(const F4 *ev, const uint32_t *last_close,
                                                   const uint32_t *sub_start, const uint32_t *sub_event,
                                                   uint32_t *sub_n, uint32_t *sub_joined, uint32_t cap,
                                                   DevStatus *st) {
    if (run_failed(st)) return;
    const unsigned long long tot = st->st_total;
    const uint32_t n_ev = st->n_events, total_subs = (uint32_t)tot;
    const uint32_t n_kept = (uint32_t)(tot >> 32) < FT_MAX_POINTS ? (uint32_t)(tot >> 32) : FT_MAX_POINTS;
    if (total_subs > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) raise_err(st, FT_EDEVICE);
        return;
    }
    for (uint32_t s = blockIdx.x * BLOCK + threadIdx.x; s < total_subs; s += gridDim.x * BLOCK) {
        const uint32_t a = sub_start[s];
        const bool kept = a < FT_MAX_POINTS;
        const bool more = s + 1 < total_subs;
        const uint32_t a1 = more ? sub_start[s + 1] : n_kept;
        const bool last_kept = kept & (!more | (a1 >= FT_MAX_POINTS));
        const uint32_t end = a1 < n_kept ? a1 : n_kept;
        // closes up to the next sub-stroke's first point (or the end) apply to this one
        const uint32_t range_end = (last_kept | !more) ? n_ev : sub_event[s + 1];
        const uint32_t lc = last_close[range_end ? range_end - 1 : 0];
        const bool has_close = (range_end > 0) & (lc > 0) & (lc - 1 >= sub_event[s]) & (lc - 1 < n_ev);
        const F4 ce = ev[has_close ? lc - 1 : 0];
        const bool joined = has_close & (ce.w == 2.0f);
        sub_n[s] = kept & (end > a) ? end - a : 0u;
        sub_joined[s] = joined ? 1u : 0u;
        if (last_kept) {
            st->st_subs = s + 1;
            st->st_items = 2u * (n_kept + s + 1) + 1u;
        }
    }
    if (total_subs == 0 && blockIdx.x == 0 && threadIdx.x == 0) st->st_items = 1u;  // the final close only
}

__device__ inline uint32_t st_item_sub(uint32_t t, const uint32_t *sub_start, uint32_t n_subs) {
    uint32_t lo = 0, hi = n_subs;  // last s with 2 (start_s + s) <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (2u * (sub_start[mid] + mid) <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

struct StrokeParams {
    uint32_t join;
    float miter, tol_sq;
};

__global__ __launch_bounds__(BLOCK) void k_st_count(const F4 *spts, const uint32_t *sub_start, const uint32_t *sub_n,
                                                    const uint32_t *sub_joined, StrokeParams sp,
                                                    unsigned long long *cnt, DevStatus *st) {
    if (run_failed(st)) return;
    const uint32_t n_items = st->st_items, n_subs = st->st_subs;
    for (uint32_t t = blockIdx.x * BLOCK + threadIdx.x; t < n_items; t += gridDim.x * BLOCK) {
        uint32_t ne = 0, nc = 0, err = 0;
        if (t + 1 == n_items) {
            ne = nc = 1;  // Plotter::fill's final close
        } else if (n_subs) {
            const uint32_t si = st_item_sub(t, sub_start, n_subs);
            StrokeSub sub;
            sub.start = sub_start[si];
            sub.n = sub_n[si];
            sub.joined = sub_joined[si] != 0;
            const uint32_t local = t - 2u * (sub.start + si);
            const bool fwd = local <= sub.n;
            const uint32_t k = fwd ? local : local - (sub.n + 1);
            if (sub.n)
                err = stroke_item<DevMath>(
                    sub, fwd, k, spts, sp.join, sp.miter, sp.tol_sq, [&](SV2) { ++ne; },
                    [&]() {
                        ++ne;
                        ++nc;
                    });
        }
        if (err) raise_err(st, err);
        cnt[t] = ((unsigned long long)ne << 32) | nc;
    }
}

__global__ __launch_bounds__(BLOCK) void k_st_emit(const F4 *spts, const uint32_t *sub_start, const uint32_t *sub_n,
                                                   const uint32_t *sub_joined, StrokeParams sp,
                                                   const unsigned long long *off, const PathDev *paths,
                                                   EventSink sink, DevStatus *st) {
    if (run_failed(st)) return;
    const unsigned long long tot = st->stroke_total;
    const uint32_t n_events = (uint32_t)(tot >> 32);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->n_events = n_events;
        st->n_closes = (uint32_t)tot;
        if (n_events > sink.cap_events) {
            atomicOr(&st->overflow, 1u);
            st->need_events = n_events;
        }
    }
    if (n_events > sink.cap_events) return;
    const uint32_t n_items = st->st_items, n_subs = st->st_subs;
    const PathDev &P = paths[0];
    uint32_t bad = 0;
    for (uint32_t t = blockIdx.x * BLOCK + threadIdx.x; t < n_items; t += gridDim.x * BLOCK) {
        EventSink em = sink;
        em.e = off[t] >> 32;
        em.run = (uint32_t)off[t];
        em.p = 0;
        if (t + 1 == n_items) {
            em.close(true);
        } else if (n_subs) {
            const uint32_t si = st_item_sub(t, sub_start, n_subs);
            StrokeSub sub;
            sub.start = sub_start[si];
            sub.n = sub_n[si];
            sub.joined = sub_joined[si] != 0;
            const uint32_t local = t - 2u * (sub.start + si);
            const bool fwd = local <= sub.n;
            const uint32_t k = fwd ? local : local - (sub.n + 1);
            if (sub.n)
                stroke_item<DevMath>(
                    sub, fwd, k, spts, sp.join, sp.miter, sp.tol_sq,
                    [&](SV2 q) { em.point(xform(P, q.x, q.y, 0.0f)); }, [&]() { em.close(true); });
        }
        bad |= em.bad;
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

// ---------------------------------------------------------------------------
// Stage 6: raster.  One block per row band and LDS row group.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(1))) uint8_t gbyte;  // raster memory
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4 gu32x4;
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ inline uint32_t wave_incl_u32(uint32_t v) { return wave_incl_dpp<false>(v); }

// Coverage of a running cell sum under a fixed fill rule (0: NonZero,
// 1: EvenOdd), as map_coverage.
template <uint32_t RULE>
__device__ __forceinline__ uint32_t cov_of(uint32_t sum) {
    if (RULE) {  // coverage_even_odd: |(s & 0xFF) - (s & 0x100)| capped at 255, s = the sum as i16
        const uint32_t c = __sad(sum & 0xFFu, sum & 0x100u, 0u);
        return c < 255u ? c : 255u;
    }
    return coverage_non_zero((int32_t)sum);
}

// SrcOver of the path's constant premultiplied colour (composite_px) depends
// on the coverage a only through S_c(a) = ch8_mul_w(widen(src_c), widen(a))
// and invw(a) = widen(255 - S_alpha(a)):
//     out_c = min(255, S_c(a) + ch8_mul_w(widen(d_c), invw(a))).
// A raster block tabulates them for a = 0..255 in LDS.  Entry, Rgba8p:
// x = S0 | S2 << 16, y = S1 | S3 << 16, z = invw, w = the S bytes as a pixel;
// Graya8p: x = S0 | S1 << 16, y = 0, z = invw, w = S0 | S1 << 8.
constexpr uint32_t LUT_ENTRIES = 256;
constexpr uint32_t LUT_BYTES = LUT_ENTRIES * sizeof(uint4);
// List entries per balanced cell-pass round: their slots share the front of
// the block's LDS with the composite table (128 fit its 4 KiB); Matte8 has no
// table and takes 256 in an 8 KiB front.
constexpr uint32_t CELL_ROUND = 128;
template <int FMT>
constexpr uint32_t cell_round() {
    return FMT == FMT_MATTE8 ? 2u * CELL_ROUND : CELL_ROUND;
}
template <int FMT>
constexpr uint32_t front_bytes() {
    return FMT == FMT_MATTE8 ? 2u * LUT_BYTES : LUT_BYTES;
}
static_assert(2 * CELL_ROUND * (sizeof(EdgeRec) + sizeof(uint32_t)) <= 2 * LUT_BYTES, "Matte8 slots fit its front");
constexpr uint32_t DIRECT_ROWS = 4;   // row groups up to this tall take the direct cell pass
static_assert(CELL_ROUND * (sizeof(EdgeRec) + sizeof(uint32_t)) <= LUT_BYTES, "edge slots fit the table's LDS");
static_assert(2 * CELL_ROUND <= BLOCK, "one entry per thread");

static inline size_t raster_lds_bytes(uint32_t fmt, size_t cells) {
    return (fmt == FMT_MATTE8 ? front_bytes<FMT_MATTE8>() : front_bytes<FMT_RGBA8P>()) + cells * sizeof(int32_t);
}

__device__ inline uint4 lut_entry(uint32_t fmt, uint32_t clr, uint32_t a) {
    const uint32_t aw = ch8_widen(a);
    if (fmt == FMT_RGBA8P) {
        uint32_t s[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) s[c] = ch8_mul_w(ch8_widen((clr >> (8 * c)) & 0xFFu), aw);
        return make_uint4(s[0] | (s[2] << 16), s[1] | (s[3] << 16), ch8_widen(255u - s[3]),
                          s[0] | (s[1] << 8) | (s[2] << 16) | (s[3] << 24));
    }
    const uint32_t s0 = ch8_mul_w(ch8_widen(clr & 0xFFu), aw), s1 = ch8_mul_w(ch8_widen((clr >> 8) & 0xFFu), aw);
    return make_uint4(s0 | (s1 << 16), 0u, ch8_widen(255u - s1), s0 | (s1 << 8));
}

// Two channels at once: d2 and s2 hold a channel in the low byte of each
// 16-bit half; returns min(255, s + (widen(d) * invw >> 16)) per half.
// widen(d) and invw are 12-bit, so each product fits 24 bits and its >> 16 is
// the product's byte 2.
__device__ __forceinline__ uint32_t blend2(uint32_t d2, uint32_t s2, uint32_t invw) {
    const uint32_t w2 = (d2 << 4) | __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, d2) >> (u16x2)(4));
    const uint32_t p0 = __umul24(w2 & 0xFFFFu, invw), p1 = __umul24(w2 >> 16, invw);
    const uint32_t t2 = __builtin_amdgcn_perm(p1, p0, 0x0C060C02u);
    const u16x2 r = __builtin_elementwise_min(__builtin_bit_cast(u16x2, s2 + t2), (u16x2)(255));
    return __builtin_bit_cast(uint32_t, r);
}

__device__ __forceinline__ uint32_t blend_rgba(uint32_t d, uint4 e) {
    const uint32_t r02 = blend2(__builtin_amdgcn_perm(0u, d, 0x0C020C00u), e.x, e.z);
    const uint32_t r13 = blend2(__builtin_amdgcn_perm(0u, d, 0x0C030C01u), e.y, e.z);
    return __builtin_amdgcn_perm(r13, r02, 0x06020400u);
}

// Two Graya8p pixels (one 32-bit word) with entries e0 (pixel 0), e1.
__device__ __forceinline__ uint32_t blend_graya2(uint32_t d, uint4 e0, uint4 e1) {
    const uint32_t r0 = blend2(__builtin_amdgcn_perm(0u, d, 0x0C010C00u), e0.x, e0.z);
    const uint32_t r1 = blend2(__builtin_amdgcn_perm(0u, d, 0x0C030C02u), e1.x, e1.z);
    return __builtin_amdgcn_perm(r1, r0, 0x06040200u);
}

// Read-modify-write formats keep the destination loads of the next
// DST_DEPTH chunks of the row segment in flight.
constexpr uint32_t DST_DEPTH = 4;
struct DstPipe {
    uint4 q0, q1, q2, q3;
};

template <int FMT>
constexpr int bpp_of() {
    return FMT == FMT_RGBA8P ? 4 : (FMT == FMT_GRAYA8P ? 2 : 1);
}

// Destination chunk k (16 bytes per lane) of the segment [xb, xe) of a row,
// or zero where the chunk is partial, misaligned or past the segment (the row
// pass then takes its per-pixel path).
template <int FMT, int CPT>
__device__ __forceinline__ uint4 dst_load(const gbyte *row, uint32_t xb, uint32_t xe, uint32_t k, uint32_t nck) {
    constexpr int BPP = bpp_of<FMT>();
    static_assert(CPT * BPP == 16, "one 16-byte chunk per lane");
    const uint32_t x = xb + k * 64u * CPT + (uint32_t)(threadIdx.x & 63) * CPT;
    const gbyte *p = row + (size_t)x * BPP;
    const bool full = (k < nck) & (x + CPT <= xe) & ((((uintptr_t)p) & 15) == 0);
    u32x4 v = (u32x4)(0u);
    if (full) v = *reinterpret_cast<const gu32x4 *>(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Coverage and the stored or composited pixels of CPT cells per lane at x,
// whose running sums are base + v[j] (q: the destination, when loaded).
template <int FMT, int CPT, uint32_t RULE, bool FRESH>
__device__ __forceinline__ void chunk_out(const uint32_t *v, uint32_t base, uint32_t x, uint32_t xe, gbyte *row,
                                          const uint4 *lut, uint32_t clr, uint4 q);

// One chunk of a row pass: CPT cells per lane at x from LDS, their prefix
// sum from `carry` (the sum of the cells left of the chunk), coverage, and
// the pixels stored or composited (q: the chunk's destination, when loaded).
template <int FMT, int CPT, uint32_t RULE, bool FRESH>
__device__ __forceinline__ void row_chunk(const int32_t *cells, uint32_t x, uint32_t xe, gbyte *row,
                                          const uint4 *lut, uint32_t clr, uint32_t &carry, uint4 q) {
    constexpr int BPP = bpp_of<FMT>();
    // A lane inside the segment holds CPT cells that are all in it: segment
    // ends inside the row are multiples of the chunk, and cells in [W, Wp)
    // are never written, so they read as 0.
    const bool in = x < xe;
    const uint32_t xr = in ? x : 0u;  // a readable address for lanes past the segment
    uint32_t c[CPT];
#pragma unroll
    for (int j = 0; j < CPT; j += 4) {
        const int4 c4 = *reinterpret_cast<const int4 *>(cells + xr + j);
        c[j] = (uint32_t)c4.x;
        c[j + 1] = (uint32_t)c4.y;
        c[j + 2] = (uint32_t)c4.z;
        c[j + 3] = (uint32_t)c4.w;
    }
    uint32_t v[CPT];
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        run += c[j];
        v[j] = run;
    }
    run = in ? run : 0u;
    const uint32_t incl = wave_incl_u32(run);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    const uint32_t base = carry + incl - run;
    carry += total;
    if (!in) return;
    chunk_out<FMT, CPT, RULE, FRESH>(v, base, x, xe, row, lut, clr, q);
}

// Rows of at most LPR * CPT pixels, LPR lanes per row (16 or 32): a
// wavefront rasters 64 / LPR rows at once in one chunk each, with the prefix
// sums segmented at LPR lanes.  `cells` and `row` are the lane's own row;
// lanes with `active` false only take part in the scan.
template <int FMT, int CPT, uint32_t RULE, bool FRESH, int LPR>
__device__ __forceinline__ void row_multi(const int32_t *cells, uint32_t W, gbyte *row, const uint4 *lut,
                                          uint32_t clr, bool active) {
    constexpr int BPP = bpp_of<FMT>();
    const uint32_t x = (uint32_t)(threadIdx.x & (LPR - 1)) * CPT;
    const bool in = active & (x < W);
    const uint32_t xr = in ? x : 0u;
    uint32_t v[CPT];
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < CPT; j += 4) {
        const int4 c4 = *reinterpret_cast<const int4 *>(cells + xr + j);
        run += (uint32_t)c4.x;
        v[j] = run;
        run += (uint32_t)c4.y;
        v[j + 1] = run;
        run += (uint32_t)c4.z;
        v[j + 2] = run;
        run += (uint32_t)c4.w;
        v[j + 3] = run;
    }
    uint4 q = make_uint4(0, 0, 0, 0);
    if constexpr (FMT != FMT_MATTE8 && !FRESH) {
        const gbyte *p = row + (size_t)xr * BPP;
        if (in & (x + CPT <= W) & ((((uintptr_t)p) & 15) == 0)) {
            const u32x4 d = *reinterpret_cast<const gu32x4 *>(p);
            q = make_uint4(d.x, d.y, d.z, d.w);
        }
    }
    run = in ? run : 0u;
    const uint32_t incl = seg_incl_dpp<false, LPR>(run);
    if (!in) return;
    chunk_out<FMT, CPT, RULE, FRESH>(v, incl - run, x, W, row, lut, clr, q);
}

template <int FMT, int CPT, uint32_t RULE, bool FRESH>
__device__ __forceinline__ void chunk_out(const uint32_t *v, uint32_t base, uint32_t x, uint32_t xe, gbyte *row,
                                          const uint4 *lut, uint32_t clr, uint4 q) {
    constexpr int BPP = bpp_of<FMT>();
    gbyte *dst = row + (size_t)x * BPP;
    constexpr uint32_t AL = CPT * BPP >= 16 ? 15 : CPT * BPP - 1;
    const bool full = (x + CPT <= xe) & ((((uintptr_t)dst) & AL) == 0);
    if (FMT == FMT_MATTE8) {
        uint32_t a[CPT];
#pragma unroll
        for (int j = 0; j < CPT; ++j) a[j] = cov_of<RULE>(base + v[j]);
        if (full) {
            if constexpr (CPT == 16) {
                uint4 pk;
                pk.x = a[0] | (a[1] << 8) | (a[2] << 16) | (a[3] << 24);
                pk.y = a[4] | (a[5] << 8) | (a[6] << 16) | (a[7] << 24);
                pk.z = a[8] | (a[9] << 8) | (a[10] << 16) | (a[11] << 24);
                pk.w = a[12] | (a[13] << 8) | (a[14] << 16) | (a[15] << 24);
                *reinterpret_cast<gu32x4 *>(dst) = (u32x4){pk.x, pk.y, pk.z, pk.w};
                return;
            } else if constexpr (CPT == 4) {
                *reinterpret_cast<__attribute__((address_space(1))) uint32_t *>(dst) =
                    a[0] | (a[1] << 8) | (a[2] << 16) | (a[3] << 24);
                return;
            }
        }
#pragma unroll
        for (int j = 0; j < CPT; ++j)
            if (x + j < xe) dst[j] = (uint8_t)a[j];
        return;
    }
    if (full) {
        uint32_t w[4] = {q.x, q.y, q.z, q.w};
        if constexpr (!FRESH) {
            // Outside the figure over a transparent destination SrcOver gives
            // 0 (S(0) = 0 and 0 * invw = 0): a wavefront whose every lane has
            // such a chunk leaves it as it is (over other pixels, the bytes
            // the coverage-0 blend changes are changed).  Inside it with an opaque
            // colour, coverage 255 gives the colour's entry whatever the
            // destination (invw(255) = widen(255 - 255) = 0): a wavefront
            // whose every pixel is covered stores that entry without the blend.
            uint32_t cany = 0u, all = 255u;
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                const uint32_t a = cov_of<RULE>(base + v[j]);
                cany |= a;
                all &= a;
            }
            if (__all(cany == 0u)) {
                if (__all((q.x | q.y | q.z | q.w) == 0u)) return;
                // Uncovered over other pixels: SrcOver with coverage 0 maps
                // each byte d to ch8_mul(d, 255) = d - [1 <= d <= 15], so a
                // lane subtracts 1 from such bytes and stores only if it did.
                uint32_t dec[4], ch = 0u;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t lo = w[j] & 0x0F0F0F0Fu, hi = (w[j] >> 4) & 0x0F0F0F0Fu;
                    dec[j] = (((lo + 0x0F0F0F0Fu) & ~(hi + 0x0F0F0F0Fu)) & 0x10101010u) >> 4;
                    ch |= dec[j];
                }
                if (ch)
                    *reinterpret_cast<gu32x4 *>(dst) =
                        (u32x4){w[0] - dec[0], w[1] - dec[1], w[2] - dec[2], w[3] - dec[3]};
                return;
            }
            // the colour's alpha byte: 3 in Rgba8p, 1 in Graya8p (two pixels per word)
            constexpr uint32_t ASH = FMT == FMT_RGBA8P ? 24u : 8u;
            if (((clr >> ASH) & 0xFFu) == 255u && __all(all == 255u)) {
                const uint32_t e = lut[255].w, c = FMT == FMT_RGBA8P ? e : (e | (e << 16));
                *reinterpret_cast<gu32x4 *>(dst) = (u32x4){c, c, c, c};
                return;
            }
        }
        if constexpr (FMT == FMT_RGBA8P) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint4 e = lut[cov_of<RULE>(base + v[j])];
                w[j] = FRESH ? e.w : blend_rgba(w[j], e);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint4 e0 = lut[cov_of<RULE>(base + v[2 * j])];
                const uint4 e1 = lut[cov_of<RULE>(base + v[2 * j + 1])];
                w[j] = FRESH ? (e0.w | (e1.w << 16)) : blend_graya2(w[j], e0, e1);
            }
        }
        *reinterpret_cast<gu32x4 *>(dst) = (u32x4){w[0], w[1], w[2], w[3]};
        return;
    }
    // partial or misaligned chunk: per pixel, composite_px on a copy
    uint32_t sw[4] = {0, 0, 0, 0};
#pragma unroll
    for (int ch = 0; ch < BPP; ++ch) sw[ch] = ch8_widen((clr >> (8 * ch)) & 0xFFu);
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        if (x + j >= xe) break;
        uint8_t px[BPP];
#pragma unroll
        for (int ch = 0; ch < BPP; ++ch) px[ch] = FRESH ? 0 : dst[j * BPP + ch];
        const uint32_t a = cov_of<RULE>(base + v[j]);
        if (FRESH)
            composite_px_fresh<BPP>(px, sw, a);
        else
            composite_px<BPP>(px, sw, a);
#pragma unroll
        for (int ch = 0; ch < BPP; ++ch) dst[j * BPP + ch] = px[ch];
    }
}

// Prefix-sum the cells [xb, xe) of one raster row (one wavefront) from
// `carry`, map coverage and store / composite CPT pixels per lane per chunk.
// `loaded`: the pipe already holds the segment's first DST_DEPTH chunks.
template <int FMT, int CPT, uint32_t RULE, bool FRESH>
__device__ __forceinline__ void row_pass(const int32_t *cells, uint32_t xb, uint32_t xe, gbyte *row,
                                         const uint4 *lut, uint32_t clr, uint32_t carry, DstPipe &pp, bool loaded) {
    constexpr bool RMW = FMT != FMT_MATTE8 && !FRESH;
    constexpr uint32_t chunk = 64u * CPT;
    const uint32_t nck = (xe - xb + chunk - 1) / chunk;
    if constexpr (RMW) {
        if (!loaded) {
            pp.q0 = dst_load<FMT, CPT>(row, xb, xe, 0, nck);
            pp.q1 = dst_load<FMT, CPT>(row, xb, xe, 1, nck);
            pp.q2 = dst_load<FMT, CPT>(row, xb, xe, 2, nck);
            pp.q3 = dst_load<FMT, CPT>(row, xb, xe, 3, nck);
        }
    }
    const uint32_t lx = (uint32_t)(threadIdx.x & 63) * CPT;
    for (uint32_t k = 0; k < nck; k += DST_DEPTH) {
        const uint32_t x0 = xb + k * chunk + lx;
        row_chunk<FMT, CPT, RULE, FRESH>(cells, x0, xe, row, lut, clr, carry, pp.q0);
        if constexpr (RMW) pp.q0 = dst_load<FMT, CPT>(row, xb, xe, k + 4, nck);
        if (k + 1 < nck) {
            row_chunk<FMT, CPT, RULE, FRESH>(cells, x0 + chunk, xe, row, lut, clr, carry, pp.q1);
            if constexpr (RMW) pp.q1 = dst_load<FMT, CPT>(row, xb, xe, k + 5, nck);
        }
        if (k + 2 < nck) {
            row_chunk<FMT, CPT, RULE, FRESH>(cells, x0 + 2 * chunk, xe, row, lut, clr, carry, pp.q2);
            if constexpr (RMW) pp.q2 = dst_load<FMT, CPT>(row, xb, xe, k + 6, nck);
        }
        if (k + 3 < nck) {
            row_chunk<FMT, CPT, RULE, FRESH>(cells, x0 + 3 * chunk, xe, row, lut, clr, carry, pp.q3);
            if constexpr (RMW) pp.q3 = dst_load<FMT, CPT>(row, xb, xe, k + 7, nck);
        }
    }
}

template <int FMT, int CPT, bool FRESH>
__device__ __forceinline__ void row_pass_rule(uint32_t rule, const int32_t *cells, uint32_t xb, uint32_t xe,
                                              gbyte *row, const uint4 *lut, uint32_t clr, uint32_t carry,
                                              DstPipe &pp, bool loaded) {
    if (rule)
        row_pass<FMT, CPT, 1, FRESH>(cells, xb, xe, row, lut, clr, carry, pp, loaded);
    else
        row_pass<FMT, CPT, 0, FRESH>(cells, xb, xe, row, lut, clr, carry, pp, loaded);
}

template <int FMT, int LPR>
__device__ __forceinline__ void row_multi_any(uint32_t rule, bool fresh, const int32_t *cells, uint32_t W,
                                              gbyte *row, const uint4 *lut, uint32_t clr, bool act) {
    constexpr int CPT = FMT == FMT_GRAYA8P ? 8 : 4;
    if (fresh & (FMT != FMT_MATTE8)) {
        if (rule)
            row_multi<FMT, CPT, 1, true, LPR>(cells, W, row, lut, clr, act);
        else
            row_multi<FMT, CPT, 0, true, LPR>(cells, W, row, lut, clr, act);
    } else {
        if (rule)
            row_multi<FMT, CPT, 1, false, LPR>(cells, W, row, lut, clr, act);
        else
            row_multi<FMT, CPT, 0, false, LPR>(cells, W, row, lut, clr, act);
    }
}

// Edge sources of a raster group: entry k of a band's list, or point k of a
// path's stored points (every point starts its segment's edge, if any).
struct BandListSrc {
    const uint32_t *list;
    uint32_t off;
    __device__ uint32_t operator()(uint32_t k) const { return list[off + k]; }
};
struct PathEdgeSrc {
    uint32_t base;
    __device__ uint32_t operator()(uint32_t k) const { return base + k; }
};

// A long (edge, row) cell walk that the direct cell pass hands to a whole
// wavefront (SPLIT): the row's set-up, its first cell's index in the
// group's LDS cells and the edge's winding sign.  The queue and its count
// take the front of the block's LDS, which the direct pass does not use and
// the composite table only fills afterwards.
struct LongWalk {
    RowCells c;
    int32_t at, ed;
};
constexpr int32_t LONG_CELLS = 16;  // walks of more cells than this are split over a wavefront
constexpr uint32_t LONG_QUEUE = (LUT_BYTES - sizeof(uint4)) / sizeof(LongWalk);
static_assert(sizeof(uint4) + LONG_QUEUE * sizeof(LongWalk) <= LUT_BYTES, "the queue fits every format's front");
constexpr uint32_t SPLIT_LIST = 2048;  // list entries a SPLIT cell pass filters at a time
// LDS a SPLIT raster group needs beyond raster_lds_bytes: the filtered list
constexpr size_t SPLIT_LIST_BYTES = SPLIT_LIST * sizeof(uint32_t);

// Rows [y0, y1) of path P's raster (one LDS group): the cnt listed edges'
// (edge, row) cell deltas into LDS (i32 atomics), the wave row scan, the rule
// map and the store / SrcOver.  Every thread of the block calls it with the
// same arguments.  Dynamic LDS: front_bytes<FMT>() for the cell pass's edge
// slots and then the composite table (Rgba8p, Graya8p), followed by the
// rows' i32 cells.
// n_pairs / n_adds: the calling thread's running counts of the (edge, row)
// pairs it walked and of the nonzero cell deltas it added (LDS atomics).
// SPLIT (small figures: few edges, rows of any width): the cell pass first
// lists the entries whose rows meet the group's (SPLIT_LIST at a time, in
// SPLIT_LIST_BYTES of LDS after the P.group_rows rows of cells), then each
// thread takes (listed edge, group row) pairs, so the work spreads over the
// block whatever the edges' sizes; walks longer than LONG_CELLS cells are
// queued (up to LONG_QUEUE of them; a lane walks the rest itself) and every
// wavefront then takes queued walks, one cell per lane (row_cell_delta).
// The listed order does not matter: the cell adds are integer adds.
template <bool PF, int FMT, class Src, bool SPLIT = false>
__device__ __forceinline__ void raster_group(const Limits &L, const PathDev &P, const PathState &S, int fwd,
                                             uint32_t y0, uint32_t y1, Src src, uint32_t cnt, bool list_ok,
                                             const EdgeRec *edges, uint4 *s_dyn, uint32_t *s_segsum,
                                             DevStatus *st, uint32_t &n_pairs_acc, uint32_t &n_adds_acc) {
    constexpr bool HAS_LUT = FMT != FMT_MATTE8;
    constexpr uint32_t CR = cell_round<FMT>();
    uint4 *const lut = s_dyn;
    EdgeRec *const s_edge = reinterpret_cast<EdgeRec *>(s_dyn);
    uint32_t *const s_pair0 = reinterpret_cast<uint32_t *>(s_edge + CR);
    int32_t *const s_cells = reinterpret_cast<int32_t *>(s_dyn + front_bytes<FMT>() / sizeof(uint4));
    const uint32_t H = P.height, W = P.width;
    const bool has_pts = S.topkey != NO_KEY && S.topidx != NO_IDX;
    const int32_t top_row = has_pts ? top_row_of(S.topkey) : 0;
    const int64_t first = !has_pts ? (int64_t)H : (top_row > 0 ? top_row : 0);
    constexpr uint32_t fmt = FMT;
    if (P.fmt != fmt) {  // the block table groups paths by format
        if (threadIdx.x == 0) raise_err(st, FT_EDEVICE);
        return;
    }
    const uint32_t bpp = fmt == FMT_RGBA8P ? 4 : (fmt == FMT_GRAYA8P ? 2 : 1);
    gbyte *const base = (gbyte *)(P.dst);
    const bool fresh = (P.flags & PATH_FRESH) != 0;
    if (fresh) {  // rows above the figure stay as a fresh raster: zero
        const uint32_t ze = first < (int64_t)y1 ? (uint32_t)first : y1;
        for (uint32_t r = y0; r < ze; ++r) {
            gbyte *row = base + (size_t)r * P.pitch;
            for (uint32_t x = threadIdx.x; x < W * bpp; x += BLOCK) row[x] = 0;
        }
    }
    const uint32_t ry0 = first > (int64_t)y0 ? (uint32_t)first : y0;
    if ((int64_t)ry0 >= (int64_t)y1) return;
    const uint32_t Wp = cell_stride(W);
    const uint32_t rows = y1 - ry0;
    // Row pass assignment.  Groups of at least 4 rows give each wavefront
    // whole rows; a group of 1-3 rows splits every row into segments, one per
    // wavefront, whose cell sums are exchanged through LDS as carry-in.
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t cpt = fmt == FMT_MATTE8 ? (W >= 1024 ? 16u : 4u) : (fmt == FMT_RGBA8P ? 4u : 8u);
    const uint32_t seg = rows >= NW ? 1u : NW / rows;
    const uint32_t chunk = 64u * cpt;
    const uint32_t seg_len = seg == 1 ? W : ((W + seg - 1) / seg + chunk - 1) / chunk * chunk;
    uint32_t r0 = wid, rstep = NW, xb = 0, xe = W;
    if (seg > 1) {
        r0 = wid / seg;
        rstep = rows;  // one row per wavefront group
        xb = (wid % seg) * seg_len;
        xe = xb + seg_len < W ? xb + seg_len : W;
    }
    // the first destination chunks of a read-modify-write format, now
    DstPipe pp;
    pp.q0 = pp.q1 = pp.q2 = pp.q3 = make_uint4(0, 0, 0, 0);
    const bool rmw = !fresh & (fmt != FMT_MATTE8);
    // Narrow rows: LPR lanes per row, 64 / LPR rows per wavefront at once.
    const uint32_t lpr = seg > 1 ? 64u : (W <= 16u * cpt ? 16u : (W <= 32u * cpt ? 32u : 64u));
    const bool pre = PF & rmw & (xb < xe) & (r0 < rows) & (lpr == 64u);
    if constexpr (PF && FMT != FMT_MATTE8) if (pre) {
        const gbyte *row = base + (size_t)(ry0 + r0) * P.pitch;
        const uint32_t nck = (xe - xb + chunk - 1) / chunk;
        if constexpr (FMT == FMT_RGBA8P) {
            pp.q0 = dst_load<FMT_RGBA8P, 4>(row, xb, xe, 0, nck);
            pp.q1 = dst_load<FMT_RGBA8P, 4>(row, xb, xe, 1, nck);
            pp.q2 = dst_load<FMT_RGBA8P, 4>(row, xb, xe, 2, nck);
            pp.q3 = dst_load<FMT_RGBA8P, 4>(row, xb, xe, 3, nck);
        } else {
            pp.q0 = dst_load<FMT_GRAYA8P, 8>(row, xb, xe, 0, nck);
            pp.q1 = dst_load<FMT_GRAYA8P, 8>(row, xb, xe, 1, nck);
            pp.q2 = dst_load<FMT_GRAYA8P, 8>(row, xb, xe, 2, nck);
            pp.q3 = dst_load<FMT_GRAYA8P, 8>(row, xb, xe, 3, nck);
        }
    }
    for (uint32_t c = threadIdx.x; c < rows * Wp / 4; c += BLOCK)
        reinterpret_cast<int4 *>(s_cells)[c] = make_int4(0, 0, 0, 0);
    uint32_t *const q_n = reinterpret_cast<uint32_t *>(s_dyn);
    LongWalk *const q = reinterpret_cast<LongWalk *>(s_dyn + 1);
    if (SPLIT && threadIdx.x == 0) *q_n = 0;
    __syncthreads();
    const int32_t shift = top_row < 0 ? top_row : 0;
    if (!list_ok && threadIdx.x == 0) raise_err(st, FT_EDEVICE);
    // Cell pass, taller groups: the block takes CR list entries at a time; their
    // (edge, row) pairs are laid out by a scan of the rows per edge and spread
    // evenly over the threads (each finds its pair's edge by binary search in
    // LDS), so an edge crossing many rows does not hold up a wavefront.  A
    // pair's row is set up in closed form (row_setup) and its cells walked.
    // Groups of at most DIRECT_ROWS rows (wide rasters) take four entries per
    // thread and walk each edge's rows in the lane: an edge there crosses at
    // most that many of the group's rows.
    constexpr int EB = 4;
    const bool direct = rows <= DIRECT_ROWS;
    if constexpr (SPLIT) {
        uint32_t *const s_list = reinterpret_cast<uint32_t *>(s_cells + (size_t)P.group_rows * Wp);
        uint32_t *const s_nl = q_n + 1;
        for (uint32_t k0 = 0; list_ok && k0 < cnt; k0 += SPLIT_LIST) {
            if (threadIdx.x == 0) *s_nl = 0;
            __syncthreads();
            const uint32_t kn = cnt - k0 < SPLIT_LIST ? cnt - k0 : SPLIT_LIST;
            for (uint32_t k = threadIdx.x; k < kn; k += BLOCK) {
                const uint32_t ei = src(k0 + k);
                const bool in = ei < L.cap_points;
                if (!in) raise_err(st, FT_EDEVICE);
                const EdgeRec E = edges[in ? ei : 0u];
                const int64_t a = (int64_t)fx_int(E.y_upper) - shift, b = (int64_t)fx_int(E.y_lower) - shift;
                if (in & ((E.flags & 1u) != 0) & (a < (int64_t)y1) & (b >= (int64_t)ry0))
                    s_list[atomicAdd(s_nl, 1u)] = ei;
            }
            __syncthreads();
            const uint32_t n_pairs = *s_nl * rows;
            for (uint32_t p = threadIdx.x; p < n_pairs; p += BLOCK) {
                const EdgeRec E = edges[s_list[p / rows]];
                const int64_t r = (int64_t)ry0 + p % rows;
                if ((r < (int64_t)fx_int(E.y_upper) - shift) | (r > (int64_t)fx_int(E.y_lower) - shift)) continue;
                const int32_t ed = ((int)((E.flags >> 1) & 1u) == fwd) ? 1 : -1;
                const RowCells c = row_setup(E, (int32_t)(r + shift), (int32_t)W);
                int32_t *const at = s_cells + (size_t)(r - ry0) * Wp + c.x_s;
                ++n_pairs_acc;
                if (row_cells_exceed(c, LONG_CELLS)) {
                    const uint32_t qi = atomicAdd(q_n, 1u);
                    if (qi < LONG_QUEUE) {
                        LongWalk w;
                        w.c = c;
                        w.at = (int32_t)((uint32_t)(r - ry0) * Wp + (uint32_t)c.x_s);
                        w.ed = ed;
                        q[qi] = w;
                        continue;
                    }
                }
                row_walk(c, [&](int32_t jj, int32_t d) {
                    if (d) {
                        atomicAdd(&at[jj], d * ed);
                        ++n_adds_acc;
                    }
                });
            }
            __syncthreads();  // before the next entries rewrite the list
        }
    }
    for (uint32_t k0 = 0; !SPLIT && direct && list_ok && k0 < cnt; k0 += (uint32_t)BLOCK * EB) {
        uint32_t ei[EB];
        bool ok[EB];
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            const uint32_t k = k0 + (uint32_t)j * BLOCK + threadIdx.x;
            ok[j] = k < cnt;
            ei[j] = src(ok[j] ? k : 0u);
        }
        EdgeRec E[EB];
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            if (ok[j] & (ei[j] >= L.cap_points)) raise_err(st, FT_EDEVICE);
            ok[j] = ok[j] & (ei[j] < L.cap_points);
            E[j] = edges[ok[j] ? ei[j] : 0u];
            ok[j] = ok[j] & ((E[j].flags & 1u) != 0);  // a point without an edge (path edge lists)
        }
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            if (!ok[j]) continue;
            const int32_t ed = ((int)((E[j].flags >> 1) & 1u) == fwd) ? 1 : -1;
            int64_t a = (int64_t)fx_int(E[j].y_upper) - shift, b = (int64_t)fx_int(E[j].y_lower) - shift;
            a = a < (int64_t)ry0 ? (int64_t)ry0 : a;
            b = b > (int64_t)y1 - 1 ? (int64_t)y1 - 1 : b;
            for (int64_t r = a; r <= b; ++r) {  // edge_row_cells, counted
                const RowCells c = row_setup(E[j], (int32_t)(r + shift), (int32_t)W);
                int32_t *const at = s_cells + (size_t)(r - ry0) * Wp + c.x_s;
                ++n_pairs_acc;
                row_walk(c, [&](int32_t jj, int32_t d) {
                    if (d) {
                        atomicAdd(&at[jj], d * ed);
                        ++n_adds_acc;
                    }
                });
            }
        }
    }
    for (uint32_t k0 = 0; !SPLIT && !direct && list_ok && k0 < cnt; k0 += CR) {
        const uint32_t k = k0 + threadIdx.x;
        bool okk = (threadIdx.x < CR) & (k < cnt);
        const uint32_t ei = src(okk ? k : 0u);
        if (okk & (ei >= L.cap_points)) raise_err(st, FT_EDEVICE);
        okk = okk & (ei < L.cap_points);
        const EdgeRec E = edges[okk ? ei : 0u];
        okk = okk & ((E.flags & 1u) != 0);  // a point without an edge (path edge lists)
        int64_t ra = (int64_t)fx_int(E.y_upper) - shift, rb = (int64_t)fx_int(E.y_lower) - shift;
        ra = ra < (int64_t)ry0 ? (int64_t)ry0 : ra;
        rb = rb > (int64_t)y1 - 1 ? (int64_t)y1 - 1 : rb;
        const uint32_t nr = (okk & (ra <= rb)) ? (uint32_t)(rb - ra + 1) : 0u;
        const uint32_t incl = wave_incl_u32(nr);
        if (lane == 63) s_segsum[wid] = incl;
        if (threadIdx.x < CR) s_edge[threadIdx.x] = E;
        __syncthreads();
        uint32_t before = 0, n_pairs = 0;
#pragma unroll
        for (int w = 0; w < (int)NW; ++w) {
            before += w < wid ? s_segsum[w] : 0u;
            n_pairs += s_segsum[w];
        }
        if (threadIdx.x < CR) s_pair0[threadIdx.x] = before + incl - nr;  // the edge's first pair
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < n_pairs; p += BLOCK) {
            // the last slot whose first pair is at or before p (slots without
            // rows share their successor's first pair)
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = CR / 2; step; step >>= 1) lo += s_pair0[lo + step] <= p ? step : 0u;
            const EdgeRec Eo = s_edge[lo];
            int32_t r_first = fx_int(Eo.y_upper) - shift;
            r_first = r_first < (int32_t)ry0 ? (int32_t)ry0 : r_first;
            const int32_t row = r_first + (int32_t)(p - s_pair0[lo]);  // raster row in [ry0, y1)
            const int32_t ed = ((int)((Eo.flags >> 1) & 1u) == fwd) ? 1 : -1;
            const RowCells c = row_setup(Eo, row + shift, (int32_t)W);
            int32_t *const cr = s_cells + (size_t)(row - (int32_t)ry0) * Wp + c.x_s;
            ++n_pairs_acc;
            row_walk(c, [&](int32_t j, int32_t d) {
                if (d) {
                    atomicAdd(&cr[j], d * ed);
                    ++n_adds_acc;
                }
            });
        }
        __syncthreads();  // before the next entries overwrite the slots
    }
    if constexpr (SPLIT) {  // the queued long walks, one per wavefront at a time
        __syncthreads();
        const uint32_t nq = *q_n < LONG_QUEUE ? *q_n : LONG_QUEUE;
        for (uint32_t qi = (uint32_t)wid; qi < nq; qi += NW) {
            const LongWalk w = q[qi];
            const int32_t m = row_cell_count(w.c);
            for (int32_t j = lane; j < m; j += 64) {
                const int32_t d = row_cell_delta(w.c, j);
                if (d) {
                    atomicAdd(&s_cells[w.at + j], d * w.ed);
                    ++n_adds_acc;
                }
            }
        }
        __syncthreads();  // the composite table overwrites the queue
    }
    if constexpr (HAS_LUT)
        for (uint32_t a = threadIdx.x; a < LUT_ENTRIES; a += BLOCK) lut[a] = lut_entry(fmt, P.clr, a);
    __syncthreads();
    uint32_t carry = 0;
    if (seg > 1) {
        const uint32_t sgi = wid % seg;
        uint32_t sum = 0;
        if (r0 < rows)
            for (uint32_t x = xb + 4u * (uint32_t)lane; x < xe; x += 256) {
                const int4 c4 = *reinterpret_cast<const int4 *>(s_cells + (size_t)r0 * Wp + x);
                sum += (uint32_t)c4.x + (x + 1 < xe ? (uint32_t)c4.y : 0u) + (x + 2 < xe ? (uint32_t)c4.z : 0u) +
                       (x + 3 < xe ? (uint32_t)c4.w : 0u);
            }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
        if (lane == 0) s_segsum[wid] = sum;
        __syncthreads();
        for (uint32_t k = 0; k < sgi; ++k) carry += s_segsum[r0 * seg + k];
    }
    if (lpr < 64u) {
        const uint32_t rpw = 64u / lpr, sub = (uint32_t)lane / lpr;
        for (uint32_t rb = (uint32_t)wid * rpw; rb < rows; rb += NW * rpw) {
            const uint32_t r = rb + sub;
            const bool act = r < rows;
            const uint32_t rr = act ? r : 0u;
            const int32_t *cells = s_cells + (size_t)rr * Wp;
            gbyte *row = base + (size_t)(ry0 + rr) * P.pitch;
            if (lpr == 16u)
                row_multi_any<FMT, 16>(P.rule, fresh, cells, W, row, lut, P.clr, act);
            else
                row_multi_any<FMT, 32>(P.rule, fresh, cells, W, row, lut, P.clr, act);
        }
        return;
    }
    bool loaded = pre;
    for (uint32_t r = r0; r < rows; r += rstep) {
        if (xb >= xe) break;
        const int32_t *cells = s_cells + (size_t)r * Wp;
        gbyte *row = base + (size_t)(ry0 + r) * P.pitch;
        if constexpr (FMT == FMT_MATTE8) {
            if (cpt == 16)
                row_pass_rule<FMT_MATTE8, 16, false>(P.rule, cells, xb, xe, row, lut, P.clr, carry, pp, false);
            else
                row_pass_rule<FMT_MATTE8, 4, false>(P.rule, cells, xb, xe, row, lut, P.clr, carry, pp, false);
        } else if constexpr (FMT == FMT_RGBA8P) {
            if (fresh)
                row_pass_rule<FMT_RGBA8P, 4, true>(P.rule, cells, xb, xe, row, lut, P.clr, carry, pp, false);
            else
                row_pass_rule<FMT_RGBA8P, 4, false>(P.rule, cells, xb, xe, row, lut, P.clr, carry, pp, loaded);
        } else {
            if (fresh)
                row_pass_rule<FMT_GRAYA8P, 8, true>(P.rule, cells, xb, xe, row, lut, P.clr, carry, pp, false);
            else
                row_pass_rule<FMT_GRAYA8P, 8, false>(P.rule, cells, xb, xe, row, lut, P.clr, carry, pp, loaded);
        }
        loaded = false;
    }
}


// The block's totals of two per-thread counts into stats[2 slot], [2 slot + 1]
// (when stats is given).  Every thread calls it.
__device__ inline void block_stats(uint32_t a, uint32_t b, unsigned long long *stats, uint32_t slot,
                                   uint32_t *s_w) {
    if (!stats) return;
    const uint32_t ia = wave_incl_u32(a), ib = wave_incl_u32(b);
    __syncthreads();
    if ((threadIdx.x & 63) == 63) {
        s_w[2 * (threadIdx.x >> 6)] = ia;
        s_w[2 * (threadIdx.x >> 6) + 1] = ib;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long ta = 0, tb = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            ta += s_w[2 * w];
            tb += s_w[2 * w + 1];
        }
        stats[2 * (size_t)slot] = ta;
        stats[2 * (size_t)slot + 1] = tb;
    }
}

// One block per (band, row group): `rblock[b]` = (global band, group, path);
// one kernel per pixel format FMT, over the format's stretch of the table
// from `first`, ordered largest first.  The block accumulates the band's listed edges for its G rows
// in LDS and scans each row into the raster; the band's other groups are
// other blocks.
template <bool PF, int FMT>
__global__ __launch_bounds__(BLOCK) void k_raster(Limits L0, const PathDev *paths, const uint4 *rblock,
                                                  uint32_t first_block, uint32_t band_shift,
                                                  const uint32_t *band_off, const uint32_t *band_list,
                                                  const EdgeRec *edges, const PathState *ps, DevStatus *st,
                                                  unsigned long long *stats) {
    extern __shared__ uint4 s_dyn[];
    __shared__ uint32_t s_segsum[BLOCK / 64];
    __shared__ uint32_t s_stat[2 * BLOCK / 64];
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    const uint4 rb = rblock[first_block + blockIdx.x];
    const uint32_t gb = rb.x, p = rb.z;
    if ((gb >= L.n_bands) | (p >= L.n_paths)) {
        if (threadIdx.x == 0) raise_err(st, FT_EDEVICE);
        return;
    }
    // the band's list: its replica segments, back to back
    const uint32_t off = band_off[gb << band_shift];
    const uint32_t end = gb + 1 < L.n_bands ? band_off[(gb + 1) << band_shift] : L.n_pairs;
    const PathDev &P = paths[p];
    const uint32_t R = P.band_rows, H = P.height;
    const uint32_t G = P.group_rows ? P.group_rows : 1u;  // rows whose cells fit in LDS at once
    const uint32_t b0 = (gb - P.band_base) * R;
    const uint32_t b1 = b0 + R < H ? b0 + R : H;
    const uint32_t y0 = b0 + rb.y * G;
    const uint32_t y1 = y0 + G < b1 ? y0 + G : b1;
    uint32_t n_pairs = 0, n_adds = 0;
    if (y0 < y1) {
        const PathState S = ps[p];
        const bool list_ok = (off <= end) & (end <= L.n_pairs);
        raster_group<PF, FMT>(L, P, S, (int)S.fwd, y0, y1, BandListSrc{band_list, off},
                              list_ok ? end - off : 0u, list_ok, edges, s_dyn, s_segsum, st, n_pairs, n_adds);
    }
    block_stats(n_pairs, n_adds, stats, first_block + blockIdx.x, s_stat);
}

// The atlas tile scheduler: a persistent grid over the work items (path,
// row group) of the paths laid out in tile mode, ordered largest first;
// block b takes items b, b + grid, b + 2 grid, ..., so every block gets an
// even share of large and small items without a shared queue counter (one
// such counter serialises the blocks' atomics on one cache line).  A
// tile-mode path has no bands: an item reads the path's own stored points as
// its edge list (one edge per point, in point order), so nothing is binned.
// The winding direction (Fig::get_dir) is evaluated per item.
template <bool PF, int FMT>
__global__ __launch_bounds__(BLOCK) void k_tiles(Limits L0, const PathDev *paths, const uint2 *items,
                                                 uint32_t first_item, uint32_t n_items, const EdgeRec *edges,
                                                 const I2 *pts, const uint32_t *pt_run, const uint32_t *run_start,
                                                 const uint32_t *run_end, const PathState *ps, DevStatus *st,
                                                 unsigned long long *stats, uint32_t stats_slot) {
    extern __shared__ uint4 s_dyn[];
    __shared__ uint32_t s_segsum[BLOCK / 64];
    __shared__ uint32_t s_stat[2 * BLOCK / 64];
    __shared__ uint32_t s_fwd;
    if (run_failed(st)) return;
    const Limits L = run_limits(st, L0.n_ops, L0.n_paths, L0.n_bands, L0.cap_points);
    uint32_t n_pairs = 0, n_adds = 0;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        if (threadIdx.x == 0) s_fwd = 0;
        const uint2 w = items[first_item + it];
        const uint32_t p = w.x;
        if (p >= L.n_paths) {
            if (threadIdx.x == 0) raise_err(st, FT_EDEVICE);
            break;
        }
        const PathDev &P = paths[p];
        const PathState S = ps[p];
        const bool has_pts = S.topkey != NO_KEY && S.topidx != NO_IDX;
        if (threadIdx.x == 0 && has_pts) {
            uint32_t bad = 0;
            s_fwd = path_forward(S, L, pts, pt_run, run_start, run_end, bad) ? 1u : 0u;
            if (bad) raise_err(st, FT_EDEVICE);
        }
        const uint32_t G = P.group_rows ? P.group_rows : 1u;
        const uint32_t y0 = w.y * G;
        const uint32_t y1 = y0 + G < P.height ? y0 + G : P.height;
        const bool list_ok = ((uint64_t)S.pt_base + S.stored) <= L.cap_points;
        __syncthreads();  // s_fwd
        if (y0 < y1)
            raster_group<PF, FMT>(L, P, S, (int)s_fwd, y0, y1, PathEdgeSrc{S.pt_base},
                                  (has_pts && list_ok) ? S.stored : 0u, list_ok, edges, s_dyn, s_segsum, st,
                                  n_pairs, n_adds);
        __syncthreads();  // the next item reuses the LDS and s_fwd
    }
    block_stats(n_pairs, n_adds, stats, stats_slot + blockIdx.x, s_stat);
}

// ---------------------------------------------------------------------------
// Small figures: the Plotter's fast path.
// ---------------------------------------------------------------------------
// A fill of at most SMALL_OPS ops whose events fit SMALL_EVENTS runs every
// stage before the raster in one block with its data in LDS: the (op, lane)
// counts and their block scan, the emit, the Fig ingest scans (pushed marks,
// kept points), compaction with the top point, Edge::new per point and the
// winding direction.  It writes the edge list and the path's state; the
// raster then runs as one block per LDS row group (k_small_raster).  Two
// launches, one upload and one readback instead of the pipeline's dozen
// launches.  A stroke does the same with the stroker's stages in between
// (flattened events with widths, stored points, sub-strokes, the outline's
// items), as long as the outline's events and closes fit the same limits.
// Anything the block does not cover (more events or sub-figures, a
// flattening error, an index check) makes it write nothing and report "not
// handled" (the raster blocks then do nothing); the caller runs the full
// pipeline, which gives the same result or the same error.
constexpr int SBLOCK = 1024;                     // threads of k_small_front
constexpr uint32_t SMALL_OPS = 256;
constexpr uint32_t SMALL_ARG_OPS = 128;          // ops passed inside the launch itself (no upload)
constexpr uint32_t SMALL_ITEMS = SMALL_OPS * OP_LANES;
constexpr uint32_t SMALL_EVENTS = 2048;
constexpr uint32_t SMALL_RUNS = SMALL_OPS + 1;    // every op closes at most one sub-figure, plus the last
constexpr uint32_t SMALL_PER_THREAD = SMALL_EVENTS / SBLOCK;
constexpr uint32_t SMALL_NOT_HANDLED = 0x10u;     // DevStatus::overflow bit
constexpr uint32_t SMALL_RASTER_BLOCKS = 64;      // raster blocks a small figure aims for
static_assert(SMALL_ITEMS == SMALL_EVENTS, "op lanes and events share the scan shape");
static_assert(SMALL_EVENTS <= SPLIT_LIST, "a small figure's edges are listed in one pass");

struct SmallLds {
    union {
        struct {  // events: flatten to compaction
            I2 xy[SMALL_EVENTS];
            uint32_t meta[SMALL_EVENTS], last[SMALL_EVENTS], kept[SMALL_EVENTS], idx[SMALL_EVENTS];
        } ev;
        struct {  // a stroke's flattened (x, y, width, kind) events and its last close so far
            F4 f[SMALL_EVENTS];
            uint32_t last[SMALL_EVENTS];
        } st;
        EdgeRec edges[SMALL_EVENTS];  // after compaction
    } a;
    union {
        uint32_t cnt_e[SMALL_ITEMS];   // (op, lane) event counts, then their exclusive scan
        I2 pts[SMALL_EVENTS];          // after the emit
    } b;
    union {
        uint32_t cnt_c[SMALL_ITEMS];   // (op, lane) close counts, then their exclusive scan
        uint32_t pt_run[SMALL_EVENTS];
    } c;
    int32_t run_first[SMALL_RUNS];
    uint32_t run_close[SMALL_RUNS], run_path[SMALL_RUNS], run_start[SMALL_RUNS], run_end[SMALL_RUNS];
    uint32_t op_path[SMALL_OPS];
    uint32_t op_item[SMALL_OPS];  // the op's first (op, lane) item
    ft_path_op ops[SMALL_OPS];
    PathDev path;
    PathState ps;
    uint32_t w[SBLOCK / 64], w2[SBLOCK / 64];
    unsigned long long wk[SBLOCK / 64];
    uint32_t flag;
};

// Block-wide exclusive scan (sum, or inclusive max when MAX) of a[0, n) in
// place, n <= SBLOCK * SMALL_PER_THREAD, each thread taking consecutive
// items; returns the total (or the maximum).  Every thread calls it.
template <bool MAX>
__device__ uint32_t small_scan(uint32_t *a, uint32_t n, uint32_t *s_w) {
    const uint32_t t0 = threadIdx.x * SMALL_PER_THREAD;
    uint32_t v[SMALL_PER_THREAD], run = 0;
#pragma unroll
    for (uint32_t j = 0; j < SMALL_PER_THREAD; ++j) {
        v[j] = t0 + j < n ? a[t0 + j] : 0u;
        run = MAX ? (v[j] > run ? v[j] : run) : run + v[j];
    }
    const uint32_t incl = wave_incl_dpp<MAX>(run);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) s_w[wid] = incl;
    __syncthreads();
    uint32_t pre = 0, total = 0;
#pragma unroll
    for (int w = 0; w < SBLOCK / 64; ++w) {
        const uint32_t x = s_w[w];
        if (MAX) {
            pre = (w < wid && x > pre) ? x : pre;
            total = x > total ? x : total;
        } else {
            pre += w < wid ? x : 0u;
            total += x;
        }
    }
    // the thread's own prefix: the block's part, then the wave's before it
    uint32_t acc;
    if (MAX) {
        const uint32_t prev = (uint32_t)__shfl_up((int)incl, 1, 64);
        acc = (threadIdx.x & 63) ? (prev > pre ? prev : pre) : pre;
    } else {
        acc = pre + incl - run;
    }
#pragma unroll
    for (uint32_t j = 0; j < SMALL_PER_THREAD; ++j) {
        if (MAX) {
            acc = v[j] > acc ? v[j] : acc;
            if (t0 + j < n) a[t0 + j] = acc;  // inclusive max
        } else {
            if (t0 + j < n) a[t0 + j] = acc;  // exclusive sum
            acc += v[j];
        }
    }
    __syncthreads();
    return total;
}

// Ops of a small figure carried inside the launch (n_ops <= SMALL_ARG_OPS).
struct SmallOps {
    ft_path_op op[SMALL_ARG_OPS];
};

// A small stroke's stroker data beside SmallLds: the pen-width marks of the
// ops and their running maximum, the stored points, the sub-strokes, and
// the first ITEM_PTS (transformed) outline points of every outline item,
// which its count computes and its emit then replays.
constexpr uint32_t ITEM_PTS = 2;
struct SmallStrokeLds {
    F4 spts[SMALL_EVENTS];
    uint32_t lastpw[SMALL_OPS];
    uint32_t sub_start[SMALL_RUNS], sub_event[SMALL_RUNS], sub_n[SMALL_RUNS], sub_joined[SMALL_RUNS];
    SV2 item_pt[ITEM_PTS][SMALL_EVENTS];
};

// Fill events of a small figure go to the event and run arrays in LDS.
__device__ inline EventSink small_fill_sink(SmallLds &S) {
    EventSink em{};
    em.mode = MODE_FILL;
    em.cap_events = SMALL_EVENTS;
    em.cap_runs = SMALL_RUNS;
    em.ev_xy = S.a.ev.xy;
    em.ev_meta = S.a.ev.meta;
    em.run_first = S.run_first;
    em.run_close = S.run_close;
    em.run_path = S.run_path;
    em.run_start = S.run_start;
    em.run_end = S.run_end;
    return em;
}

// The op of (op, lane) item t: the last op whose first item is at or before
// t (every op has at least one item).
__device__ inline uint32_t small_item_op(const uint32_t *op_item, uint32_t n_ops, uint32_t t) {
    uint32_t lo = 0, hi = n_ops;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (op_item[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

constexpr uint32_t SMALL_SPLIT = 5;  // a small figure's curves take 32 lanes when the items fit

// Stages 1-2 in one block: the (op, lane) event and close counts, their block
// scans, and the emit through `em`.  A curve's subdivision tree is split
// SMALL_SPLIT levels down over 2^SMALL_SPLIT lanes (FLAT_SPLIT if those
// items would not fit SMALL_ITEMS), every other op takes one lane: the
// points and their order are the same for any split.  False (in every
// thread, nothing emitted) when flattening fails or the events or closes
// exceed their limits.
// A small fill keeps each curve lane's first SMALL_LEAF points in LDS
// (leaf[k * SMALL_ITEMS + t], with the lane's leaf count in n_leaf[t]), so
// that its emit is a copy unless the lane has more.
constexpr uint32_t SMALL_LEAF = 2;

__device__ bool small_flatten(SmallLds &S, uint32_t n_ops, uint32_t mode, const uint32_t *lastpw, EventSink em,
                              uint32_t max_runs, uint32_t &n_events, uint32_t &n_runs, uint32_t &bad,
                              I2 *leaf = nullptr, uint16_t *n_leaf = nullptr) {
    const uint32_t tid = threadIdx.x;
    static_assert(SMALL_OPS <= SBLOCK, "one op per thread");
    const uint32_t tag = tid < n_ops ? S.ops[tid].tag : TAG_CLOSE;
    const bool curve = (tag == TAG_QUAD) | (tag == TAG_CUBIC);
    const uint32_t n_curves = (uint32_t)__syncthreads_count(curve ? 1 : 0);
    const uint32_t sp = n_curves * ((1u << SMALL_SPLIT) - 1u) + n_ops <= SMALL_ITEMS ? SMALL_SPLIT : FLAT_SPLIT;
    if (tid < SMALL_OPS) S.op_item[tid] = tid < n_ops ? (curve ? 1u << sp : 1u) : 0u;
    __syncthreads();
    const uint32_t n_items = small_scan<false>(S.op_item, n_ops, S.w);
    // item t's op, lane and the op's last lane
    auto lane_of = [&](uint32_t t, uint32_t &i, uint32_t &s, uint32_t &last) {
        i = small_item_op(S.op_item, n_ops, t);
        const uint32_t first = S.op_item[i];
        const uint32_t end = i + 1 < n_ops ? S.op_item[i + 1] : n_items;
        s = t - first;
        last = end - first - 1;
    };
    uint32_t err = 0;
    for (uint32_t t = tid; t < SMALL_ITEMS; t += SBLOCK) {
        OpCount c = {0, 0, 0, 0, 0};
        if (t < n_items) {
            uint32_t i, s, last;
            lane_of(t, i, s, last);
            c = count_op(S.ops, S.op_path, &S.path, i, s, mode, lastpw, leaf ? leaf + t : nullptr, SMALL_ITEMS, sp,
                         last, SMALL_LEAF);
        }
        if (n_leaf) n_leaf[t] = (uint16_t)(c.leaves < 0xFFFFu ? c.leaves : 0xFFFFu);
        err |= c.err;
        S.b.cnt_e[t] = c.ne;
        S.c.cnt_c[t] = c.nc;
    }
    if (err) S.flag = 1;
    __syncthreads();
    n_events = small_scan<false>(S.b.cnt_e, n_items, S.w);
    n_runs = small_scan<false>(S.c.cnt_c, n_items, S.w2);
    if (S.flag | (n_events > SMALL_EVENTS) | (n_runs > max_runs)) return false;
    for (uint32_t t = tid; t < n_items; t += SBLOCK) {
        const unsigned long long off = ((unsigned long long)S.b.cnt_e[t] << 32) | S.c.cnt_c[t];
        uint32_t i, s, last;
        lane_of(t, i, s, last);
        bad |= emit_op(S.ops, S.op_path, &S.path, i, s, mode, lastpw, off, em, leaf ? leaf + t : nullptr,
                       n_leaf ? n_leaf[t] : 0u, SMALL_ITEMS, sp, last, SMALL_LEAF);
    }
    __syncthreads();
    return true;
}

// The stroker (the k_st_* stages) in one block, from the flattened events in
// S.a.st.f to the outline's fill events and runs.  False (in every thread)
// when the sub-strokes, outline items, events or closes exceed their limits
// or a join fails to flatten; nothing usable is written then.
__device__ bool small_stroker(SmallLds &S, SmallStrokeLds &T, uint32_t n_sev, StrokeParams sp, uint32_t &n_events,
                              uint32_t &n_runs, uint32_t &bad) {
    const uint32_t tid = threadIdx.x;
    // flags (k_st_flags): pushed -> stored point index, first point of a run
    // -> sub-stroke index, closes -> the last close so far
    for (uint32_t e = tid; e < n_sev; e += SBLOCK) {
        const StrokeFlags f = stroke_flags(e, S.a.st.f);
        S.b.cnt_e[e] = f.pushed ? 1u : 0u;
        S.c.cnt_c[e] = f.first ? 1u : 0u;
        S.a.st.last[e] = f.is_pt ? 0u : e + 1;
    }
    __syncthreads();
    const uint32_t n_pts = small_scan<false>(S.b.cnt_e, n_sev, S.w);  // below the 65,535 cap
    const uint32_t n_subs = small_scan<false>(S.c.cnt_c, n_sev, S.w2);
    (void)small_scan<true>(S.a.st.last, n_sev, S.w);
    if (n_subs > SMALL_RUNS) return false;
    // compaction (k_st_compact)
    for (uint32_t e = tid; e < n_sev; e += SBLOCK) {
        const StrokeFlags f = stroke_flags(e, S.a.st.f);
        const uint32_t sidx = S.b.cnt_e[e], sub = S.c.cnt_c[e];
        if (f.pushed) T.spts[sidx] = S.a.st.f[e];
        if (f.first) {
            T.sub_start[sub] = sidx;
            T.sub_event[sub] = e;
        }
    }
    __syncthreads();
    // sub-strokes (k_st_subs; every point is kept)
    for (uint32_t s = tid; s < n_subs; s += SBLOCK) {
        const uint32_t a = T.sub_start[s];
        const bool more = s + 1 < n_subs;
        const uint32_t end = more ? T.sub_start[s + 1] : n_pts;
        const uint32_t range_end = more ? T.sub_event[s + 1] : n_sev;
        const uint32_t lc = S.a.st.last[range_end ? range_end - 1 : 0];
        const bool has_close = (range_end > 0) & (lc > 0) & (lc - 1 >= T.sub_event[s]) & (lc - 1 < n_sev);
        const F4 ce = S.a.st.f[has_close ? lc - 1 : 0];
        T.sub_n[s] = end > a ? end - a : 0u;
        T.sub_joined[s] = (has_close & (ce.w == 2.0f)) ? 1u : 0u;
    }
    const uint32_t n_items = n_subs ? 2u * (n_pts + n_subs) + 1u : 1u;
    if (n_items > SMALL_EVENTS) return false;
    __syncthreads();
    // outline items: line and close counts (k_st_count)
    auto sub_of = [&](uint32_t t, StrokeSub &sub, bool &fwd, uint32_t &k) {
        const uint32_t si = st_item_sub(t, T.sub_start, n_subs);
        sub.start = T.sub_start[si];
        sub.n = T.sub_n[si];
        sub.joined = T.sub_joined[si] != 0;
        const uint32_t local = t - 2u * (sub.start + si);
        fwd = local <= sub.n;
        k = fwd ? local : local - (sub.n + 1);
    };
    uint32_t err = 0;
    for (uint32_t t = tid; t < n_items; t += SBLOCK) {
        uint32_t ne = 0, nc = 0;
        if (t + 1 == n_items) {
            ne = nc = 1;  // Plotter::fill's final close
        } else if (n_subs) {
            StrokeSub sub;
            bool fwd;
            uint32_t k;
            sub_of(t, sub, fwd, k);
            // an item's points all come before its close (at most one)
            if (sub.n)
                err |= stroke_item<DevMath>(
                    sub, fwd, k, T.spts, sp.join, sp.miter, sp.tol_sq,
                    [&](SV2 q) {
                        if (ne < ITEM_PTS) {
                            const WP w = xform(S.path, q.x, q.y, 0.0f);
                            T.item_pt[ne][t].x = w.x;
                            T.item_pt[ne][t].y = w.y;
                        }
                        ++ne;
                    },
                    [&]() {
                        ++ne;
                        ++nc;
                    });
        }
        S.b.cnt_e[t] = ne;
        S.c.cnt_c[t] = nc;
    }
    if (err) S.flag = 1;
    __syncthreads();
    n_events = small_scan<false>(S.b.cnt_e, n_items, S.w);
    n_runs = small_scan<false>(S.c.cnt_c, n_items, S.w2);
    if (S.flag | (n_events > SMALL_EVENTS) | (n_runs > SMALL_RUNS)) return false;
    // the outline's events (k_st_emit), transformed again as Plotter::fill
    // does: an item of at most ITEM_PTS points replays what its count kept
    const EventSink sink = small_fill_sink(S);
    for (uint32_t t = tid; t < n_items; t += SBLOCK) {
        EventSink em = sink;
        em.e = S.b.cnt_e[t];
        em.run = S.c.cnt_c[t];
        em.p = 0;
        const uint32_t e1 = t + 1 < n_items ? S.b.cnt_e[t + 1] : n_events;
        const uint32_t r1 = t + 1 < n_items ? S.c.cnt_c[t + 1] : n_runs;
        const uint32_t n_cl = r1 - (uint32_t)em.run, n_pt = (e1 - (uint32_t)em.e) - n_cl;
        if (t + 1 == n_items) {
            em.close(true);
        } else if (n_pt <= ITEM_PTS) {
            for (uint32_t j = 0; j < n_pt; ++j) {
                WP w;
                w.x = T.item_pt[j][t].x;
                w.y = T.item_pt[j][t].y;
                w.w = 0.0f;
                em.point(w);
            }
            if (n_cl) em.close(true);
        } else if (n_subs) {
            StrokeSub sub;
            bool fwd;
            uint32_t k;
            sub_of(t, sub, fwd, k);
            if (sub.n)
                stroke_item<DevMath>(
                    sub, fwd, k, T.spts, sp.join, sp.miter, sp.tol_sq,
                    [&](SV2 q) { em.point(xform(S.path, q.x, q.y, 0.0f)); }, [&]() { em.close(true); });
        }
        bad |= em.bad;
    }
    __syncthreads();
    return true;
}

// ops: the figure's ops in device memory, or null when `inl` carries them.
// STROKE: Plotter::stroke (flatten with widths, the stroker, then the
// outline's fill); else Plotter::fill.
template <bool STROKE>
__global__ __launch_bounds__(SBLOCK) void k_small_front(const ft_path_op *ops, SmallOps inl, uint32_t n_ops,
                                                        PathDev P, StrokeParams sp, EdgeRec *edges, PathState *ps,
                                                        DevStatus *st) {
    __shared__ SmallLds S;
    const uint32_t tid = threadIdx.x;
    if (tid < sizeof(DevStatus) / 4) reinterpret_cast<uint32_t *>(st)[tid] = 0;
    if (tid == 0) {
        S.path = P;
        S.flag = 0;
    }
    const uint32_t n_in = n_ops < SMALL_OPS ? n_ops : SMALL_OPS;
    for (uint32_t i = tid; i < SMALL_OPS; i += SBLOCK) S.op_path[i] = 0;
    for (uint32_t i = tid; i < n_in; i += SBLOCK) S.ops[i] = ops ? ops[i] : inl.op[i < SMALL_ARG_OPS ? i : 0];
    __syncthreads();
    // 1-2. flatten (and for a stroke, the stroker) into the fill events in LDS
    uint32_t n_events = 0, n_runs = 0, bad = 0;
    bool ok;
    if constexpr (STROKE) {
        __shared__ SmallStrokeLds T;
        // the last PenWidth op up to each op (a running maximum of i + 1); the
        // pipeline's exclusive form differs only at PenWidth ops, whose width
        // is never asked for
        for (uint32_t i = tid; i < SMALL_OPS; i += SBLOCK)
            T.lastpw[i] = (i < n_in && S.ops[i].tag == TAG_PEN_WIDTH) ? i + 1 : 0u;
        __syncthreads();
        (void)small_scan<true>(T.lastpw, n_in, S.w);
        EventSink em{};
        em.mode = MODE_STROKE;
        em.cap_events = SMALL_EVENTS;
        em.cap_runs = SMALL_EVENTS;
        em.ev_f = S.a.st.f;
        uint32_t n_sev = 0, n_scl = 0;
        ok = small_flatten(S, n_ops, MODE_STROKE, T.lastpw, em, SMALL_EVENTS, n_sev, n_scl, bad) &&
             small_stroker(S, T, n_sev, sp, n_events, n_runs, bad);
    } else {
        __shared__ I2 s_leaf[SMALL_LEAF * SMALL_ITEMS];
        __shared__ uint16_t s_nleaf[SMALL_ITEMS];
        ok = small_flatten(S, n_ops, MODE_FILL, nullptr, small_fill_sink(S), SMALL_RUNS, n_events, n_runs, bad,
                           s_leaf, s_nleaf);
    }
    if (!ok) {
        if (tid == 0) atomicOr(&st->overflow, SMALL_NOT_HANDLED);
        return;
    }
    Limits L{};
    L.n_ops = n_ops;
    L.n_paths = 1;
    L.n_bands = 0;
    L.n_events = n_events;
    L.n_runs = n_runs;
    L.cap_points = SMALL_EVENTS;
    L.n_pairs = 0;
    // 3. Fig ingest: pushed marks and their running maximum, then kept points
    for (uint32_t e = tid; e < n_events; e += SBLOCK)
        ingest_event(e, L, S.a.ev.meta, S.a.ev.xy, S.run_first, S.run_path, S.a.ev.last, bad);
    __syncthreads();
    (void)small_scan<true>(S.a.ev.last, n_events, S.w);
    for (uint32_t e = tid; e < n_events; e += SBLOCK) {
        const uint32_t k = kept_event(e, L, S.a.ev.meta, S.a.ev.xy, S.run_first, S.run_close, S.a.ev.last, bad);
        S.a.ev.kept[e] = k;
        S.a.ev.idx[e] = k;
    }
    __syncthreads();
    const uint32_t n_kept = small_scan<false>(S.a.ev.idx, n_events, S.w);
    // 4. the path's counts (no cap below 65,535 points) and compaction with
    // the top (y, x) key
    if (tid == 0) {
        PathState ps;
        ps.topkey = NO_KEY;
        ps.topidx = NO_IDX;
        ps.offered = n_events - n_runs;
        ps.repop = 0;
        ps.pt_base = 0;
        ps.n_edges = 0;
        ps.stored = n_kept;
        ps.fwd = 0;
        S.ps = ps;
    }
    __syncthreads();
    unsigned long long kmin = NO_KEY;
    for (uint32_t e = tid; e < n_events; e += SBLOCK) {
        TopCand tc;
        compact_event(e, L, S.a.ev.meta, S.a.ev.xy, S.a.ev.kept, S.a.ev.idx, S.run_first, S.run_path, S.run_start,
                      S.run_end, S.b.pts, S.c.pt_run, &S.ps, bad, &tc);
        if (tc.path != NO_IDX && tc.key < kmin) kmin = tc.key;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long u = __shfl_xor(kmin, o, 64);
        kmin = u < kmin ? u : kmin;
    }
    if ((tid & 63) == 0) S.wk[tid >> 6] = kmin;
    __syncthreads();
    if (tid == 0) {
        unsigned long long m = S.wk[0];
        for (int w = 1; w < SBLOCK / 64; ++w) m = S.wk[w] < m ? S.wk[w] : m;
        S.ps.topkey = m;
    }
    __syncthreads();
    // 5. Edge::new per stored point (the events' LDS now holds the edges)
    EdgeRec E[SMALL_PER_THREAD];
    uint32_t tmin = NO_IDX;
#pragma unroll
    for (uint32_t j = 0; j < SMALL_PER_THREAD; ++j) {
        const uint32_t i = tid + j * SBLOCK;
        E[j].flags = 0;
        if (i < n_kept) {
            const EdgeEval r = edge_eval(i, L, S.c.pt_run, S.run_path, S.run_start, S.run_end, &S.path, S.b.pts,
                                         &S.ps, bad);
            E[j] = r.E;
            if (r.top && r.top_local < tmin) tmin = r.top_local;
        }
    }
    __syncthreads();  // every event read is done: the edges take their LDS
#pragma unroll
    for (uint32_t j = 0; j < SMALL_PER_THREAD; ++j) {
        const uint32_t i = tid + j * SBLOCK;
        if (i < n_kept) S.a.edges[i] = E[j];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t u = (uint32_t)__shfl_xor((int)tmin, o, 64);
        tmin = u < tmin ? u : tmin;
    }
    if ((tid & 63) == 0) S.w[tid >> 6] = tmin;
    if (bad) S.flag = 1;
    __syncthreads();
    if (tid == 0) {
        uint32_t m = S.w[0];
        for (int w = 1; w < SBLOCK / 64; ++w) m = S.w[w] < m ? S.w[w] : m;
        S.ps.topidx = m;
        uint32_t b2 = 0;
        const bool has = S.ps.topkey != NO_KEY && m != NO_IDX;
        S.ps.fwd = has ? (path_forward(S.ps, L, S.b.pts, S.c.pt_run, S.run_start, S.run_end, b2) ? 1u : 0u) : 0u;
        if (b2) S.flag = 1;
    }
    __syncthreads();
    if (S.flag) {  // an index check failed: let the full pipeline run (and report)
        if (tid == 0) atomicOr(&st->overflow, SMALL_NOT_HANDLED);
        return;
    }
    // 6. the edge list and the path's state for the raster blocks
    for (uint32_t i = tid; i < n_kept; i += SBLOCK) edges[i] = S.a.edges[i];
    if (tid == 0) {
        ps[0] = S.ps;
        st->n_events = n_events;
        st->n_closes = n_runs;
        st->n_kept = n_kept;
    }
}

// The raster of a small figure: block b rasters row group b (P.group_rows
// rows) from the edge list k_small_front wrote.
template <int FMT>
__global__ __launch_bounds__(BLOCK) void k_small_raster(PathDev P, const EdgeRec *edges, const PathState *ps,
                                                        DevStatus *st) {
    extern __shared__ uint4 s_dyn[];
    __shared__ uint32_t s_segsum[BLOCK / 64];
    if (run_failed(st)) return;  // not handled, or an error: nothing is drawn
    Limits L{};
    L.n_ops = 0;
    L.n_paths = 1;
    L.n_bands = 0;
    L.n_events = st->n_events;
    L.n_runs = st->n_closes;
    L.cap_points = SMALL_EVENTS;
    L.n_pairs = 0;
    const uint32_t n_kept = st->n_kept;
    const PathState S = ps[0];
    const uint32_t G = P.group_rows ? P.group_rows : 1u;
    const uint32_t y0 = blockIdx.x * G;
    const uint32_t y1 = y0 + G < P.height ? y0 + G : P.height;
    const bool has_pts = S.topkey != NO_KEY && S.topidx != NO_IDX && n_kept <= SMALL_EVENTS;
    uint32_t n_pairs = 0, n_adds = 0;
    if (y0 < y1)
        raster_group<false, FMT, PathEdgeSrc, true>(L, P, S, (int)S.fwd, y0, y1, PathEdgeSrc{0u},
                                                    has_pts ? n_kept : 0u, true, edges, s_dyn, s_segsum, st, n_pairs,
                                                    n_adds);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define FT_HIP(call)                                   \
    do {                                               \
        if ((call) != hipSuccess) return FT_EDEVICE;   \
    } while (0)

// Eager runs whose look-back kernels take work in block-index order (no
// tickets) assume that no other such kernel waits beside them.  Per device,
// the workspace that launched such a run last is remembered; a run of
// another workspace first waits on the host for that one's stream, so runs of
// one workspace follow each other freely and runs of different workspaces
// take turns.
struct BlockOrderChain {
    std::mutex mu;
    const Workspace *last[64] = {};
    hipStream_t stream[64] = {};
};
static BlockOrderChain g_block_order;

static void block_order_forget(const Workspace &ws) {
    std::lock_guard<std::mutex> lock(g_block_order.mu);
    for (int d = 0; d < 64; ++d)
        if (g_block_order.last[d] == &ws) {
            g_block_order.last[d] = nullptr;
            g_block_order.stream[d] = nullptr;
        }
}

static void graph_drop(Workspace &ws) {
    if (ws.gexec) (void)hipGraphExecDestroy(ws.gexec);
    ws.gexec = nullptr;
}

// Arena mode: the next n elements of the arena (256-byte aligned); the arena
// is only rewound by ws_set_batch.
template <typename T>
static ft_status arena_carve(Workspace &ws, T *&ptr, size_t n) {
    constexpr size_t ALIGN = 256;
    const size_t at = (ws.arena_used + ALIGN - 1) / ALIGN * ALIGN;
    const size_t bytes = (n ? n : 1) * sizeof(T);
    if (at > ws.arena_bytes || bytes > ws.arena_bytes - at) {
        ptr = nullptr;
        return FT_ENOMEM;
    }
    ptr = reinterpret_cast<T *>(ws.arena + at);
    ws.arena_used = at + bytes;
    return FT_OK;
}

template <typename T>
static ft_status dev_grow(Workspace &ws, T *&ptr, size_t &cap, size_t need) {
    if (need <= cap && ptr) return FT_OK;
    size_t n = need < 1 ? 1 : need;
    if (ws.arena) {
        const ft_status s = arena_carve(ws, ptr, n);
        cap = s ? 0 : n;
        return s;
    }
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    if (hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)) != hipSuccess) {
        cap = 0;
        return FT_ENOMEM;
    }
    cap = n;
    return FT_OK;
}

template <typename T>
static ft_status dev_alloc(Workspace &ws, T *&ptr, size_t n) {
    if (ws.arena) return arena_carve(ws, ptr, n);
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    if (hipMalloc(reinterpret_cast<void **>(&ptr), (n ? n : 1) * sizeof(T)) != hipSuccess) return FT_ENOMEM;
    return FT_OK;
}

// Blocks of k_flatten the device holds at once (the one-pass flattening is
// used for grids that fit: one wave of blocks, no block waiting for a slot):
// its occupancy per compute unit, at most four, times the CUs.
static ft_status flatten_resident(Workspace &ws) {
    int per_cu = 0, cus = 0;
    FT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_flatten, BLOCK, 0));
    FT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ws.device));
    const uint32_t g = (uint32_t)((cus > 0 ? cus : 0) * (per_cu < 4 ? (per_cu > 0 ? per_cu : 0) : 4));
    ws.flat_resident = g < SCAN_GRID_MAX ? g : SCAN_GRID_MAX;
    return FT_OK;
}

ft_status ws_init_arena(Workspace &ws, void *arena, size_t bytes, hipStream_t stream) {
    if (!arena || (reinterpret_cast<uintptr_t>(arena) & 255u)) return FT_EINVAL;
    FT_HIP(hipGetDevice(&ws.device));
    ws.scan_resident = scan_resident_blocks();
    if (ft_status s = flatten_resident(ws)) return s;
    int lds_max = 0;
    FT_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ws.device));
    ws.lds_max = (size_t)lds_max;
    ws.stream = stream;
    ws.arena = reinterpret_cast<uint8_t *>(arena);
    ws.arena_bytes = bytes;
    ws.arena_used = 0;
    ws.h_status = &ws.h_status_store;
    ws.tickets = true;  // the caller's stream: other waiting kernels may run beside
    return FT_OK;
}

// The status + path-state blocks (device and pinned host) for n_paths paths;
// the previous blocks are freed.
constexpr size_t STATUS_BYTES = 256;
static_assert(sizeof(DevStatus) <= STATUS_BYTES, "the status fits in front of the path states");

static ft_status sp_grow(Workspace &ws, size_t n_paths) {
    const size_t bytes = STATUS_BYTES + n_paths * sizeof(PathState);
    uint8_t *d = nullptr, *h = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&d), bytes) != hipSuccess) return FT_ENOMEM;
    if (hipHostMalloc(reinterpret_cast<void **>(&h), bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(d);
        return FT_ENOMEM;
    }
    memset(h, 0, bytes);
    if (ws.d_sp) (void)hipFree(ws.d_sp);
    if (ws.h_sp) (void)hipHostFree(ws.h_sp);
    ws.d_sp = d;
    ws.h_sp = h;
    ws.cap_sp_paths = n_paths;
    ws.d_status = reinterpret_cast<DevStatus *>(d);
    ws.h_status = reinterpret_cast<DevStatus *>(h);
    ws.d_pstate = reinterpret_cast<PathState *>(d + STATUS_BYTES);
    ws.h_pstate = reinterpret_cast<PathState *>(h + STATUS_BYTES);
    return FT_OK;
}

ft_status ws_init(Workspace &ws) {
    FT_HIP(hipGetDevice(&ws.device));
    ws.scan_resident = scan_resident_blocks();
    if (ft_status s = flatten_resident(ws)) return s;
    int lds_max = 0;
    FT_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ws.device));
    ws.lds_max = (size_t)lds_max;
    FT_HIP(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking));
    return sp_grow(ws, 16);
}

void ws_free(Workspace &ws) {
    block_order_forget(ws);
    if (ws.arena) {  // nothing was allocated; the stream is the caller's
        graph_drop(ws);
        ws = Workspace();
        return;
    }
    // (the layout's upload pointers point into d_lay, the status and path
    // states into d_sp)
    void *bufs[] = {ws.d_lay, ws.d_op_cnt, ws.d_op_off, ws.d_sub_off, ws.d_leaf, ws.d_op_pw, ws.d_op_lastpw,
                    ws.d_ev_xy, ws.d_ev_f, ws.d_ev_meta, ws.d_ev_kept, ws.d_ev_idx, ws.d_ev_last,
                    ws.d_st_flags, ws.d_st_off, ws.d_st_last, ws.d_spts, ws.d_sub_start, ws.d_sub_event, ws.d_sub_n,
                    ws.d_sub_joined, ws.d_it_cnt, ws.d_it_off,
                    ws.d_run_first, ws.d_run_close, ws.d_run_path, ws.d_run_start, ws.d_run_end, ws.d_pts,
                    ws.d_pt_run, ws.d_edges, ws.d_pt_band, ws.d_band_cnt, ws.d_band_off, ws.d_band_cur, ws.d_band_list,
                    ws.d_scan_state, ws.d_top_rep, ws.d_stats, ws.d_sp};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (ws.h_sp) (void)hipHostFree(ws.h_sp);
    if (ws.h_pin) (void)hipHostFree(ws.h_pin);
    if (ws.ev) {
        for (int i = 0; i < 3 * Workspace::TIMING_RING; ++i)
            if (ws.ev[i]) (void)hipEventDestroy(ws.ev[i]);
        delete[] ws.ev;
    }
    graph_drop(ws);
    ws_pipe_detach(ws);
    if (ws.stream && ws.stream_owned) (void)hipStreamDestroy(ws.stream);
    ws = Workspace();
}

ft_status ws_pipe_attach(Workspace &ws, hipStream_t main, hipStream_t front) {
    if (ws.arena || !main) return FT_EINVAL;
    graph_drop(ws);
    if (!ws.fstream) {
        if (front) {
            ws.fstream = front;
            ws.fstream_owned = false;
        } else {
            FT_HIP(hipStreamCreateWithFlags(&ws.fstream, hipStreamNonBlocking));
            ws.fstream_owned = true;
        }
        FT_HIP(hipEventCreateWithFlags(&ws.ev_front, hipEventDisableTiming));
        FT_HIP(hipEventCreateWithFlags(&ws.ev_free, hipEventDisableTiming));
    }
    if (ws.stream != main) {
        // the old stream's work (a layout upload) completes before it goes
        block_order_forget(ws);
        if (ws.stream) {
            FT_HIP(hipStreamSynchronize(ws.stream));
            if (ws.stream_owned) FT_HIP(hipStreamDestroy(ws.stream));
        }
        ws.stream = main;
        ws.stream_owned = false;
    }
    ws.tickets = true;  // front ends of several workspaces may wait side by side
    // the first front end waits for everything on the raster stream so far
    FT_HIP(hipEventRecord(ws.ev_free, ws.stream));
    return FT_OK;
}

void ws_pipe_detach(Workspace &ws) {
    if (ws.fstream && ws.fstream_owned) (void)hipStreamDestroy(ws.fstream);
    if (ws.ev_front) (void)hipEventDestroy(ws.ev_front);
    if (ws.ev_free) (void)hipEventDestroy(ws.ev_free);
    ws.fstream = nullptr;
    ws.ev_front = ws.ev_free = nullptr;
    ws.tickets = ws.arena != nullptr;
}

static ft_status ensure_event_capacity(Workspace &ws, size_t need) {
    if (need <= ws.cap_events && ws.d_ev_meta) return FT_OK;
    size_t cap = need < 4096 ? 4096 : need;
    ft_status s;
    if (ws.arena) ws.d_ev_f = nullptr;  // arena workspaces only fill: no stroke-mode events
    if ((s = dev_alloc(ws, ws.d_ev_xy, cap)) || (!ws.arena && (s = dev_alloc(ws, ws.d_ev_f, cap))) ||
        (s = dev_alloc(ws, ws.d_ev_meta, cap)) ||
        (s = dev_alloc(ws, ws.d_ev_kept, cap)) || (s = dev_alloc(ws, ws.d_ev_idx, cap)) || (s = dev_alloc(ws, ws.d_ev_last, cap)) ||
        (s = dev_alloc(ws, ws.d_run_first, cap)) || (s = dev_alloc(ws, ws.d_run_close, cap)) ||
        (s = dev_alloc(ws, ws.d_run_path, cap)) || (s = dev_alloc(ws, ws.d_run_start, cap)) ||
        (s = dev_alloc(ws, ws.d_run_end, cap)) || (s = dev_alloc(ws, ws.d_pts, cap)) || (s = dev_alloc(ws, ws.d_pt_run, cap)) ||
        (s = dev_alloc(ws, ws.d_edges, cap)) || (s = dev_alloc(ws, ws.d_pt_band, cap))) {
        ws.cap_events = 0;
        return s;
    }
    ws.cap_events = cap;
    return FT_OK;
}

static ft_status ensure_stroke_capacity(Workspace &ws) {
    if (ws.cap_stroke == ws.cap_events && ws.d_st_flags) return FT_OK;
    const size_t cap = ws.cap_events, items = 4 * cap + 1;
    ft_status s;
    if ((s = dev_alloc(ws, ws.d_st_flags, cap)) || (s = dev_alloc(ws, ws.d_st_off, cap)) ||
        (s = dev_alloc(ws, ws.d_st_last, cap)) || (s = dev_alloc(ws, ws.d_spts, cap)) || (s = dev_alloc(ws, ws.d_sub_start, cap)) ||
        (s = dev_alloc(ws, ws.d_sub_event, cap)) || (s = dev_alloc(ws, ws.d_sub_n, cap)) ||
        (s = dev_alloc(ws, ws.d_sub_joined, cap)) || (s = dev_alloc(ws, ws.d_it_cnt, items)) ||
        (s = dev_alloc(ws, ws.d_it_off, items))) {
        ws.cap_stroke = 0;
        return s;
    }
    ws.cap_stroke = cap;
    return FT_OK;
}

static ft_status ensure_scan_capacity(Workspace &ws) {
    size_t n = 4 * ws.cap_events + 1;  // stroke outline items bound every other scan length
    if (ws.n_ops > n) n = ws.n_ops;
    if ((ws.n_bands << ws.band_shift) > n) n = ws.n_bands << ws.band_shift;
    size_t tiles = (n + SCAN_TILE_MIN - 1) / SCAN_TILE_MIN + 1;
    // look-back words for any scan grid
    if (2 * tiles < lookback_words(SCAN_GRID_MAX)) tiles = lookback_words(SCAN_GRID_MAX) / 2;
    if (tiles <= ws.cap_tiles && ws.d_scan_state) return FT_OK;
    ft_status s;
    if ((s = dev_alloc(ws, ws.d_scan_state, 2 * tiles * SCAN_SLOTS))) return s;
    ws.cap_tiles = tiles;
    return FT_OK;
}

ft_status ws_reserve_like(Workspace &dst, const Workspace &src) {
    if (dst.arena) return FT_EINVAL;
    ft_status s;
    if (src.cap_events > dst.cap_events && (s = ensure_event_capacity(dst, src.cap_events))) return s;
    if ((s = ensure_scan_capacity(dst))) return s;
    if (src.cap_pairs > dst.cap_pairs) {
        if ((s = dev_alloc(dst, dst.d_band_list, src.cap_pairs))) return s;
        dst.cap_pairs = src.cap_pairs;
    }
    dst.ingest_pass = src.ingest_pass;
    graph_drop(dst);
    return FT_OK;
}

ft_status ws_upload_ops(Workspace &ws, const ft_path_op *ops, size_t n_ops) {
    if (n_ops != ws.n_ops) return FT_EINVAL;
    if (n_ops) FT_HIP(hipMemcpyAsync(ws.d_ops, ops, n_ops * sizeof(ft_path_op), hipMemcpyHostToDevice, ws.stream));
    return FT_OK;
}

// Arena mode: forget every carved buffer (the next layout carves afresh) and
// carve the status block first.
static ft_status arena_rewind(Workspace &ws) {
    ws.arena_used = 0;
    void **ptrs[] = {(void **)&ws.d_lay, (void **)&ws.d_ops, (void **)&ws.d_op_path, (void **)&ws.d_op_order,
                     (void **)&ws.d_paths,
                     (void **)&ws.d_band_path, (void **)&ws.d_op_cnt, (void **)&ws.d_op_off, (void **)&ws.d_sub_off,
                     (void **)&ws.d_leaf, (void **)&ws.d_op_pw, (void **)&ws.d_op_lastpw, (void **)&ws.d_ev_xy,
                     (void **)&ws.d_ev_f, (void **)&ws.d_ev_meta, (void **)&ws.d_ev_kept, (void **)&ws.d_ev_idx,
                     (void **)&ws.d_ev_last, (void **)&ws.d_st_flags, (void **)&ws.d_st_off, (void **)&ws.d_st_last,
                     (void **)&ws.d_spts, (void **)&ws.d_sub_start, (void **)&ws.d_sub_event, (void **)&ws.d_sub_n,
                     (void **)&ws.d_sub_joined, (void **)&ws.d_it_cnt, (void **)&ws.d_it_off,
                     (void **)&ws.d_run_first, (void **)&ws.d_run_close, (void **)&ws.d_run_path,
                     (void **)&ws.d_rblock, (void **)&ws.d_run_start, (void **)&ws.d_run_end, (void **)&ws.d_pts,
                     (void **)&ws.d_pt_run, (void **)&ws.d_edges, (void **)&ws.d_pt_band, (void **)&ws.d_band_cnt,
                     (void **)&ws.d_band_off, (void **)&ws.d_band_cur, (void **)&ws.d_band_list,
                     (void **)&ws.d_scan_state, (void **)&ws.d_pstate, (void **)&ws.d_top_rep,
                     (void **)&ws.d_items, (void **)&ws.d_stats, (void **)&ws.d_status};
    for (void **p : ptrs) *p = nullptr;
    ws.cap_ops = ws.cap_paths = ws.cap_events = ws.cap_stroke = ws.cap_lay = 0;
    ws.cap_band_slots = ws.cap_pairs = ws.cap_tiles = 0;
    return arena_carve(ws, ws.d_status, 1);
}

// Arena mode: the rest of the arena, after the layout's fixed buffers and the
// scan state, split between the event-sized buffers (88 B per event: no
// stroke-mode events) and the band lists (4 B per pair), 3 : 1 in bytes.
static ft_status arena_split(Workspace &ws) {
    constexpr size_t EVENT_BYTES = 88, SLACK = 32 * 256;  // alignment of the carved buffers
    const size_t left = ws.arena_bytes > ws.arena_used ? ws.arena_bytes - ws.arena_used : 0;
    // scan state for the largest event capacity this arena could hold
    const size_t ev_max = left / EVENT_BYTES;
    size_t n = 4 * ev_max + 1;
    if (ws.n_ops > n) n = ws.n_ops;
    if ((ws.n_bands << ws.band_shift) > n) n = ws.n_bands << ws.band_shift;
    size_t tiles = (n + SCAN_TILE_MIN - 1) / SCAN_TILE_MIN + 1;
    if (2 * tiles < lookback_words(SCAN_GRID_MAX)) tiles = lookback_words(SCAN_GRID_MAX) / 2;
    const size_t scan_bytes = 2 * tiles * SCAN_SLOTS * sizeof(unsigned long long);
    if (left < scan_bytes + SLACK + 4096 * (EVENT_BYTES + 4)) return FT_ENOMEM;
    const size_t room = left - scan_bytes - SLACK;
    const size_t cap_events = room / 4 * 3 / EVENT_BYTES;
    ft_status s;
    if ((s = dev_alloc(ws, ws.d_scan_state, 2 * tiles * SCAN_SLOTS))) return s;
    ws.cap_tiles = tiles;
    if ((s = ensure_event_capacity(ws, cap_events))) return s;
    const size_t rest = ws.arena_bytes - ws.arena_used - 256;
    if ((s = dev_alloc(ws, ws.d_band_list, rest / 4))) return s;
    ws.cap_pairs = rest / 4;
    return FT_OK;
}

constexpr int LAY_SECTIONS = 7;  // ops, op -> path, op order, paths, band -> path, raster blocks, tile items

// One-pass flattening (k_flatten): one item per thread, every block resident.
static inline bool flatten_in_one_pass(const Workspace &ws, size_t n_ops) {
    const size_t blocks = (n_ops * OP_LANES + BLOCK - 1) / BLOCK;
    return blocks <= ws.flat_resident;
}

// The tile scheduler's persistent grid for pixel format f: as many blocks as
// the device holds at once with `cells` LDS cells each (at most the items).
static uint32_t tile_grid_for(const Workspace &ws, int f, uint32_t cells) {
    const size_t lds = raster_lds_bytes((uint32_t)f, cells);
    int per_cu = 0;
    hipError_t e;
    if (f == FT_MATTE8)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiles<false, FMT_MATTE8>, BLOCK, lds);
    else if (f == FT_RGBA8P)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiles<true, FMT_RGBA8P>, BLOCK, lds);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tiles<true, FMT_GRAYA8P>, BLOCK, lds);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    const uint64_t cus = ws.scan_resident ? ws.scan_resident / 4 : 64;
    const uint64_t g = cus * (uint64_t)per_cu;
    return (uint32_t)std::min<uint64_t>(g, ws.it_count[f]);
}

ft_status ws_set_batch(Workspace &ws, const ft_path_op *ops, size_t n_ops, const ft_fill_desc *paths,
                       size_t n_paths, uint32_t extra_flags) {
    if (n_ops >= 0x7FFFFFFFu || n_paths >= 0x7FFFFFFFu) return FT_EINVAL;
    graph_drop(ws);
    ws.ingest_pass = false;
    // host staging kept in the workspace: the uploads below need not complete
    // before this returns (the run that follows is ordered after them); the
    // staging is only rewritten once the previous layout's uploads are done
    if (ws.uploads_pending) FT_HIP(hipStreamSynchronize(ws.stream));
    ws.uploads_pending = false;
    if (ws.arena) {
        const ft_status s = arena_rewind(ws);
        if (s) return s;
    }
    std::vector<uint32_t> &op_path = ws.h_op_path;
    std::vector<PathDev> &pd = ws.h_paths;
    std::vector<uint32_t> &band_path = ws.h_band_path;
    std::vector<uint4> &rblock = ws.h_rblock;
    op_path.assign(n_ops, NO_IDX);
    pd.assign(n_paths, PathDev());
    band_path.clear();
    rblock.clear();
    std::vector<uint4> rblock_f[3];  // per pixel format (ft_format), then concatenated
    std::vector<uint2> items_f[3];
    std::vector<uint64_t> rcost_f[3];  // estimated work of each raster block
    std::vector<uint64_t> cost_f[3];  // estimated work of each item, for the largest-first order
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
        P = path_from_desc(d, extra_flags, ws.tuning);
        // Tile mode: a path of at most tuning.tile_groups LDS groups (rows
        // whose cells fit lds_cells) is rastered by the tile scheduler, one
        // work item per group, with no bands.
        const uint32_t wp = cell_stride(d.width);
        const uint32_t gt = wp ? std::max(1u, ws.tuning.lds_cells / wp) : 1u;
        const uint64_t tgroups = d.height ? ((uint64_t)d.height + gt - 1) / gt : 0;
        if (ws.tuning.tile_groups && d.width && d.height && tgroups <= ws.tuning.tile_groups) {
            P.flags |= PATH_TILE;
            P.n_bands = 0;
            P.band_rows = d.height;
            P.group_rows = gt;
            P.band_base = (uint32_t)band_path.size();
            const uint32_t cells = std::min(gt, d.height) * wp;
            if (cells > lds) lds = cells;
            for (uint32_t g = 0; g < (uint32_t)tgroups; ++g) {
                const uint32_t rg = std::min(gt, d.height - g * gt);
                items_f[d.fmt].push_back(make_uint2((uint32_t)k, g));
                // edges in the group's rows (about 16 points per op, spread over the
                // path's rows, a few rows per edge) and the group's pixels
                cost_f[d.fmt].push_back((uint64_t)d.op_count * 64u * rg / d.height + (uint64_t)d.width * rg);
            }
            continue;
        }
        uint32_t rows, group, nb, cells;
        band_geometry(d.width, d.height, ws.tuning, &rows, &group, &nb, &cells);
        nb = P.n_bands;
        P.band_base = (uint32_t)band_path.size();
        if (nb && cells > lds) lds = cells;
        band_path.insert(band_path.end(), nb, (uint32_t)k);
        if (nb) {
            const uint32_t G = P.group_rows;
            const uint32_t groups = (P.band_rows + G - 1) / G;
            for (uint32_t b = 0; b < nb; ++b) {
                const uint32_t rows = std::min(P.band_rows, d.height - b * P.band_rows);
                for (uint32_t g = 0; g < groups && g * G < rows; ++g) {
                    const uint32_t rg = std::min(G, rows - g * G);
                    rblock_f[d.fmt].push_back(make_uint4(P.band_base + b, g, (uint32_t)k, 0u));
                    rcost_f[d.fmt].push_back((uint64_t)d.op_count * 64u * rg / d.height + (uint64_t)d.width * rg);
                }
            }
        }
    }
    for (int f = 0; f < 3; ++f) {  // largest first within each format (a single path's blocks are alike)
        ws.rb_first[f] = (uint32_t)rblock.size();
        ws.rb_count[f] = (uint32_t)rblock_f[f].size();
        if (n_paths < 2) {
            rblock.insert(rblock.end(), rblock_f[f].begin(), rblock_f[f].end());
            continue;
        }
        std::vector<uint32_t> order(rblock_f[f].size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (uint32_t)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return rcost_f[f][a] > rcost_f[f][b]; });
        for (uint32_t i : order) rblock.push_back(rblock_f[f][i]);
    }
    ws.h_items.clear();
    for (int f = 0; f < 3; ++f) {
        std::vector<uint32_t> order(items_f[f].size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (uint32_t)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return cost_f[f][a] > cost_f[f][b]; });
        ws.it_first[f] = (uint32_t)ws.h_items.size();
        ws.it_count[f] = (uint32_t)order.size();
        for (uint32_t i : order) ws.h_items.push_back(items_f[f][i]);
    }
    if (ws.h_items.size() >= 0x7FFFFFFFu) return FT_ENOTSUP;
    if (rblock.size() >= 0x7FFFFFFFu) return FT_ENOTSUP;
    // the tuning asks for more LDS than a block has
    if (raster_lds_bytes(FMT_MATTE8, lds) > ws.lds_max || raster_lds_bytes(FMT_RGBA8P, lds) > ws.lds_max)
        return FT_ENOTSUP;
    ft_status s = FT_OK;
    if (n_ops > ws.cap_ops || !ws.d_op_cnt) {  // per-op scratch
        ws.cap_ops = n_ops ? n_ops : 1;
        if ((s = dev_alloc(ws, ws.d_op_cnt, ws.cap_ops)) ||
            (s = dev_alloc(ws, ws.d_op_off, ws.cap_ops)) || (s = dev_alloc(ws, ws.d_op_pw, ws.cap_ops)) ||
            (s = dev_alloc(ws, ws.d_sub_off, ws.cap_ops * OP_LANES)) ||
            (s = dev_alloc(ws, ws.d_leaf, ws.cap_ops * OP_LANES * LEAF_CACHE)) ||
            (s = dev_alloc(ws, ws.d_op_lastpw, ws.cap_ops))) {
            ws.cap_ops = 0;
            return s;
        }
    }
    if (n_paths > ws.cap_paths || !ws.d_pstate) {  // per-path results
        const size_t cap = n_paths ? n_paths : 1;
        if (ws.arena)
            s = dev_alloc(ws, ws.d_pstate, cap);
        else if (cap > ws.cap_sp_paths)
            s = sp_grow(ws, cap);
        if (s || (s = dev_alloc(ws, ws.d_top_rep, cap * TOP_REPS))) {
            ws.cap_paths = 0;
            return s;
        }
        ws.cap_paths = cap;
    }
    size_t nb_total = band_path.size();
    // replicas per band counter: enough that the slots spread over about
    // BAND_SLOTS_MIN counters, at most 2^BAND_SHIFT_MAX
    uint32_t shift = 0;
    while (shift < BAND_SHIFT_MAX && (nb_total << shift) < BAND_SLOTS_MIN) ++shift;
    if ((nb_total << shift) >= 0x7FFFFFFFu) return FT_ENOTSUP;
    const size_t n_slots = nb_total << shift;
    if (n_slots > ws.cap_band_slots || !ws.d_band_cnt) {
        size_t cap = n_slots ? n_slots : 1;
        if ((s = dev_alloc(ws, ws.d_band_cnt, cap)) || (s = dev_alloc(ws, ws.d_band_off, cap)) ||
            (s = dev_alloc(ws, ws.d_band_cur, cap)))
            return s;
        ws.cap_band_slots = cap;
    }
    ws.band_shift = shift;
    ws.n_rblocks = rblock.size();
    // the upload block: its sections at 256-byte aligned offsets (the
    // pinned staging below uses the same offsets)
    const size_t sec_bytes[LAY_SECTIONS] = {n_ops * sizeof(ft_path_op), n_ops * 4, n_ops * 4,
                                            n_paths * sizeof(PathDev), nb_total * 4, rblock.size() * sizeof(uint4),
                                            ws.h_items.size() * sizeof(uint2)};
    size_t sec_off[LAY_SECTIONS], lay_bytes = 0;
    for (int k = 0; k < LAY_SECTIONS; ++k) {
        sec_off[k] = lay_bytes;
        lay_bytes += (sec_bytes[k] + 255) / 256 * 256;
    }
    if (lay_bytes > ws.cap_lay || !ws.d_lay) {
        const size_t cap = ws.arena ? lay_bytes : lay_bytes + lay_bytes / 4 + 256;
        if ((s = dev_alloc(ws, ws.d_lay, cap))) {
            ws.cap_lay = 0;
            return s;
        }
        ws.cap_lay = cap;
    }
    ws.d_ops = reinterpret_cast<ft_path_op *>(ws.d_lay + sec_off[0]);
    ws.d_op_path = reinterpret_cast<uint32_t *>(ws.d_lay + sec_off[1]);
    ws.d_op_order = reinterpret_cast<uint32_t *>(ws.d_lay + sec_off[2]);
    ws.d_paths = reinterpret_cast<PathDev *>(ws.d_lay + sec_off[3]);
    ws.d_band_path = reinterpret_cast<uint32_t *>(ws.d_lay + sec_off[4]);
    ws.d_rblock = reinterpret_cast<uint4 *>(ws.d_lay + sec_off[5]);
    ws.d_items = reinterpret_cast<uint2 *>(ws.d_lay + sec_off[6]);
    for (int f = 0; f < 3; ++f) ws.tile_grid[f] = ws.it_count[f] ? tile_grid_for(ws, f, lds) : 0u;
    ws.n_ops = n_ops;
    ws.n_paths = n_paths;
    ws.n_bands = nb_total;
    ws.lds_cells = lds;
    if (ws.arena) {
        ws.h_pstate_store.resize(n_paths);
        ws.h_pstate = ws.h_pstate_store.data();
    }
    if (ws.arena) {
        if ((s = arena_split(ws))) return s;
    } else {
        if (!ws.d_ev_meta) {
            if ((s = ensure_event_capacity(ws, 16 * n_ops + 1024))) return s;
        }
        if (!ws.d_band_list) {
            size_t cap = 4 * (16 * n_ops + 1024);
            if ((s = dev_alloc(ws, ws.d_band_list, cap))) return s;
            ws.cap_pairs = cap;
        }
        if ((s = ensure_scan_capacity(ws))) return s;
    }
    // Fill runs whose (op, lane) items fit one resident grid flatten in one
    // pass (k_flatten), which takes them in op order.  Larger batches flatten
    // their ops cubics first, then quads, then the rest (stable within each),
    // so that a wavefront's ops take the same branch.  Any order is correct:
    // events land at their scanned offsets.
    ws.op_ordered = n_ops >= OP_ORDER_MIN && !flatten_in_one_pass(ws, n_ops);
    if (ws.op_ordered) {
        ws.h_op_order.resize(n_ops);
        size_t at = 0;
        for (uint32_t pass = 0; pass < 3; ++pass)
            for (size_t i = 0; i < n_ops; ++i) {
                const uint32_t tag = ops[i].tag;
                const uint32_t cls = tag == FT_OP_CUBIC ? 0u : (tag == FT_OP_QUAD ? 1u : 2u);
                if (cls == pass) ws.h_op_order[at++] = (uint32_t)i;
            }
    }
    // The layout's uploads: staged into one pinned host buffer at the upload
    // block's offsets and sent with one copy (a copy from pageable memory goes
    // through the runtime's own staging, a host round trip each).  Arena
    // workspaces, which may not allocate, copy each section from its source.
    const void *sec_src[LAY_SECTIONS] = {ops, op_path.data(), ws.h_op_order.data(), pd.data(),
                                         band_path.data(), rblock.data(), ws.h_items.data()};
    size_t sec_len[LAY_SECTIONS];
    for (int k = 0; k < LAY_SECTIONS; ++k) sec_len[k] = sec_bytes[k];
    if (!ws.op_ordered) sec_len[2] = 0;
    if (ws.arena) {
        for (int k = 0; k < LAY_SECTIONS; ++k)
            if (sec_len[k])
                FT_HIP(hipMemcpyAsync(ws.d_lay + sec_off[k], sec_src[k], sec_len[k], hipMemcpyHostToDevice,
                                      ws.stream));
    } else if (lay_bytes) {
        if (lay_bytes > ws.cap_pin) {
            if (ws.h_pin) (void)hipHostFree(ws.h_pin);
            ws.h_pin = nullptr;
            ws.cap_pin = 0;
            if (hipHostMalloc(reinterpret_cast<void **>(&ws.h_pin), lay_bytes, hipHostMallocDefault) != hipSuccess) {
                ws.h_pin = nullptr;
                return FT_ENOMEM;
            }
            ws.cap_pin = lay_bytes;
        }
        for (int k = 0; k < LAY_SECTIONS; ++k)
            if (sec_len[k]) memcpy(ws.h_pin + sec_off[k], sec_src[k], sec_len[k]);
        FT_HIP(hipMemcpyAsync(ws.d_lay, ws.h_pin, lay_bytes, hipMemcpyHostToDevice, ws.stream));
    }
    ws.uploads_pending = true;
    return FT_OK;
}

static inline uint32_t grid_for(size_t n) {
    size_t g = (n + BLOCK - 1) / BLOCK;
    if (g < 1) g = 1;
    return (uint32_t)(g > GRID_MAX ? GRID_MAX : g);
}

static ft_status ws_enqueue(Workspace &ws, uint32_t mode, hipEvent_t *ev, bool with_pstate, bool copy_status,
                            bool keep_failed);

// Enqueue one run.  In graph mode a fill run is captured once into a HIP
// graph (the launch sequence has no memset nodes and no host waits) and then
// replayed; any change of buffers or layout drops the graph.
ft_status ws_launch(Workspace &ws, uint32_t mode) {
    ws.edges_counted = false;
    ws.pstate_copied = false;
    // block-index look-back: take turns with the other workspaces' runs
    std::unique_lock<std::mutex> turn;
    if (!ws.tickets && !ws.fstream) {
        turn = std::unique_lock<std::mutex>(g_block_order.mu);
        const int d = ws.device & 63;
        if (g_block_order.last[d] && g_block_order.last[d] != &ws)
            FT_HIP(hipStreamSynchronize(g_block_order.stream[d]));
        g_block_order.last[d] = &ws;
        g_block_order.stream[d] = ws.stream;
    }
    hipEvent_t *ev = nullptr;
    if (ws.timing && mode == MODE_FILL && ws.n_paths && ws.n_pending < Workspace::TIMING_RING) {
        if (!ws.ev) {
            ws.ev = new hipEvent_t[3 * Workspace::TIMING_RING]();
            for (int i = 0; i < 3 * Workspace::TIMING_RING; ++i) FT_HIP(hipEventCreate(&ws.ev[i]));
        }
        ev = ws.ev + 3 * ws.n_pending++;
    }
    if (!(ws.graph && mode == MODE_FILL) || ws.fstream) {
        // An eager run leaves its status (and path states) on the device:
        // ws_finish copies them once for all the launches before it, and a
        // failed launch stays failed until then (keep_failed).
        const bool keep = ws.status_owed;
        const ft_status st = ws_enqueue(ws, mode, ev, false, false, keep);
        ws.status_owed = true;
        return st;
    }
    if (!ws.gexec) {
        hipGraph_t g = nullptr;
        FT_HIP(hipStreamBeginCapture(ws.stream, hipStreamCaptureModeThreadLocal));
        const ft_status st = ws_enqueue(ws, mode, nullptr, false, true, false);
        const hipError_t e = hipStreamEndCapture(ws.stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess || !g) return FT_EDEVICE;
        const hipError_t ie = hipGraphInstantiate(&ws.gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            ws.gexec = nullptr;
            return FT_EDEVICE;
        }
    }
    if (ev) FT_HIP(hipEventRecord(ev[0], ws.stream));
    FT_HIP(hipGraphLaunch(ws.gexec, ws.stream));
    if (ev) {  // the raster kernel is not timed inside a graph: ev[1] = ev[2]
        FT_HIP(hipEventRecord(ev[1], ws.stream));
        FT_HIP(hipEventRecord(ev[2], ws.stream));
    }
    return FT_OK;
}

static ft_status ws_enqueue(Workspace &ws, uint32_t mode, hipEvent_t *ev, bool with_pstate, bool copy_status,
                            bool keep_failed) {
    if (mode == MODE_STROKE) {
        const ft_status e = ensure_stroke_capacity(ws);
        if (e) return e;
    }
    // `s` is the front end's stream until hand_over() moves to the raster
    // stream (one stream unless the workspace is pipelined)
    const hipStream_t rs = ws.stream;
    hipStream_t s = rs;
    if (ws.fstream) {
        if (ws.uploads_pending) FT_HIP(hipEventRecord(ws.ev_free, rs));  // a new layout's uploads
        FT_HIP(hipStreamWaitEvent(ws.fstream, ws.ev_free, 0));
        s = ws.fstream;
    }
    auto hand_over = [&]() -> hipError_t {
        if (s == rs) return hipSuccess;
        s = rs;
        const hipError_t e = hipEventRecord(ws.ev_front, ws.fstream);
        return e != hipSuccess ? e : hipStreamWaitEvent(rs, ws.ev_front, 0);
    };
    const uint32_t n_ops = (uint32_t)ws.n_ops;
    const size_t tiles = ws.cap_tiles;
    unsigned long long *lb = ws.d_scan_state;  // look-back words: 2 * tiles per scan slot
    DevStatus *st = ws.d_status;
    const uint32_t n_lb = (uint32_t)(2 * tiles * SCAN_SLOTS);
    // partial top keys per path: batches of few paths spread their compaction
    // blocks' atomics; with many paths they are spread already
    unsigned long long *const top_rep = ws.n_paths < TOP_FEW_PATHS ? ws.d_top_rep : nullptr;
    // Timing: the first and the raster kernels carry the events in their
    // dispatch (ev[0] at k_clear's start, ev[1] / ev[2] around the raster), so
    // no marker packet stalls the stream between kernels.
    const size_t n_slots = ws.n_bands << ws.band_shift;
    hipExtLaunchKernelGGL(k_clear, dim3(grid_for(std::max(std::max(ws.n_paths * TOP_REPS, n_slots), (size_t)n_lb))),
                          dim3(BLOCK), 0, s, ev ? ev[0] : nullptr, nullptr, 0, ws.d_pstate, ws.d_top_rep,
                          (uint32_t)ws.n_paths, ws.d_band_cnt, ws.d_band_cur, (uint32_t)n_slots, lb, n_lb, st,
                          keep_failed ? 1u : 0u);
    const uint32_t *lastpw = nullptr;
    if (mode != MODE_FILL && n_ops) {
        k_pen_width_mark<<<grid_for(n_ops), BLOCK, 0, s>>>(ws.d_ops, n_ops, ws.d_op_pw);
        scan_launch<uint32_t, true>(s, ws.scan_resident, ws.d_op_pw, ws.d_op_lastpw, nullptr, n_ops, n_ops,
                                    lb + 10 * tiles, nullptr, st, false, false, ws.tickets);
        lastpw = ws.d_op_lastpw;
    }
    if (n_ops) {
        const uint32_t *order = ws.op_ordered ? ws.d_op_order : nullptr;
        I2 *const leaf = mode == MODE_FILL ? ws.d_leaf : nullptr;
        EventSink em{};
        em.cap_events = ws.cap_events;
        em.cap_runs = (uint32_t)ws.cap_events;
        em.mode = mode;
        em.ev_xy = ws.d_ev_xy;
        em.ev_f = ws.d_ev_f;
        em.ev_meta = ws.d_ev_meta;
        em.run_first = ws.d_run_first;
        em.run_close = ws.d_run_close;
        em.run_path = ws.d_run_path;
        em.run_start = ws.d_run_start;
        em.run_end = ws.d_run_end;
        const size_t n_items = (size_t)n_ops * OP_LANES;
        if (mode == MODE_FILL && !order && flatten_in_one_pass(ws, n_ops)) {
            k_flatten<<<(uint32_t)((n_items + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(
                ws.d_ops, ws.d_op_path, ws.d_paths, n_ops, ws.d_op_off, lb, em, st, ws.tickets);
        } else {
            k_count<<<grid_for(n_items), BLOCK, 0, s>>>(ws.d_ops, ws.d_op_path, ws.d_paths, n_ops, mode, lastpw, order,
                                                        ws.d_op_cnt, ws.d_sub_off, leaf, ws.cap_ops * OP_LANES, st);
            scan_launch<unsigned long long, false>(s, ws.scan_resident, ws.d_op_cnt, ws.d_op_off, nullptr, n_ops,
                                                   n_ops, lb, &st->op_total, st, false, false, ws.tickets);
            k_emit<<<grid_for(n_items), BLOCK, 0, s>>>(ws.d_ops, ws.d_op_path, ws.d_paths, n_ops, mode, lastpw, order,
                                                       ws.d_op_off, ws.d_sub_off, leaf, ws.cap_ops * OP_LANES, em, st);
        }
    }
    if (mode == MODE_STROKE) {
        const uint32_t ge = grid_for(ws.cap_events);
        const uint32_t cap = (uint32_t)ws.cap_events;
        const size_t cap_items = 4 * ws.cap_events + 1;
        StrokeParams sp;
        sp.join = ws.join;
        sp.miter = ws.miter;
        sp.tol_sq = ws.n_paths ? ws.stroke_tol_sq : 0.0f;
        k_st_flags<<<ge, BLOCK, 0, s>>>(ws.d_ev_f, ws.d_st_flags, ws.d_st_last, st);
        scan_launch<unsigned long long, false>(s, ws.scan_resident, ws.d_st_flags, ws.d_st_off, &st->n_events, 0, cap,
                                               lb + 12 * tiles, &st->st_total, st, true, false, ws.tickets);
        scan_launch<uint32_t, true>(s, ws.scan_resident, ws.d_st_last, ws.d_st_last, &st->n_events, 0, cap,
                                    lb + 14 * tiles, nullptr, st, true, true, ws.tickets);
        k_st_compact<<<ge, BLOCK, 0, s>>>(ws.d_ev_f, ws.d_st_off, ws.d_spts, ws.d_sub_start, ws.d_sub_event, cap, st);
        k_st_subs<<<ge, BLOCK, 0, s>>>(ws.d_ev_f, ws.d_st_last, ws.d_sub_start, ws.d_sub_event, ws.d_sub_n,
                                       ws.d_sub_joined, cap, st);
        k_st_count<<<grid_for(cap_items), BLOCK, 0, s>>>(ws.d_spts, ws.d_sub_start, ws.d_sub_n, ws.d_sub_joined, sp,
                                                         ws.d_it_cnt, st);
        scan_launch<unsigned long long, false>(s, ws.scan_resident, ws.d_it_cnt, ws.d_it_off, &st->st_items, 0,
                                               (uint32_t)cap_items, lb + 16 * tiles, &st->stroke_total, st, true,
                                               false, ws.tickets);
        EventSink em{};
        em.cap_events = ws.cap_events;
        em.cap_runs = cap;
        em.mode = MODE_FILL;
        em.ev_xy = ws.d_ev_xy;
        em.ev_f = ws.d_ev_f;
        em.ev_meta = ws.d_ev_meta;
        em.run_first = ws.d_run_first;
        em.run_close = ws.d_run_close;
        em.run_path = ws.d_run_path;
        em.run_start = ws.d_run_start;
        em.run_end = ws.d_run_end;
        k_st_emit<<<grid_for(cap_items), BLOCK, 0, s>>>(ws.d_spts, ws.d_sub_start, ws.d_sub_n, ws.d_sub_joined, sp,
                                                        ws.d_it_off, ws.d_paths, em, st);
    }
    if (mode == MODE_FILL || mode == MODE_STROKE) {
        const uint32_t ge = grid_for(ws.cap_events);
        Limits L0{};
        L0.n_ops = n_ops;
        L0.n_paths = (uint32_t)ws.n_paths;
        L0.n_bands = (uint32_t)ws.n_bands;
        L0.cap_points = (uint32_t)ws.cap_events;
        // The event items of the two ingest scans are computed inside the scans
        // while every scan block holds one tile, else by a pass of their own.
        const bool in_scan =
            scan_grid_for<uint32_t, KeptSrc::ITEMS>(ws.cap_events, ws.scan_resident).tiles_per_block <= 1;
        if (!ws.ingest_pass) {
            // kept flags (each run's first point and closing walk found
            // locally) -> point indices
            const KeptLocalSrc kl{L0, ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_close, ws.d_ev_kept, 0u,
                                  nullptr};
            if (in_scan) {
                scan_launch_src<uint32_t, false>(s, ws.scan_resident, kl, ws.d_ev_idx, &st->n_events, 0,
                                                 (uint32_t)ws.cap_events, lb + 2 * tiles, &st->n_kept, st, true,
                                                 false, ws.tickets);
            } else {
                k_kept_local<<<ge, BLOCK, 0, s>>>(L0, ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_close,
                                                  ws.d_ev_kept, st);
                scan_launch<uint32_t, false>(s, ws.scan_resident, ws.d_ev_kept, ws.d_ev_idx, &st->n_events, 0,
                                             (uint32_t)ws.cap_events, lb + 2 * tiles, &st->n_kept, st, true, false, ws.tickets);
            }
        } else {
            // pushed-point marks -> their running maximum (the last pushed
            // point so far)
            if (in_scan) {
                const IngestSrc ing{L0, ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_path, 0u};
                scan_launch_src<uint32_t, true>(s, ws.scan_resident, ing, ws.d_ev_last, &st->n_events, 0,
                                                (uint32_t)ws.cap_events, lb + 8 * tiles, nullptr, st, true, true, ws.tickets);
            } else {
                k_ingest<<<ge, BLOCK, 0, s>>>(L0, ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_path,
                                              ws.d_ev_last, st);
                scan_launch<uint32_t, true>(s, ws.scan_resident, ws.d_ev_last, ws.d_ev_last, &st->n_events, 0,
                                            (uint32_t)ws.cap_events, lb + 8 * tiles, nullptr, st, true, true, ws.tickets);
            }
            // kept flags -> point indices
            const KeptSrc kp{L0, ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_close, ws.d_ev_last,
                             ws.d_ev_kept, 0u};
            if (in_scan) {
                scan_launch_src<uint32_t, false>(s, ws.scan_resident, kp, ws.d_ev_idx, &st->n_events, 0,
                                                 (uint32_t)ws.cap_events, lb + 2 * tiles, &st->n_kept, st, true,
                                                 false, ws.tickets);
            } else {
                k_kept<<<ge, BLOCK, 0, s>>>(L0, ws.d_ev_meta, ws.d_ev_xy, ws.d_run_first, ws.d_run_close,
                                            ws.d_ev_last, ws.d_ev_kept, st);
                scan_launch<uint32_t, false>(s, ws.scan_resident, ws.d_ev_kept, ws.d_ev_idx, &st->n_events, 0,
                                             (uint32_t)ws.cap_events, lb + 2 * tiles, &st->n_kept, st, true, false, ws.tickets);
            }
        }
        k_cap<<<grid_for(ws.n_paths * 64), BLOCK, 0, s>>>(L0, ws.d_paths, ws.d_op_off, ws.d_ev_meta, ws.d_ev_xy,
                                                     ws.d_ev_kept, ws.d_ev_idx, ws.d_run_first, ws.d_pstate, st);
        k_compact<<<ge, BLOCK, 0, s>>>(L0, ws.d_ev_meta, ws.d_ev_xy, ws.d_ev_kept, ws.d_ev_idx, ws.d_run_first,
                                       ws.d_run_path, ws.d_run_start, ws.d_run_end, ws.d_pts, ws.d_pt_run, ws.d_pstate,
                                       top_rep, st);
        const bool wide = ws.cap_events >= BIN_WIDE_MIN;
        (wide ? k_edges<4> : k_edges<1>)<<<ge, BLOCK, 0, s>>>(
            L0, ws.band_shift, ws.d_pt_run, ws.d_run_path, ws.d_run_start, ws.d_run_end, ws.d_paths, ws.d_pts,
            ws.d_edges, ws.d_pt_band, ws.d_band_cnt, ws.d_pstate, top_rep, st);
        if (ws.n_bands) {
            // pairs per (band, replica) -> list segment offsets and the pair total
            const uint32_t n_slots = (uint32_t)(ws.n_bands << ws.band_shift);
            scan_launch<uint32_t, false>(s, ws.scan_resident, ws.d_band_cnt, ws.d_band_off, nullptr, n_slots, n_slots,
                                         lb + 4 * tiles, &st->n_pairs, st, true, false, ws.tickets);
            (wide ? k_bin<4> : k_bin<1>)<<<ge, BLOCK, 0, s>>>(
                L0, ws.band_shift, ws.d_pt_band, ws.d_band_off, ws.d_band_cur, ws.d_band_list, ws.cap_pairs, ws.d_pts,
                ws.d_pt_run, ws.d_run_start, ws.d_run_end, ws.d_pstate, st);
        }
        // The raster kernels: one per pixel format with band blocks, then the
        // tile scheduler per format with tile-mode items; the timing events
        // bracket the first and the last of them.
        FT_HIP(hand_over());
        struct RasterLaunch {
            int f;
            bool tile;
        };
        RasterLaunch rl[6];
        int n_rl = 0;
        for (int f = 0; f < 3; ++f)
            if (ws.n_bands && ws.rb_count[f]) rl[n_rl++] = {f, false};
        for (int f = 0; f < 3; ++f)
            if (ws.it_count[f] && ws.tile_grid[f]) rl[n_rl++] = {f, true};
        const bool pf = ws.tuning.dst_prefetch != 0;
        unsigned long long *const stats = ws.stats ? ws.d_stats : nullptr;
        uint32_t tile_slot = (uint32_t)ws.n_rblocks;  // per-block stats: band blocks, then tile blocks
        for (int k = 0; k < n_rl; ++k) {
            const int f = rl[k].f;
            hipEvent_t e0 = (ev && k == 0) ? ev[1] : nullptr, e1 = (ev && k == n_rl - 1) ? ev[2] : nullptr;
            const uint32_t lds = (uint32_t)raster_lds_bytes((uint32_t)f, ws.lds_cells);
            if (rl[k].tile) {
                const dim3 grid(ws.tile_grid[f]);
                const uint32_t nf = ws.it_first[f], ni = ws.it_count[f];
                if (f == FT_MATTE8)
                    hipExtLaunchKernelGGL(k_tiles<false, FMT_MATTE8>, grid, dim3(BLOCK), lds, s, e0, e1, 0, L0,
                                          ws.d_paths, ws.d_items, nf, ni, ws.d_edges, ws.d_pts, ws.d_pt_run,
                                          ws.d_run_start, ws.d_run_end, ws.d_pstate, st, stats, tile_slot);
                else if (f == FT_RGBA8P)
                    hipExtLaunchKernelGGL(pf ? k_tiles<true, FMT_RGBA8P> : k_tiles<false, FMT_RGBA8P>, grid,
                                          dim3(BLOCK), lds, s, e0, e1, 0, L0, ws.d_paths, ws.d_items, nf, ni,
                                          ws.d_edges, ws.d_pts, ws.d_pt_run, ws.d_run_start, ws.d_run_end,
                                          ws.d_pstate, st, stats, tile_slot);
                else
                    hipExtLaunchKernelGGL(pf ? k_tiles<true, FMT_GRAYA8P> : k_tiles<false, FMT_GRAYA8P>, grid,
                                          dim3(BLOCK), lds, s, e0, e1, 0, L0, ws.d_paths, ws.d_items, nf, ni,
                                          ws.d_edges, ws.d_pts, ws.d_pt_run, ws.d_run_start, ws.d_run_end,
                                          ws.d_pstate, st, stats, tile_slot);
                tile_slot += ws.tile_grid[f];
                continue;
            }
            const dim3 grid(ws.rb_count[f]);
            if (f == FT_MATTE8)
                hipExtLaunchKernelGGL(k_raster<false, FMT_MATTE8>, grid, dim3(BLOCK), lds, s, e0, e1, 0, L0,
                                      ws.d_paths, ws.d_rblock, ws.rb_first[f], ws.band_shift, ws.d_band_off,
                                      ws.d_band_list, ws.d_edges, ws.d_pstate, st, stats);
            else if (f == FT_RGBA8P)
                hipExtLaunchKernelGGL(pf ? k_raster<true, FMT_RGBA8P> : k_raster<false, FMT_RGBA8P>, grid,
                                      dim3(BLOCK), lds, s, e0, e1, 0, L0, ws.d_paths, ws.d_rblock, ws.rb_first[f],
                                      ws.band_shift, ws.d_band_off, ws.d_band_list, ws.d_edges, ws.d_pstate, st,
                                      stats);
            else
                hipExtLaunchKernelGGL(pf ? k_raster<true, FMT_GRAYA8P> : k_raster<false, FMT_GRAYA8P>, grid,
                                      dim3(BLOCK), lds, s, e0, e1, 0, L0, ws.d_paths, ws.d_rblock, ws.rb_first[f],
                                      ws.band_shift, ws.d_band_off, ws.d_band_list, ws.d_edges, ws.d_pstate, st,
                                      stats);
        }
        if (ev && n_rl == 0) {
            FT_HIP(hipEventRecord(ev[1], s));
            FT_HIP(hipEventRecord(ev[2], s));
        }
    }
    FT_HIP(hipGetLastError());
    FT_HIP(hand_over());
    if (copy_status) {
        const size_t back = with_pstate ? STATUS_BYTES + ws.n_paths * sizeof(PathState) : sizeof(DevStatus);
        FT_HIP(hipMemcpyAsync(ws.h_status, ws.d_status, back, hipMemcpyDeviceToHost, s));
    }
    if (ws.fstream) FT_HIP(hipEventRecord(ws.ev_free, rs));  // this workspace's buffers are free again
    return FT_OK;
}

ft_status ws_finish(Workspace &ws, bool read_pstate) {
    if (ws.status_owed) {  // the eager launches' status, with the path states when they share its block
        const bool with_ps = read_pstate && ws.d_sp && ws.n_paths;
        const size_t back = with_ps ? STATUS_BYTES + ws.n_paths * sizeof(PathState) : sizeof(DevStatus);
        FT_HIP(hipMemcpyAsync(ws.h_status, ws.d_status, back, hipMemcpyDeviceToHost, ws.stream));
        ws.pstate_copied = with_ps;
        ws.status_owed = false;
    }
    if (read_pstate && ws.n_paths && !ws.pstate_copied)
        FT_HIP(hipMemcpyAsync(ws.h_pstate, ws.d_pstate, ws.n_paths * sizeof(PathState), hipMemcpyDeviceToHost,
                              ws.stream));
    ws.pstate_copied = false;
    FT_HIP(hipStreamSynchronize(ws.stream));
    if (ws.fstream) FT_HIP(hipStreamSynchronize(ws.fstream));  // a pipelined run's status copy
    ws.uploads_pending = false;
    for (int k = 0; k < ws.n_pending; ++k) {
        float total = 0.0f, raster = 0.0f;
        hipEvent_t *ev = ws.ev + 3 * k;
        FT_HIP(hipEventElapsedTime(&total, ev[0], ev[2]));
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

ft_status ws_count_edges(Workspace &ws) {
    if (!ws.edges_counted && ws.n_paths) {
        Limits L0{};
        L0.n_ops = (uint32_t)ws.n_ops;
        L0.n_paths = (uint32_t)ws.n_paths;
        L0.n_bands = (uint32_t)ws.n_bands;
        L0.cap_points = (uint32_t)ws.cap_events;
        k_edge_count<<<grid_for(ws.cap_events), BLOCK, 0, ws.stream>>>(L0, ws.d_edges, ws.d_pt_run, ws.d_run_path,
                                                                        ws.d_pstate, ws.d_status);
        FT_HIP(hipGetLastError());
        FT_HIP(hipMemcpyAsync(ws.h_pstate, ws.d_pstate, ws.n_paths * sizeof(PathState), hipMemcpyDeviceToHost,
                              ws.stream));
        FT_HIP(hipStreamSynchronize(ws.stream));
        ws.edges_counted = true;
    }
    return FT_OK;
}

ft_status ws_raster_stats(Workspace &ws, uint64_t *pairs, uint64_t *adds) {
    size_t slots = ws.n_rblocks;
    for (int f = 0; f < 3; ++f) slots += ws.tile_grid[f];
    std::vector<unsigned long long> h(2 * (slots ? slots : 1), 0ull);
    ft_status s;
    if ((s = dev_alloc(ws, ws.d_stats, h.size()))) return s;
    FT_HIP(hipMemcpyAsync(ws.d_stats, h.data(), h.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                          ws.stream));
    const bool graph = ws.graph;
    ws.graph = false;  // an eager run with the counting kernels' outputs
    ws.stats = true;
    s = ws_run(ws, MODE_FILL, false);
    ws.stats = false;
    ws.graph = graph;
    if (s) return s;
    FT_HIP(hipMemcpy(h.data(), ws.d_stats, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    uint64_t a = 0, b = 0;
    for (size_t i = 0; i < slots; ++i) {
        a += h[2 * i];
        b += h[2 * i + 1];
    }
    if (pairs) *pairs = a;
    if (adds) *adds = b;
    return FT_OK;
}

ft_status ws_small_fill(Workspace &ws, const ft_path_op *ops, size_t n_ops, const ft_fill_desc &d, bool *handled,
                        uint64_t *offered, uint64_t *stored, uint32_t mode) {
    *handled = false;
    if (ws.arena || !n_ops || n_ops > SMALL_OPS || d.fmt > FT_GRAYA8P || !d.width || !d.height ||
        d.width > FT_MAX_WIDTH || (mode != MODE_FILL && mode != MODE_STROKE))
        return FT_OK;
    // the figure's raster geometry: row groups as in tile mode, no bands,
    // short enough that the raster has SMALL_RASTER_BLOCKS blocks or more,
    // and at most DIRECT_ROWS rows (the direct cell pass, which splits long
    // cell walks over wavefronts)
    PathDev P = path_from_desc(d, 0u, ws.tuning);
    const uint32_t wp = cell_stride(d.width);
    const uint32_t g_lds = std::max(1u, ws.tuning.lds_cells / wp);
    const uint32_t g_par = std::max(1u, (d.height + SMALL_RASTER_BLOCKS - 1) / SMALL_RASTER_BLOCKS);
    const uint32_t G = std::min(std::min(d.height, DIRECT_ROWS), std::min(g_lds, g_par));
    const size_t lds = raster_lds_bytes(d.fmt, (size_t)G * wp) + SPLIT_LIST_BYTES;
    if (lds > ws.lds_max) return FT_OK;
    P.flags |= PATH_TILE;
    P.n_bands = 0;
    P.band_base = 0;
    P.band_rows = d.height;
    P.group_rows = G;
    const uint32_t groups = (d.height + G - 1) / G;
    graph_drop(ws);
    if (ws.uploads_pending) FT_HIP(hipStreamSynchronize(ws.stream));
    ws.uploads_pending = false;
    ft_status s;
    if (!ws.d_edges || ws.cap_events < SMALL_EVENTS)
        if ((s = ensure_event_capacity(ws, SMALL_EVENTS))) return s;
    // the ops: inside the launch when they fit, else through the pinned
    // staging into the upload block
    SmallOps inl;  // the launch copies its arguments
    const ft_path_op *d_ops = nullptr;
    const size_t bytes = n_ops * sizeof(ft_path_op);
    if (n_ops <= SMALL_ARG_OPS) {
        memcpy(inl.op, ops, bytes);
    } else {
        if (bytes > ws.cap_lay || !ws.d_lay) {
            if ((s = dev_alloc(ws, ws.d_lay, bytes + 256))) {
                ws.cap_lay = 0;
                return s;
            }
            ws.cap_lay = bytes + 256;
        }
        if (bytes > ws.cap_pin) {
            if (ws.h_pin) (void)hipHostFree(ws.h_pin);
            ws.h_pin = nullptr;
            ws.cap_pin = 0;
            if (hipHostMalloc(reinterpret_cast<void **>(&ws.h_pin), bytes, hipHostMallocDefault) != hipSuccess) {
                ws.h_pin = nullptr;
                return FT_ENOMEM;
            }
            ws.cap_pin = bytes;
        }
        memcpy(ws.h_pin, ops, bytes);
        d_ops = reinterpret_cast<const ft_path_op *>(ws.d_lay);
        FT_HIP(hipMemcpyAsync(ws.d_lay, ws.h_pin, bytes, hipMemcpyHostToDevice, ws.stream));
    }
    // the layout's pointers into the upload block are stale from here on: the
    // next full-pipeline call lays the batch out again
    ws.n_ops = ws.n_paths = ws.n_bands = 0;
    StrokeParams sp;
    sp.join = ws.join;
    sp.miter = ws.miter;
    sp.tol_sq = ws.stroke_tol_sq;
    if (mode == MODE_STROKE)
        k_small_front<true><<<1, SBLOCK, 0, ws.stream>>>(d_ops, inl, (uint32_t)n_ops, P, sp, ws.d_edges,
                                                         ws.d_pstate, ws.d_status);
    else
        k_small_front<false><<<1, SBLOCK, 0, ws.stream>>>(d_ops, inl, (uint32_t)n_ops, P, sp, ws.d_edges,
                                                          ws.d_pstate, ws.d_status);
    if (d.fmt == FT_MATTE8)
        k_small_raster<FMT_MATTE8><<<groups, BLOCK, lds, ws.stream>>>(P, ws.d_edges, ws.d_pstate, ws.d_status);
    else if (d.fmt == FT_RGBA8P)
        k_small_raster<FMT_RGBA8P><<<groups, BLOCK, lds, ws.stream>>>(P, ws.d_edges, ws.d_pstate, ws.d_status);
    else
        k_small_raster<FMT_GRAYA8P><<<groups, BLOCK, lds, ws.stream>>>(P, ws.d_edges, ws.d_pstate, ws.d_status);
    FT_HIP(hipGetLastError());
    FT_HIP(hipMemcpyAsync(ws.h_status, ws.d_status, STATUS_BYTES + sizeof(PathState), hipMemcpyDeviceToHost,
                          ws.stream));
    FT_HIP(hipStreamSynchronize(ws.stream));
    const DevStatus &h = *ws.h_status;
    if (h.overflow & SMALL_NOT_HANDLED) return FT_OK;
    if (h.err) return (ft_status)h.err;
    *handled = true;
    if (offered) *offered = ws.h_pstate[0].offered;
    if (stored) *stored = ws.h_pstate[0].stored;
    return FT_OK;
}

ft_status ws_run(Workspace &ws, uint32_t mode, bool read_pstate) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        ft_status s = ws_launch(ws, mode);
        if (s) return s;
        s = ws_finish(ws, read_pstate);
        if (s != (ft_status)0x100) return s;
        const DevStatus &h = *ws.h_status;
        graph_drop(ws);
        if (h.overflow & OV_LONG_WALK) ws.ingest_pass = true;  // again, with the ingest pass
        if (!(h.overflow & ~OV_LONG_WALK)) continue;
        if (ws.arena) return FT_ENOMEM;  // the arena's capacities are fixed
        if (h.overflow & 1u) {
            if ((s = ensure_event_capacity(ws, (size_t)h.need_events + h.need_events / 4 + 1024))) return s;
            if ((s = ensure_scan_capacity(ws))) return s;
        }
        if (h.overflow & 2u) {
            size_t cap = (size_t)h.need_pairs + h.need_pairs / 4 + 1024;
            if ((s = dev_alloc(ws, ws.d_band_list, cap))) return s;
            ws.cap_pairs = cap;
        }
    }
    return FT_ENOMEM;
}

}  // namespace ft
