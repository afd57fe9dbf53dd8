// ft_scan.h -- device building blocks shared by the gfx950 kernels: launch
// constants, the status block helpers and a single-pass prefix scan with
// decoupled look-back.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/footile_hip.h"
#include "ft_types.h"

namespace ft {

constexpr int BLOCK = 256;
constexpr uint32_t GRID_MAX = 2048;     // grid-stride kernels: 8 blocks per CU
// Items per thread of a scan tile: 16 for 32-bit items, 8 for 64-bit ones
// (the same registers).
template <typename T>
constexpr uint32_t scan_items() {
    return sizeof(T) == 8 ? 8u : 16u;
}
template <typename T>
constexpr uint32_t scan_tile() {
    return (uint32_t)BLOCK * scan_items<T>();
}
// Sizes look-back state: a scan of n items over tiles of at least this many
// items has at most n / SCAN_TILE_MIN + 1 blocks (sources with smaller tiles
// run where the state holds SCAN_GRID_MAX blocks).
constexpr uint32_t SCAN_TILE_MIN = (uint32_t)BLOCK * 8u;

// Scan blocks certainly resident at once on the current device: four per
// compute unit (a scan block needs far less than a quarter of one).
inline uint32_t scan_resident_blocks() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 0;  // unknown: scans fall back to a 64-block grid
    return (uint32_t)cus * 4u;
}
constexpr uint32_t SCAN_GRID_MAX = 1024;  // largest scan grid
// look-back state instances per run: 0 op counts, 1 kept points, 2 band
// offsets, 3 band differences, 4 pushed points' running maximum, 5 PenWidth
// indices, 6 stroker point/sub-stroke flags, 7 stroker last closes, 8 stroke
// outline items
constexpr uint32_t SCAN_SLOTS = N_SCAN_SLOTS;
static_assert(SCAN_SLOTS == 9, "scan slot assignment");
constexpr uint32_t SPIN_LIMIT = 1u << 24;

// Device atomics for the stage code.
struct DevAtom {
    __device__ static void max_i32(int32_t *p, int32_t v) { atomicMax(p, v); }
    __device__ static void max_u32(uint32_t *p, uint32_t v) { atomicMax(p, v); }
    __device__ static void min_u32(uint32_t *p, uint32_t v) { atomicMin(p, v); }
    __device__ static uint32_t add_u32(uint32_t *p, uint32_t v) { return atomicAdd(p, v); }
    __device__ static void add_cell(int32_t *p, int32_t v) { atomicAdd(p, v); }
};

__device__ inline void raise_err(DevStatus *st, uint32_t code) { atomicCAS(&st->err, 0u, code); }
// A kernel's early exit after an earlier kernel of the run failed: plain
// loads (the previous kernels have completed), both issued before either is
// tested, so the check costs one memory round trip.
__device__ inline bool run_failed(const DevStatus *st) {
    const uint32_t e = st->err, o = st->overflow;
    return (e | o) != 0u;
}

// ---------------------------------------------------------------------------
// Single-pass prefix scan with decoupled look-back.  Block b publishes its
// aggregate and then its inclusive prefix in two 64-bit words, each a value
// below 2^63 with bit 63 as its ready mark, written and read with relaxed
// agent-scope atomics: a reader that sees the mark has the value, so no
// fence (and no L2 write-back or invalidate) is needed.
// ---------------------------------------------------------------------------
template <typename T, bool IS_MAX>
__device__ inline T scan_op_t(T a, T b) {
    if constexpr (IS_MAX)
        return a > b ? a : b;
    else
        return a + b;
}

// Inclusive scan over the 64 lanes of a wavefront with DPP: row_shr 1, 2, 4
// and 8 within each row of 16 lanes, then row_bcast:15 (rows 1 and 3 add the
// last lane of the row before) and row_bcast:31 (rows 2 and 3 add lane 31).
// Lanes without a source read 0, the identity of + and of unsigned max.
// All 64 lanes must be active.
// SEG = 16 or 32 stops after the row shifts or after row_bcast:15 and gives
// independent scans of each 16- or 32-lane segment.
template <bool IS_MAX, int SEG>
__device__ inline uint32_t seg_incl_dpp(uint32_t v) {
    static_assert(SEG == 16 || SEG == 32 || SEG == 64, "segments of 16, 32 or 64 lanes");
    auto op = [](uint32_t a, uint32_t b) { return IS_MAX ? (a > b ? a : b) : a + b; };
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    if constexpr (SEG >= 32) v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    if constexpr (SEG >= 64) v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return v;
}

template <bool IS_MAX>
__device__ inline uint32_t wave_incl_dpp(uint32_t v) {
    return seg_incl_dpp<IS_MAX, 64>(v);
}

template <typename T, bool IS_MAX>
__device__ inline T wave_incl_t(T v) {
    if constexpr (sizeof(T) == 4) {
        return (T)wave_incl_dpp<IS_MAX>((uint32_t)v);
    } else {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const T t = __shfl_up(v, o, 64);
            if (lane >= o) v = scan_op_t<T, IS_MAX>(v, t);
        }
        return v;
    }
}

// Host side: blocks of a scan over n_cap items and the tiles each block owns
// (a grid the device can hold at once: no block waits for a slot).
struct ScanGrid {
    uint32_t blocks, tiles_per_block;
};

template <typename T, uint32_t ITEMS = scan_items<T>()>
inline ScanGrid scan_grid_for(size_t n_cap, uint32_t resident) {
    constexpr uint64_t TILE = (uint64_t)BLOCK * ITEMS;
    const uint64_t tiles = n_cap ? (n_cap + TILE - 1) / TILE : 1;
    uint64_t g = resident ? resident : 64;  // unknown device: a grid any GPU holds
    if (g > SCAN_GRID_MAX) g = SCAN_GRID_MAX;
    if (g > tiles) g = tiles;
    ScanGrid r;
    r.blocks = (uint32_t)g;
    r.tiles_per_block = (uint32_t)((tiles + g - 1) / g);
    return r;
}

// Look-back state of one scan (zero before the run): lb[0] hands out the
// blocks' tickets, then two words per ticket.  With tickets a block works on
// the items of the ticket it takes on arrival, not of its block index: every
// block it waits on took its ticket earlier, is running or done, and
// publishes its aggregate without waiting on anything.  So a scan completes
// however few of its blocks the device holds at once and whatever runs
// beside it (other streams' kernels that wait too).  One thread takes the
// ticket; the atomics of a whole grid on one word cost a few microseconds.
// Without tickets (block index = ticket) the grid must be resident at once
// beside nothing else that waits: a workspace whose runs own the device.
__device__ inline uint32_t lookback_ticket(unsigned long long *lb) { return (uint32_t)atomicAdd(lb, 1ull); }
// Look-back words of a scan of `blocks` blocks.
constexpr size_t lookback_words(size_t blocks) { return 2 * blocks + 2; }

// Decoupled look-back for ticket b, called by one whole wavefront: publishes
// the block's reduction btot, returns the reduction of tickets 0 .. b-1 (to
// every lane) and publishes the block's inclusive prefix.  lb[2 j + 2] is
// ticket j's aggregate and lb[2 j + 3] its inclusive prefix, each LB_READY |
// value; lane l reads ticket top - l, 64 tickets per round trip.
constexpr unsigned long long LB_READY = 1ull << 63;
template <typename T, bool IS_MAX>
__device__ inline T lookback_prefix(unsigned long long *lb, uint32_t b, T btot, DevStatus *st) {
    lb += 2;  // past the ticket counter
    const int lane = threadIdx.x & 63;
    T prefix = 0;  // identity of + and of max over unsigned values
    if (lane == 0 && ((unsigned long long)btot & LB_READY)) raise_err(st, FT_ERANGE);  // 63-bit values only
    if (b == 0) {
        if (lane == 0)
            __hip_atomic_store(&lb[1], LB_READY | (unsigned long long)btot, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return prefix;
    }
    if (lane == 0)
        __hip_atomic_store(&lb[2 * (size_t)b], LB_READY | (unsigned long long)btot, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    int64_t top = (int64_t)b - 1;
    uint32_t spins = 0;
    for (;;) {
        const int64_t j = top - lane;
        unsigned long long wi = LB_READY, wa = 0;  // past block 0: an inclusive zero
        if (j >= 0) {
            wi = __hip_atomic_load(&lb[2 * (size_t)j + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wa = __hip_atomic_load(&lb[2 * (size_t)j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const bool has_i = (wi & LB_READY) != 0, has_a = (wa & LB_READY) != 0;
        const unsigned long long im = __ballot(has_i), zm = __ballot(!has_i & !has_a);
        const int fi = im ? __ffsll((long long)im) - 1 : 64;  // nearest inclusive
        const unsigned long long need = fi >= 63 ? ~0ull : ((2ull << fi) - 1ull);
        if (zm & need) {
            if (++spins > SPIN_LIMIT) {  // never expected: report, do not hang
                if (lane == 0) {
                    atomicOr(&st->spin_fail, 1u);
                    raise_err(st, FT_EDEVICE);
                }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        T vv = 0;
        if (lane <= fi) vv = (T)((has_i ? wi : wa) & ~LB_READY);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vv = scan_op_t<T, IS_MAX>(vv, __shfl_xor(vv, o, 64));
        prefix = scan_op_t<T, IS_MAX>(prefix, vv);
        if (fi < 64) break;
        top -= 64;
    }
    if (lane == 0)
        __hip_atomic_store(&lb[2 * (size_t)b + 1],
                           LB_READY | (unsigned long long)scan_op_t<T, IS_MAX>(prefix, btot), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    return prefix;
}

// Items of a scan read from an array.  A scan's source gives item i's value;
// `first` is true on the one visit of pass 1, where a source may also store
// per-item side outputs.  bind() runs after the status check; `bad` set by a
// source raises FT_EDEVICE.
template <typename T>
struct ArraySrc {
    static constexpr uint32_t ITEMS = scan_items<T>();
    const T *in;
    uint32_t bad;
    __device__ void bind(const DevStatus *) {}
    __device__ T value(uint64_t i, bool) { return in[i]; }
};

// Prefix scan (sum or running maximum) of src's items [0, n) into out, n = *n_dev if
// given, else n_host, at most n_cap.  The block with ticket b owns the
// tiles_per_block tiles from b * tiles_per_block: it reduces them (coalesced
// loads), takes its prefix by decoupled look-back over the tickets before it
// (lane l of the first wavefront reads ticket top - l, 64 per round trip), and then
// scans the tiles again, transposing each through LDS so that both the loads
// and the stores stay coalesced.  A single-tile block keeps its items in
// registers between the passes.  `total` receives the whole reduction.
template <typename T, bool IS_MAX, class Src>
__global__ __launch_bounds__(BLOCK) void k_scan(Src src, T *out, const uint32_t *n_dev, uint32_t n_host,
                                                uint32_t n_cap, uint32_t tiles_per_block, unsigned long long *lb,
                                                T *total, DevStatus *st, bool check_failed, bool inclusive,
                                                bool tickets) {
    constexpr uint32_t ITEMS = Src::ITEMS, TILE = (uint32_t)BLOCK * ITEMS, ROW = ITEMS + 1;
    __shared__ T s_t[BLOCK * ROW];  // a tile, one padded row of ITEMS per thread
    __shared__ uint32_t s_fail, s_n, s_tk;
    __shared__ T s_wave[BLOCK / 64];
    __shared__ T s_prefix;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {  // status, length and the block's ticket in one round trip
        const uint32_t e = st->err, o = st->overflow;
        const uint32_t nn = n_dev ? *n_dev : n_host;
        s_tk = tickets ? lookback_ticket(lb) : blockIdx.x;
        s_fail = (check_failed & ((e | o) != 0u)) ? 1u : 0u;
        s_n = nn;
    }
    __syncthreads();
    if (s_fail) return;
    src.bind(st);
    const uint32_t b = s_tk;
    const uint64_t n = s_n < n_cap ? s_n : n_cap;
    const uint64_t n_tiles = (n + TILE - 1) / TILE;
    const uint64_t t0 = (uint64_t)b * tiles_per_block;
    if (t0 >= n_tiles) return;  // no items: no later block has any either
    const uint64_t t1 = t0 + tiles_per_block < n_tiles ? t0 + tiles_per_block : n_tiles;
    const bool single = t1 - t0 == 1;
    // pass 1: the chunk's reduction
    T xr[ITEMS];
    T acc = 0;
    for (uint64_t t = t0; t < t1; ++t) {
        const uint64_t base = t * TILE;
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const uint64_t i = base + (uint64_t)j * BLOCK + tid;
            xr[j] = i < n ? src.value(i, true) : (T)0;
        }
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) acc = scan_op_t<T, IS_MAX>(acc, xr[j]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc = scan_op_t<T, IS_MAX>(acc, __shfl_xor(acc, o, 64));
    if (lane == 0) s_wave[wid] = acc;
    __syncthreads();
    T btot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) btot = scan_op_t<T, IS_MAX>(btot, s_wave[w]);
    // the chunk's prefix
    if (wid == 0) {
        const T prefix = lookback_prefix<T, IS_MAX>(lb, b, btot, st);
        if (lane == 0) {
            s_prefix = prefix;
            if (t1 == n_tiles && total) *total = scan_op_t<T, IS_MAX>(prefix, btot);
        }
    }
    __syncthreads();
    T running = s_prefix;
    // pass 2: scan each tile from the running prefix
    for (uint64_t t = t0; t < t1; ++t) {
        const uint64_t base = t * TILE;
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const uint32_t k = j * BLOCK + tid;
            const uint64_t i = base + k;
            const T x = single ? xr[j] : (i < n ? src.value(i, false) : (T)0);
            s_t[(k / ITEMS) * ROW + k % ITEMS] = x;
        }
        __syncthreads();
        T v[ITEMS];
        T run = 0;
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const T before = run;
            run = scan_op_t<T, IS_MAX>(run, s_t[tid * ROW + j]);
            v[j] = inclusive ? run : before;
        }
        const T winc = wave_incl_t<T, IS_MAX>(run);
        T wexc = __shfl_up(winc, 1, 64);
        if (lane == 0) wexc = 0;
        if (lane == 63) s_wave[wid] = winc;
        __syncthreads();
        T wpre = 0, ttot = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) {
            if (w < wid) wpre = scan_op_t<T, IS_MAX>(wpre, s_wave[w]);
            ttot = scan_op_t<T, IS_MAX>(ttot, s_wave[w]);
        }
        const T pre = scan_op_t<T, IS_MAX>(running, scan_op_t<T, IS_MAX>(wpre, wexc));
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) s_t[tid * ROW + j] = scan_op_t<T, IS_MAX>(pre, v[j]);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < ITEMS; ++j) {
            const uint32_t k = j * BLOCK + tid;
            const uint64_t i = base + k;
            if (i < n) out[i] = s_t[(k / ITEMS) * ROW + k % ITEMS];
        }
        running = scan_op_t<T, IS_MAX>(running, ttot);
        __syncthreads();
    }
    if (src.bad) raise_err(st, FT_EDEVICE);
}

// Launches k_scan over a resident grid (see scan_grid_for).
template <typename T, bool IS_MAX, class Src>
inline void scan_launch_src(hipStream_t s, uint32_t resident, const Src &src, T *out, const uint32_t *n_dev,
                            uint32_t n_host, uint32_t n_cap, unsigned long long *lb, T *total, DevStatus *st,
                            bool check_failed, bool inclusive, bool tickets = true) {
    const ScanGrid g = scan_grid_for<T, Src::ITEMS>(n_cap, resident);
    k_scan<T, IS_MAX, Src><<<g.blocks, BLOCK, 0, s>>>(src, out, n_dev, n_host, n_cap, g.tiles_per_block, lb, total,
                                                      st, check_failed, inclusive, tickets);
}

template <typename T, bool IS_MAX>
inline void scan_launch(hipStream_t s, uint32_t resident, const T *in, T *out, const uint32_t *n_dev,
                        uint32_t n_host, uint32_t n_cap, unsigned long long *lb, T *total, DevStatus *st,
                        bool check_failed, bool inclusive, bool tickets = true) {
    ArraySrc<T> src{in, 0u};
    scan_launch_src<T, IS_MAX>(s, resident, src, out, n_dev, n_host, n_cap, lb, total, st, check_failed, inclusive,
                               tickets);
}

}  // namespace ft
