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
// (the same registers).  Look-back state is sized by the smallest tile.
template <typename T>
constexpr uint32_t scan_items() {
    return sizeof(T) == 8 ? 8u : 16u;
}
template <typename T>
constexpr uint32_t scan_tile() {
    return (uint32_t)BLOCK * scan_items<T>();
}
constexpr uint32_t SCAN_TILE_MIN = (uint32_t)BLOCK * 8u;

// Scan blocks certainly resident at once on the current device: four per
// compute unit (a scan block needs far less than a quarter of one).
inline uint32_t scan_resident_blocks() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 0;  // unknown: ticket order only
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
// Single-pass prefix scan with decoupled look-back.  Tiles are taken in
// ticket order, so a tile only waits on tiles already owned by running blocks,
// each of which publishes its aggregate before it waits on anything.
// ---------------------------------------------------------------------------
// Scan inputs.  load() reads item i (and may return a 32-bit side value);
// commit() then performs the item's side effects.  A block loads all of its
// items before it commits any, so the loads of a tile go out together.  A
// pure source has no side effects, so a block may load its likely tile before
// it holds the tile's ticket.  bind() runs once per block before any item.
template <typename T>
struct ArraySrc {
    static constexpr bool kPure = true;
    const T *in;
    __device__ void bind(const DevStatus *) {}
    __device__ T load(uint32_t i, uint32_t &, uint32_t &) const { return in[i]; }
    __device__ void commit(uint32_t, T, uint32_t) const {}
};

template <typename T>
inline ArraySrc<T> array_src(const T *in) {
    return ArraySrc<T>{in};
}

template <typename T, bool IS_MAX>
__device__ inline T scan_op_t(T a, T b) {
    if constexpr (IS_MAX)
        return a > b ? a : b;
    else
        return a + b;
}

template <typename T, bool IS_MAX>
__device__ inline T wave_incl_t(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T t = __shfl_up(v, o, 64);
        if (lane >= o) v = scan_op_t<T, IS_MAX>(v, t);
    }
    return v;
}

template <typename T, bool IS_MAX, class Src>
__global__ __launch_bounds__(BLOCK) void k_scan(Src src, T *out, const uint32_t *n_dev, uint32_t n_host,
                                                uint32_t n_cap, uint32_t *flags, unsigned long long *agg,
                                                unsigned long long *incl, uint32_t *ticket, T *total,
                                                DevStatus *st, bool check_failed, bool inclusive,
                                                uint32_t resident) {
    constexpr uint32_t SCAN_ITEMS = scan_items<T>(), SCAN_TILE = scan_tile<T>();
    __shared__ uint32_t s_tile, s_fail, s_n;
    __shared__ T s_wave[BLOCK / 64];
    __shared__ T s_prefix;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // One pass: the grid covers every tile and fits on the device at once
    // (`resident` blocks, from the host), so block b owns tile b and every
    // tile's predecessors are running.  Otherwise blocks take tiles in ticket
    // order, so a tile only waits on tiles owned by running blocks.
    const bool one_pass = ((uint64_t)gridDim.x * SCAN_TILE >= n_cap) & (gridDim.x <= resident);
    if (threadIdx.x == 0) {  // tile, status and length in one round trip
        const uint32_t t = one_pass ? blockIdx.x : atomicAdd(ticket, 1u);
        const uint32_t e = st->err, o = st->overflow;
        const uint32_t nn = n_dev ? *n_dev : n_host;
        s_tile = t;
        s_fail = (check_failed & ((e | o) != 0u)) ? 1u : 0u;
        s_n = nn;
    }
    // the loads of a pure source's likely tile (tickets mostly follow block
    // order) go out with the ticket; items past n are masked once n is known
    T x[SCAN_ITEMS];
    uint32_t aux[SCAN_ITEMS];
    uint32_t loaded = NO_IDX, bad = 0;
    if constexpr (Src::kPure) {
        loaded = blockIdx.x;
        const uint64_t g0 = (uint64_t)loaded * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
        for (uint32_t j = 0; j < SCAN_ITEMS; ++j)
            x[j] = (g0 + j < n_cap) ? src.load((uint32_t)(g0 + j), bad, aux[j]) : (T)0;
    }
    __syncthreads();
    if (s_fail) return;
    const uint32_t n = s_n < n_cap ? s_n : n_cap;
    src.bind(st);
    uint32_t tile = s_tile;
    for (;;) {
        const uint64_t base = (uint64_t)tile * SCAN_TILE;
        if (base >= n) break;
        const uint64_t mine = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
        if (tile != loaded) {
#pragma unroll
            for (uint32_t j = 0; j < SCAN_ITEMS; ++j)
                x[j] = (mine + j < n) ? src.load((uint32_t)(mine + j), bad, aux[j]) : (T)0;
            if constexpr (!Src::kPure) {
#pragma unroll
                for (uint32_t j = 0; j < SCAN_ITEMS; ++j)
                    if (mine + j < n) src.commit((uint32_t)(mine + j), x[j], aux[j]);
            }
        }
        T v[SCAN_ITEMS];
        T run = 0;
#pragma unroll
        for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
            const T xi = (mine + j < n) ? x[j] : (T)0;
            const T before = run;
            run = scan_op_t<T, IS_MAX>(run, xi);
            v[j] = inclusive ? run : before;  // within the thread
        }
        const T winc = wave_incl_t<T, IS_MAX>(run);
        T wexc = __shfl_up(winc, 1, 64);
        if (lane == 0) wexc = 0;
        if (lane == 63) s_wave[wid] = winc;
        __syncthreads();
        T wpre = 0, btot = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w) {
            if (w < wid) wpre = scan_op_t<T, IS_MAX>(wpre, s_wave[w]);
            btot = scan_op_t<T, IS_MAX>(btot, s_wave[w]);
        }
        // Look-back by the first wavefront: lane l reads the state of tile
        // (top - l), so 64 predecessors are examined per round trip.  A round
        // ends at the nearest published inclusive value; every tile before it
        // must have published at least its aggregate.
        if (wid == 0) {
            T prefix = 0;  // identity of + and of max over unsigned values
            if (tile == 0) {
                if (lane == 0) {
                    __hip_atomic_store(&incl[0], (unsigned long long)btot, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&flags[0], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                if (lane == 0) {
                    __hip_atomic_store(&agg[tile], (unsigned long long)btot, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&flags[tile], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
                int64_t top = (int64_t)tile - 1;
                uint32_t spins = 0;
                for (;;) {
                    const int64_t j = top - lane;
                    const uint32_t f =
                        j >= 0 ? __hip_atomic_load(&flags[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : 2u;
                    const unsigned long long im = __ballot(f == 2u), zm = __ballot(f == 0u);
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
                    if (lane <= fi && j >= 0)
                        vv = (T)(f == 2u
                                     ? __hip_atomic_load(&incl[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                     : __hip_atomic_load(&agg[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) vv = scan_op_t<T, IS_MAX>(vv, __shfl_xor(vv, o, 64));
                    prefix = scan_op_t<T, IS_MAX>(prefix, vv);
                    if (fi < 64) break;
                    top -= 64;
                }
                if (lane == 0) {
                    __hip_atomic_store(&incl[tile], (unsigned long long)scan_op_t<T, IS_MAX>(prefix, btot),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&flags[tile], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            if (lane == 0) {
                s_prefix = prefix;
                if (base + SCAN_TILE >= n && total) *total = scan_op_t<T, IS_MAX>(prefix, btot);
            }
        }
        __syncthreads();
        const T pre = scan_op_t<T, IS_MAX>(s_prefix, scan_op_t<T, IS_MAX>(wpre, wexc));
#pragma unroll
        for (uint32_t j = 0; j < SCAN_ITEMS; ++j)
            if (mine + j < n) out[mine + j] = scan_op_t<T, IS_MAX>(pre, v[j]);
        if (one_pass) break;
        loaded = NO_IDX;
        __syncthreads();
        if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
        __syncthreads();
        tile = s_tile;
    }
    if (bad) raise_err(st, FT_EDEVICE);
}

}  // namespace ft
