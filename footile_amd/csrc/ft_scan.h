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
constexpr uint32_t SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = BLOCK * SCAN_ITEMS;
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
__device__ inline bool run_failed(const DevStatus *st) {
    return __hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
           __hip_atomic_load(&st->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
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
                                                DevStatus *st, bool check_failed, bool inclusive) {
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
            const unsigned long long before = run;
            run = scan_op<IS_MAX>(run, x);
            v[j] = inclusive ? run : before;  // within the thread
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
        // Look-back by the first wavefront: lane l reads the state of tile
        // (top - l), so 64 predecessors are examined per round trip.  A round
        // ends at the nearest published inclusive value; every tile before it
        // must have published at least its aggregate.
        if (wid == 0) {
            unsigned long long prefix = 0;  // identity of + and of max over unsigned values
            if (tile == 0) {
                if (lane == 0) {
                    __hip_atomic_store(&incl[0], btot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&flags[0], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                if (lane == 0) {
                    __hip_atomic_store(&agg[tile], btot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
                    unsigned long long v = 0;
                    if (lane <= fi && j >= 0)
                        v = f == 2u ? __hip_atomic_load(&incl[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : __hip_atomic_load(&agg[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) v = scan_op<IS_MAX>(v, __shfl_xor(v, o, 64));
                    prefix = scan_op<IS_MAX>(prefix, v);
                    if (fi < 64) break;
                    top -= 64;
                }
                if (lane == 0) {
                    __hip_atomic_store(&incl[tile], scan_op<IS_MAX>(prefix, btot), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&flags[tile], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            if (lane == 0) {
                s_prefix = prefix;
                if (base + SCAN_TILE >= n && total) *total = (T)scan_op<IS_MAX>(prefix, btot);
            }
        }
        __syncthreads();
        const unsigned long long pre = scan_op<IS_MAX>(s_prefix, scan_op<IS_MAX>(wpre, wexc));
#pragma unroll
        for (uint32_t j = 0; j < SCAN_ITEMS; ++j)
            if (mine + j < n) out[mine + j] = (T)scan_op<IS_MAX>(pre, v[j]);
        __syncthreads();
    }
}

}  // namespace ft
