// ft_device.h -- device-side arithmetic shared by the gfx950 kernels.
//
// 16.16 fixed point with the release-build semantics of footile's Fixed
// (src/fixed.rs): i32 add/sub wrap, mul/div go through i64 and truncate back,
// f32 -> Fixed truncates toward zero with saturation (NaN -> 0).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FT_HD __host__ __device__
#else
#define FT_HD
#endif

namespace ft {

constexpr int32_t FX_ONE = 65536;
constexpr int32_t FX_HALF = 32768;
constexpr int32_t FX_MASK = 0xFFFF;

FT_HD inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
FT_HD inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

// Fixed * Fixed, src/fixed.rs:40-47
FT_HD inline int32_t fx_mul(int32_t a, int32_t b) {
    return (int32_t)(uint32_t)(uint64_t)(((int64_t)a * (int64_t)b) >> 16);
}
// Fixed / Fixed, src/fixed.rs:49-56 (callers never divide by zero)
FT_HD inline int32_t fx_div(int32_t a, int32_t b) {
    return (int32_t)(uint32_t)(uint64_t)(((int64_t)a * 65536) / (int64_t)b);
}
FT_HD inline int32_t fx_abs(int32_t a) { return a < 0 ? (int32_t)(0u - (uint32_t)a) : a; }
FT_HD inline int32_t fx_floor(int32_t a) { return (int32_t)((uint32_t)a & 0xFFFF0000u); }
FT_HD inline int32_t fx_ceil(int32_t a) { return fx_floor(wadd(a, FX_ONE - 1)); }
FT_HD inline int32_t fx_round(int32_t a) { return fx_floor(wadd(a, FX_HALF)); }
FT_HD inline int32_t fx_fract(int32_t a) { return a & FX_MASK; }
FT_HD inline int32_t fx_avg(int32_t a, int32_t b) { return wadd(a, b) >> 1; }
FT_HD inline int32_t fx_int(int32_t a) { return a >> 16; }  // row_of / min_pix

// `(f * 65536.0) as i32`, src/fixed.rs:88-93
FT_HD inline int32_t fx_from_f32(float f) {
    float v = f * 65536.0f;
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v < -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

// pixel_cov, src/fig.rs:677-682: round(fcov << 8) as 0..=256
FT_HD inline int32_t pixel_cov(int32_t fcov) {
    return (int32_t)(int16_t)(fx_round((int32_t)((uint32_t)fcov << 8)) >> 16);
}

// FxPt::widdershins, src/fig.rs:116-119
FT_HD inline bool widdershins(int32_t ax, int32_t ay, int32_t bx, int32_t by) {
    return fx_mul(ax, by) > fx_mul(bx, ay);
}

// pix 0.14 Ch8 multiply: both operands widened to 12 bits, product >> 16.
FT_HD inline uint32_t ch8_widen(uint32_t v) { return (v << 4) | (v >> 4); }
FT_HD inline uint32_t ch8_mul_w(uint32_t lw, uint32_t rw) { return (lw * rw) >> 16; }
FT_HD inline uint32_t sat_add_u8(uint32_t a, uint32_t b) {
    uint32_t s = a + b;
    return s > 255u ? 255u : s;
}

// Map a running signed-area sum (taken mod 2^16 as i16) to coverage 0..255,
// src/fig.rs:629-671 and src/imgbuf.rs:54-66,157-167.
FT_HD inline uint32_t coverage_non_zero(int32_t sum) {
    int32_t s = (int32_t)(int16_t)(uint16_t)(uint32_t)sum;
    return s < 0 ? 0u : (s > 255 ? 255u : (uint32_t)s);
}
FT_HD inline uint32_t coverage_even_odd(int32_t sum) {
    int32_t s = (int32_t)(int16_t)(uint16_t)(uint32_t)sum;
    int32_t v = s & 0xFF;
    int32_t odd = s & 0x100;
    int32_t c = v - odd;  // -256..255: no i16 wrap
    c = c < 0 ? -c : c;
    return c > 255 ? 255u : (uint32_t)c;
}

}  // namespace ft
