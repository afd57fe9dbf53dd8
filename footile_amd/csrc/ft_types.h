// ft_types.h -- plain data shared by the device pipeline and its host build.
//
// No HIP types here: the same layouts are used by the gfx950 kernels and by
// the bounds-checked host build of the stage code (tests/cpu_mirror).
#pragma once

#include <stdint.h>

#include "../../include/footile_hip.h"

namespace ft {

struct I2 {  // Fixed point (x, y), 8 bytes (same layout as HIP int2)
    int32_t x, y;
};

struct U2 {
    uint32_t x, y;
};
struct F4 {  // stroke-mode event (x, y, width, kind), 16 bytes
    float x, y, z, w;
};

// MODE_FILL: Plotter::fill.  MODE_STROKE_FLATTEN: flatten only, writing
// (x, y, width, kind) events.  MODE_STROKE: Plotter::stroke (flatten, the
// stroker stages, then the fill of the outline).
enum : uint32_t { MODE_FILL = 0, MODE_STROKE_FLATTEN = 1, MODE_STROKE = 2 };
enum : uint32_t { TAG_CLOSE = 0, TAG_MOVE, TAG_LINE, TAG_QUAD, TAG_CUBIC, TAG_PEN_WIDTH };
enum : uint32_t { FMT_MATTE8 = 0, FMT_RGBA8P = 1, FMT_GRAYA8P = 2 };

constexpr uint32_t CLOSE_BIT = 0x80000000u;
constexpr uint32_t NO_IDX = 0xFFFFFFFFu;
constexpr unsigned long long NO_KEY = ~0ull;

// Per-path descriptor as the kernels see it.
struct PathDev {
    uint32_t op_begin, op_count;
    uint32_t rule, fmt;
    uint64_t dst;
    uint32_t width, height, pitch;
    uint32_t clr;  // packed bytes c0 | c1<<8 | c2<<16 | c3<<24
    float tr[6];
    float tol_sq;
    float s_width0;
    uint32_t band_base, band_rows, n_bands;
    uint32_t group_rows;  // rows per raster block (LDS group) within a band
    uint32_t flags;       // PATH_FRESH | PATH_RAW
};

enum : uint32_t {
    PATH_FRESH = 1u,  // region starts cleared: rows above the figure are zeroed
    PATH_RAW = 2u,    // points go to Fig::add_point untransformed (Fig-level entry)
    PATH_WHOLE = 4u,  // the path owns every event of the run (a stroke outline)
    PATH_TILE = 8u    // rastered by the tile scheduler from its own points (no bands)
};

// Per-path results of a run (cleared at the start of each run).
struct PathState {
    unsigned long long topkey;  // min over points of (y, x) as a biased u64
    uint32_t topidx;            // min local index among points at topkey
    uint32_t offered;           // points offered to Fig::add_point
    uint32_t repop;             // 1 if the cap re-pop dropped local point 65534
    uint32_t pt_base;           // global index of the path's first point
    uint32_t n_edges;           // non-horizontal segments
    uint32_t stored;            // points stored in the Fig after the cap
    uint32_t fwd;               // Fig::get_dir at the top point (set before the raster)
};

constexpr uint32_t N_SCAN_SLOTS = 9;

// DevStatus::overflow bit: a run's closing walk was too long for the kept
// flags' local form; the run is launched again with the ingest pass.
constexpr uint32_t OV_LONG_WALK = 0x20u;

struct DevStatus {
    uint32_t err;        // first ft_status error raised by a kernel
    uint32_t overflow;   // bit 0: events, bit 1: band pairs, OV_LONG_WALK
    uint32_t n_events, n_closes;
    uint32_t n_kept, n_pairs;
    uint32_t need_events, need_pairs;
    unsigned long long op_total;      // packed (events << 32 | closes) of all ops
    unsigned long long st_total;      // stroker: points stored << 32 | sub-strokes, before the cap
    unsigned long long stroke_total;  // stroker outline: events << 32 | closes
    uint32_t st_subs, st_items;       // sub-strokes kept; outline items
    uint32_t spin_fail;
    uint32_t pad[3];
};

struct EdgeRec {
    int32_t y_upper, y_lower, x_bot0, inv, step;
    uint32_t flags;  // bit0: edge present, bit1: Forward direction
};

}  // namespace ft
