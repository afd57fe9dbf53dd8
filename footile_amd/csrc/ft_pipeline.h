// ft_pipeline.h -- host-side interface of the device fill pipeline.
//
// One "run" renders a batch of independent fills (a single Plotter::fill is a
// batch of one).  All stages run on the GPU from device-resident ops:
//   count -> scan -> emit          (op interpretation + De Casteljau flatten)
//   ingest -> max-scan -> kept -> scan -> cap -> compact -> top
//                                  (Fig::add_point/close, 65,535 cap, top point)
//   edges -> scan -> scan -> bin   (Edge::new per segment, row-band binning)
//   raster                         (signed-area cells in LDS, row scan,
//                                   NonZero/EvenOdd map, Matte/Rgba store)
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/footile_hip.h"
#include "ft_stages.h"
#include "ft_types.h"

namespace ft {

struct Workspace {
    hipStream_t stream = nullptr;
    bool stream_owned = true;  // ws_free destroys `stream`
    // Pipelined runs (ws_pipe_attach): the front end (every kernel before the
    // raster) goes to `fstream` once this workspace's previous raster and
    // status copy are done (`ev_free`, recorded on `stream`), and the raster
    // follows on `stream` after the front end (`ev_front`).  Workspaces
    // sharing `stream`, each with a front-end stream of its own, then overlap
    // their front ends with each other and with an earlier raster, while the
    // rasters stay in launch order.  Null: one stream.  (Front ends may run
    // side by side because their scans and the one-pass flattening hand out
    // look-back tickets on arrival: no block waits on a block that has not
    // started.)
    hipStream_t fstream = nullptr;
    bool fstream_owned = false;
    // look-back kernels take arrival tickets (pipelined and arena
    // workspaces, whose runs may overlap other waiting kernels); otherwise
    // block index order, over grids the device holds at once
    bool tickets = false;
    hipEvent_t ev_front = nullptr, ev_free = nullptr;
    int device = 0;
    size_t lds_max = 65536;  // LDS bytes per block on the device
    uint32_t scan_resident = 0;  // scan blocks resident at once (0: unknown)
    uint32_t flat_resident = 0;  // k_flatten blocks resident at once (0: k_flatten unused)
    // batch layout
    size_t n_ops = 0, n_paths = 0, n_bands = 0;
    size_t cap_ops = 0, cap_paths = 0;
    // The layout's uploads (ops, op -> path, op order, paths, band -> path,
    // raster blocks, tile items) are sections of one device block, filled
    // from the pinned staging by one copy; the pointers below point into it.
    uint8_t *d_lay = nullptr;
    size_t cap_lay = 0;
    ft_path_op *d_ops = nullptr;
    uint32_t *d_op_path = nullptr;
    uint32_t *d_op_order = nullptr;  // flattening order of the ops (a permutation)
    std::vector<uint32_t> h_op_order;
    // host staging of the layout's uploads (op -> path, paths, band -> path,
    // raster blocks), alive until the next layout
    std::vector<uint32_t> h_op_path, h_band_path;
    std::vector<PathDev> h_paths;
    std::vector<uint4> h_rblock;
    bool uploads_pending = false;  // a layout's uploads may not have run yet (cleared by a stream sync)
    uint8_t *h_pin = nullptr;      // pinned staging of the layout's uploads
    size_t cap_pin = 0;
    bool op_ordered = false;
    // the kept flags take the ingest pass (pushed marks and their running
    // maximum) instead of their local form: set for the rest of a layout
    // once a run's closing walk was too long
    bool ingest_pass = false;
    PathDev *d_paths = nullptr;
    uint32_t *d_band_path = nullptr;
    uint32_t lds_cells = 0;  // dynamic LDS cells needed by the raster kernel
    // raster blocks: (global band, row group) per block
    size_t n_rblocks = 0;
    uint32_t rb_first[3] = {0, 0, 0}, rb_count[3] = {0, 0, 0};  // raster blocks per pixel format
    uint4 *d_rblock = nullptr;  // raster blocks: (global band, row group, path, 0)
    // tile scheduler work items (path, row group) of tile-mode paths, per
    // pixel format, largest first; and the persistent grid per format
    std::vector<uint2> h_items;
    uint2 *d_items = nullptr;
    uint32_t it_first[3] = {0, 0, 0}, it_count[3] = {0, 0, 0};
    uint32_t tile_grid[3] = {0, 0, 0};
    // raster statistics (ws_raster_stats): per raster block, (edge, row)
    // pairs walked and nonzero cell deltas added
    bool stats = false;
    unsigned long long *d_stats = nullptr;
    // per-op scratch
    unsigned long long *d_op_cnt = nullptr, *d_op_off = nullptr;
    unsigned long long *d_sub_off = nullptr;  // per (op, lane): offset within the op (+ leaf count)
    I2 *d_leaf = nullptr;                      // per (op, lane): its first LEAF_CACHE points (fill runs)
    uint32_t *d_op_pw = nullptr, *d_op_lastpw = nullptr;
    // events and everything sized by them
    size_t cap_events = 0;
    I2 *d_ev_xy = nullptr;
    F4 *d_ev_f = nullptr;  // stroke-mode events
    uint32_t *d_ev_meta = nullptr, *d_ev_kept = nullptr, *d_ev_idx = nullptr;
    uint32_t *d_ev_last = nullptr;  // pushed-point marks, then their running maximum
    int32_t *d_run_first = nullptr;
    uint32_t *d_run_close = nullptr;
    uint32_t *d_run_path = nullptr, *d_run_start = nullptr, *d_run_end = nullptr;
    I2 *d_pts = nullptr;
    uint32_t *d_pt_run = nullptr;
    EdgeRec *d_edges = nullptr;
    U2 *d_pt_band = nullptr;  // per stored point: its edge's band range
    // stroker stages (MODE_STROKE), sized by cap_events
    size_t cap_stroke = 0;
    unsigned long long *d_st_flags = nullptr, *d_st_off = nullptr;
    uint32_t *d_st_last = nullptr;
    F4 *d_spts = nullptr;
    uint32_t *d_sub_start = nullptr, *d_sub_event = nullptr, *d_sub_n = nullptr, *d_sub_joined = nullptr;
    unsigned long long *d_it_cnt = nullptr, *d_it_off = nullptr;  // 4 cap_events + 1 items
    uint32_t join = 0;  // ft_join
    float miter = 4.0f, stroke_tol_sq = 0.09f;
    // bands
    // pairs, list offsets and fill cursors per (band, replica) slot: band g,
    // replica r at (g << band_shift) + r
    uint32_t *d_band_cnt = nullptr, *d_band_off = nullptr, *d_band_cur = nullptr;
    size_t cap_band_slots = 0;
    uint32_t band_shift = 0;
    size_t cap_pairs = 0;
    uint32_t *d_band_list = nullptr;
    // scan look-back state, SCAN_SLOTS instances
    size_t cap_tiles = 0;
    unsigned long long *d_scan_state = nullptr;  // scan look-back words, cleared by k_clear
    // results
    PathState *d_pstate = nullptr;
    unsigned long long *d_top_rep = nullptr;  // per path: TOP_REPS partial top keys
    // Outside arena mode the status and the path states are one device block
    // and one pinned host block (status first, STATUS_BYTES long), so that a
    // run's final copy can bring both back at once.
    DevStatus *d_status = nullptr;
    DevStatus *h_status = nullptr;  // pinned
    PathState *h_pstate = nullptr;  // host copy of d_pstate (pinned, or h_pstate_store in arena mode)
    uint8_t *d_sp = nullptr, *h_sp = nullptr;
    size_t cap_sp_paths = 0;
    std::vector<PathState> h_pstate_store;
    bool pstate_copied = false;  // the status copy since the last launch included the path states
    bool status_owed = false;    // eager launches left their status on the device: ws_finish copies it
    bool edges_counted = false;  // PathState::n_edges holds the last run's counts
    // band / LDS geometry of the next ws_set_batch
    Tuning tuning;
    // graph mode: fill runs replay a captured graph of the launch sequence
    bool graph = false;
    hipGraphExec_t gexec = nullptr;
    // optional event timing of whole runs and of the raster kernel alone:
    // launch k records into ring slot k (up to TIMING_RING launches between
    // two finishes); finish adds up every recorded slot.
    static constexpr int TIMING_RING = 256;
    bool timing = false;
    int n_pending = 0;
    hipEvent_t *ev = nullptr;  // TIMING_RING x 3 events (start, raster start, raster end), created on first use
    double t_total_ms = 0.0, t_raster_ms = 0.0;
    uint64_t n_timed = 0;
    // Arena mode (ws_init_arena): every device buffer is carved from
    // [arena, arena + arena_bytes), nothing is allocated or freed, each
    // ws_set_batch carves afresh, and a run that outgrows its carved
    // capacities fails with FT_ENOMEM instead of growing.  The stream is the
    // caller's and the status is read back into h_status_store.
    uint8_t *arena = nullptr;
    size_t arena_bytes = 0, arena_used = 0;
    DevStatus h_status_store{};
};

// Allocate the stream and status buffers.
ft_status ws_init(Workspace &ws);
// Pipelined runs: give `ws` the front-end stream `front` (null: a new one that
// `ws` owns) and its two events, and make `main` its raster stream (a stream
// of its own is destroyed; `main` then belongs to the workspace that created
// it).  Detach returns to one stream (the caller has drained both).
ft_status ws_pipe_attach(Workspace &ws, hipStream_t main, hipStream_t front = nullptr);
void ws_pipe_detach(Workspace &ws);
// Grow dst's event, scan and pair capacities to src's, and take over its
// kept-flag form (for a twin workspace of the same layout).
ft_status ws_reserve_like(Workspace &dst, const Workspace &src);
// Arena mode: no allocation; device buffers come from `arena` (256-byte
// aligned, `bytes` long), work goes to `stream`.
ft_status ws_init_arena(Workspace &ws, void *arena, size_t bytes, hipStream_t stream);
void ws_free(Workspace &ws);
// Upload a batch layout: ops + path descriptors (host arrays).
// extra_flags (PATH_RAW, ...) are or-ed into every path's flags.
ft_status ws_set_batch(Workspace &ws, const ft_path_op *ops, size_t n_ops, const ft_fill_desc *paths,
                       size_t n_paths, uint32_t extra_flags = 0);
// Re-upload only the ops (same count and layout) -- for a plotter reusing ws.
ft_status ws_upload_ops(Workspace &ws, const ft_path_op *ops, size_t n_ops);
// Enqueue a run (MODE_FILL: the fills; MODE_STROKE: the strokes, with the
// workspace's join, miter and stroke tolerance).  An eager run's status and
// path states stay on the device until ws_finish.
ft_status ws_launch(Workspace &ws, uint32_t mode);
// Wait and read the status block; returns the run's status.
ft_status ws_finish(Workspace &ws, bool read_pstate);
// launch + finish, growing capacities and relaunching on overflow.
ft_status ws_run(Workspace &ws, uint32_t mode, bool read_pstate);
// One fill (mode MODE_FILL) or stroke (MODE_STROKE, with the workspace's
// join, miter and stroke tolerance) of host ops as a small figure when it is
// small enough (ops, events, sub-figures and, for a stroke, the outline in
// one block's LDS): sets *handled and the offered / stored point counts;
// otherwise *handled is false and nothing was drawn.  The workspace's batch
// layout is invalid afterwards.
ft_status ws_small_fill(Workspace &ws, const ft_path_op *ops, size_t n_ops, const ft_fill_desc &d, bool *handled,
                        uint64_t *offered, uint64_t *stored, uint32_t mode = MODE_FILL);
// Run the batch once (eagerly) counting the raster's (edge, row) pairs and
// its nonzero cell deltas (the LDS atomics of the cell pass).
ft_status ws_raster_stats(Workspace &ws, uint64_t *pairs, uint64_t *adds);
// Count the last fill run's non-horizontal edges per path into
// PathState::n_edges (and h_pstate).
ft_status ws_count_edges(Workspace &ws);

}  // namespace ft
