// ft_noalloc.hip -- the batch fill without allocation (ft_batch_run_noalloc).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc): the same launch
// sequence as ft_batch_new + ft_batch_run, with every device buffer the
// kernels need beyond their inputs and outputs (ops and layout staging,
// events, points, edges, band lists, scan state, status) carved from one
// fixed array per device, `scratch`.
//
// `scratch` is a file-scope device array: each GPU the library runs on holds
// its own copy until the process ends.  In host code its name is a host-side
// stand-in, not a device address, so its address on each device is looked up
// once (hipGetSymbolAddress, with that device current) and only that pointer
// is handed to the kernels and the copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/footile_hip.h"
#include "ft_pipeline.h"

static __device__ __align__(256) uint8_t scratch[FT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

struct DeviceSlot {
    std::mutex lock;           // one call at a time per device in this process
    void *addr = nullptr;      // this device's address of `scratch`
    ft::Workspace ws;          // arena-mode workspace over it
    bool ready = false;
};

DeviceSlot g_slot[MAX_DEVICES];

// First call on a device (that device current): the device address of
// `scratch`, and the workspace over it.
ft_status slot_setup(DeviceSlot &d, hipStream_t stream) {
    if (!d.ready) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FT_EDEVICE;
        d.addr = ptr;
        const ft_status s = ft::ws_init_arena(d.ws, d.addr, FT_NOALLOC_SCRATCH_BYTES, stream);
        if (s) return s;
        d.ready = true;
    }
    d.ws.stream = stream;
    return FT_OK;
}

}  // namespace

extern "C" ft_status ft_batch_run_noalloc(const ft_path_op *ops, size_t n_ops, const ft_fill_desc *paths,
                                          size_t n_paths, void *stream, uint64_t *offered, uint64_t *stored) {
    if ((n_ops && !ops) || (n_paths && !paths)) return FT_EINVAL;
    for (size_t i = 0; i < n_ops; ++i)
        if (ops[i].tag > FT_OP_PEN_WIDTH) return FT_EINVAL;
    if (n_ops > FT_NOALLOC_MAX_OPS || n_paths > FT_NOALLOC_MAX_PATHS) return FT_ENOMEM;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return FT_EDEVICE;
    DeviceSlot &d = g_slot[dev];
    std::lock_guard<std::mutex> hold(d.lock);
    ft_status s = slot_setup(d, reinterpret_cast<hipStream_t>(stream));
    if (s) return s;
    // lays the batch out in the scratch (FT_ENOMEM when it does not fit) and
    // runs it; a run that outgrows the scratch returns FT_ENOMEM
    if ((s = ft::ws_set_batch(d.ws, ops, n_ops, paths, n_paths))) return s;
    if ((s = ft::ws_run(d.ws, ft::MODE_FILL, offered || stored))) return s;
    for (size_t i = 0; i < n_paths; ++i) {
        if (offered) offered[i] = d.ws.h_pstate[i].offered;
        if (stored) stored[i] = d.ws.h_pstate[i].stored;
    }
    return FT_OK;
}
