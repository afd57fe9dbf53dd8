"""ctypes binding of libfootile_hip.so (include/footile_hip.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_FILE = os.path.join(_HERE, "libfootile_hip.so")

STATUS = {0: "FT_OK", 1: "FT_EINVAL", 2: "FT_ENOMEM", 3: "FT_EDEVICE", 4: "FT_ERANGE", 5: "FT_ENOTSUP"}


class FootileError(RuntimeError):
    """A non-OK ft_status from the library."""

    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {STATUS.get(code, code)}")


class FillDescC(ctypes.Structure):
    """ft_fill_desc"""
    _fields_ = [
        ("op_begin", ctypes.c_uint64), ("op_count", ctypes.c_uint32), ("rule", ctypes.c_uint32),
        ("dst", ctypes.c_uint64), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
        ("pitch", ctypes.c_uint32), ("fmt", ctypes.c_uint32), ("clr", ctypes.c_uint8 * 4),
        ("transform", ctypes.c_float * 6), ("tolerance", ctypes.c_float),
        ("pen_width", ctypes.c_float), ("fresh", ctypes.c_uint32),
    ]


_lib = None


def library_path():
    return _LIB_FILE


def _declare(L):
    vp, sz, u32, f32, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_float, ctypes.c_int
    pvp = ctypes.POINTER(ctypes.c_void_p)
    pu64 = ctypes.POINTER(ctypes.c_uint64)
    sig = {
        "ft_plotter_new": ([u32, u32, i, pvp], i),
        "ft_plotter_new_with_pixels": ([u32, u32, i, vp, pvp], i),
        "ft_plotter_free": ([vp], None),
        "ft_plotter_size": ([vp, ctypes.POINTER(u32), ctypes.POINTER(u32)], i),
        "ft_plotter_set_tolerance": ([vp, f32], i),
        "ft_plotter_set_transform": ([vp, vp], i),
        "ft_plotter_set_join": ([vp, i, f32], i),
        "ft_plotter_fill": ([vp, i, vp, sz, vp], i),
        "ft_plotter_stroke": ([vp, vp, sz, vp], i),
        "ft_plotter_read": ([vp, vp, sz], i),
        "ft_plotter_write": ([vp, vp, sz], i),
        "ft_plotter_device_ptr": ([vp, pvp, ctypes.POINTER(sz)], i),
        "ft_plotter_pen_width": ([vp, ctypes.POINTER(f32)], i),
        "ft_plotter_last_counts": ([vp, pu64, pu64], i),
        "ft_fig_fill": ([i, vp, vp, sz, i, u32, u32, vp, vp], i),
        "ft_batch_new": ([vp, sz, vp, sz, pvp], i),
        "ft_batch_free": ([vp], None),
        "ft_batch_launch": ([vp], i),
        "ft_batch_sync": ([vp], i),
        "ft_batch_run": ([vp], i),
        "ft_batch_counts": ([vp, vp, vp, vp, vp], i),
        "ft_batch_stream": ([vp, pvp], i),
        "ft_batch_set_tuning": ([vp, vp], i),
        "ft_batch_get_tuning": ([vp, vp], i),
        "ft_batch_set_graph": ([vp, i], i),
        "ft_batch_set_pipeline": ([vp, i], i),
        "ft_batch_set_timing": ([vp, i], i),
        "ft_batch_timing": ([vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), pu64], i),
        "ft_batch_run_noalloc": ([vp, sz, vp, sz, vp, vp, vp], i),
        "ft_batch_raster_stats": ([vp, pu64, pu64], i),
        "ft_device_alloc": ([sz, pvp], i),
        "ft_device_free": ([vp], i),
        "ft_memcpy_h2d": ([vp, vp, sz], i),
        "ft_memcpy_d2h": ([vp, vp, sz], i),
        "ft_memset_d": ([vp, i, sz], i),
        "ft_version": ([], ctypes.c_char_p),
        "ft_diff_new": ([vp, pvp], i),
        "ft_diff_free": ([vp], None),
        "ft_diff_info": ([vp, pu64, pu64, pu64], i),
        "ft_diff_forward": ([vp, vp, vp, vp], i),
        "ft_diff_loss": ([vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double)], i),
        "ft_diff_backward": ([vp, vp, vp, vp, vp, vp], i),
        "ft_diff_step": ([vp, vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double)], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def lib():
    """Load libfootile_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_FILE):
            raise FootileError(-1, f"{_LIB_FILE} is missing; build it with "
                                   "`python -m footile_amd.build`")
        L = ctypes.CDLL(_LIB_FILE)
        _declare(L)
        _lib = L
    return _lib


def check(code, what):
    if code != 0:
        raise FootileError(code, what)
