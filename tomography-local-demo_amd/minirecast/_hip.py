"""ctypes binding of libminirecast_hip.so (the C ABI in include/minirecast_hip.h).

The library is built in-tree next to this file.  Loading it never touches the GPU;
creating a context does.  There is no CPU fallback: when the library is missing,
every call that needs it raises RuntimeError.
"""
import ctypes
import os
import threading

import numpy as np

LIB_NAME = "libminirecast_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

MR_OK, MR_EVALUE, MR_ERUNTIME, MR_ETYPE = 0, -1, -2, -3
WINDOWS = {None: 0, "none": 0, "ram-lak": 0, "shepp-logan": 1, "hann": 2, "cosine": 3}
STORAGE = {"f32": 0, "float32": 0, "f16": 1, "float16": 1, "bf16": 2, "bfloat16": 2}
KIND_CONE, KIND_PARALLEL = 1, 2

# every symbol the header declares: (name, restype, argtypes)
_c = ctypes
_P = _c.c_void_p
SIGNATURES = [
    ("mr_last_error", _c.c_char_p, []),
    ("mr_abi_version", _c.c_int, []),
    ("mr_device_count", _c.c_int, [_c.POINTER(_c.c_int)]),
    ("mr_ramp_response", _c.c_int, [_c.c_int, _c.c_double, _c.c_int, _P, _c.POINTER(_c.c_int)]),
    ("mr_ctx_create", _c.c_int, [_c.c_int, _c.POINTER(_P)]),
    ("mr_ctx_destroy", _c.c_int, [_P]),
    ("mr_set_geometry_cone", _c.c_int, [_P, _c.c_int, _c.c_int, _c.c_int, _P, _c.c_int, _c.c_int]),
    ("mr_set_geometry_parallel", _c.c_int,
     [_P, _c.c_int, _c.c_int, _c.c_int, _P, _c.c_double, _c.c_double, _c.c_int, _c.c_int]),
    ("mr_push_views", _c.c_int, [_P, _P, _c.c_int, _P, _c.c_int, _c.POINTER(_c.c_int)]),
    ("mr_store_views", _c.c_int, [_P, _P, _c.c_int, _P, _c.c_int, _c.POINTER(_c.c_int)]),
    ("mr_set_preprocess", _c.c_int, [_P, _c.c_double, _c.c_double]),
    ("mr_set_raw_correction", _c.c_int, [_P, _P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_int,
                                         _c.c_int, _c.c_int, _c.c_int]),
    ("mr_push_raw", _c.c_int, [_P, _P, _c.c_int, _P, _c.c_int, _c.POINTER(_c.c_int)]),
    ("mr_last_tile_kernel_ms", _c.c_int, [_P, _c.POINTER(_c.c_double)]),
    ("mr_last_tile_kernel_launches", _c.c_int, [_P, _c.POINTER(_c.c_int)]),
    ("mr_last_bp_kernel", _c.c_int, [_P, _c.POINTER(_c.c_int)]),
    ("mr_nonzero_percentiles", _c.c_int, [_P, _P, _c.c_int64, _c.c_int, _P, _c.c_int, _P,
                                          _c.POINTER(_c.c_int64)]),
    ("mr_clip_histogram", _c.c_int, [_P, _P, _c.c_int64, _c.c_int, _c.c_double, _c.c_double,
                                     _c.c_int, _P, _P]),
    ("mr_clip_histogram_static", _c.c_int, [_P, _P, _c.c_int64, _c.c_double, _c.c_double,
                                            _c.c_int, _P, _P]),
    ("mr_set_bp_config", _c.c_int, [_P, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    ("mr_set_bp_views_per_pass", _c.c_int, [_P, _c.c_int]),
    ("mr_get_bp_views_per_pass", _c.c_int, [_P, _c.POINTER(_c.c_int)]),
    ("mr_get_bp_config", _c.c_int, [_P, _P]),
    ("mr_set_ring_copy_budget", _c.c_int, [_P, _c.c_int64]),
    ("mr_device_bytes", _c.c_int, [_P, _P]),
    ("mr_autotune", _c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_int64, _c.c_int, _c.c_int, _P,
                               _c.POINTER(_c.c_float)]),
    ("mr_preprocess", _c.c_int,
     [_P, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_double, _c.c_double, _P]),
    ("mr_num_views", _c.c_int, [_P, _c.POINTER(_c.c_int)]),
    ("mr_bp_points", _c.c_int, [_P, _P, _c.c_int64, _c.c_double, _c.c_int, _P]),
    ("mr_bp_tasks", _c.c_int, [_P, _P, _c.c_int, _c.c_int, _P, _c.c_int64, _c.c_int]),
    ("mr_cor_scores", _c.c_int, [_P, _P, _c.c_int, _c.c_int, _P]),
    ("mr_bp_tasks_adjoint", _c.c_int,
     [_P, _P, _c.c_int, _c.c_int, _P, _c.c_int64, _c.c_int, _P, _c.c_int]),
    ("mr_bp_tasks_dshift", _c.c_int, [_P, _P, _c.c_int, _c.c_int, _P, _c.c_int64, _c.c_int]),
    ("mr_filter_adjoint", _c.c_int, [_P, _P, _c.c_int, _P, _c.c_int, _P, _c.c_int]),
    ("mr_reset_views", _c.c_int, [_P]),
    ("mr_plan_create", _c.c_int, [_P, _P, _c.c_int, _c.c_int64, _c.POINTER(_P)]),
    ("mr_plan_destroy", _c.c_int, [_P]),
    ("mr_plan_run", _c.c_int, [_P, _c.c_int, _P]),
    ("mr_plan_stream", _c.c_int, [_P, _P, _c.c_int, _P, _P]),
    ("mr_plan_update", _c.c_int, [_P, _P]),
    ("mr_plan_cor_scores", _c.c_int, [_P, _c.c_int, _P]),
    ("mr_plan_output", _c.c_int, [_P, _c.POINTER(_P), _c.POINTER(_c.c_int64)]),
    ("mr_plan_last_ms", _c.c_int, [_P, _c.POINTER(_c.c_double)]),
    ("mr_filter_rows", _c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _P]),
    ("mr_read_views", _c.c_int, [_P, _c.c_int, _c.c_int, _P]),
    ("mr_phantom_cone", _c.c_int, [_P, _P, _c.c_int, _P, _c.c_int, _c.c_int, _c.c_int, _P]),
    ("mr_phantom_parallel", _c.c_int,
     [_P, _P, _c.c_int, _P, _c.c_int, _c.c_int, _c.c_int, _c.c_double, _c.c_double,
      _c.c_double, _P]),
    ("mr_device_alloc", _c.c_int, [_P, _c.c_int64, _c.POINTER(_P)]),
    ("mr_device_free", _c.c_int, [_P, _P]),
    ("mr_memcpy", _c.c_int, [_P, _P, _P, _c.c_int64, _c.c_int]),
    ("mr_synchronize", _c.c_int, [_P]),
    ("mr_last_kernel_ms", _c.c_int, [_P, _c.POINTER(_c.c_double)]),
    ("mr_l2_gather_bandwidth", _c.c_int, [_P, _c.c_int, _c.c_int, _c.c_int64, _c.POINTER(_c.c_double),
                                          _c.POINTER(_c.c_double)]),
]

# mr_task (include/minirecast_hip.h)
TASK_DTYPE = np.dtype([
    ("o", "<f8", 3), ("a", "<f8", 3), ("b", "<f8", 3), ("c", "<f8", 3),
    ("mo", "<f8", 3), ("me1", "<f8", 3), ("me2", "<f8", 3), ("box", "<f8", 3),
    ("mhalf", "<f8"), ("mstep", "<f8"), ("cor_shift", "<f8"),
    ("out_offset", "<i8"), ("stride_q", "<i8"), ("stride_p", "<i8"), ("stride_s", "<i8"),
    ("nq", "<i4"), ("np", "<i4"), ("ns", "<i4"), ("mask_box", "<i4"),
])
assert TASK_DTYPE.itemsize == 264

_lib = None
_lib_lock = threading.Lock()


class HipUnavailable(RuntimeError):
    """The HIP library or a GPU is not available."""


def _share_torch_hip_runtime():
    """Make this library and PyTorch use one HIP runtime.

    The library needs libamdhip64.so.7.  PyTorch ships its own copy and its extension
    modules load it by another file name; if the system copy were loaded first, the
    process would hold two HIP/HSA runtimes and PyTorch's would not see the GPU.
    Loading PyTorch's copy first (without importing torch) lets the library bind to it
    through the shared soname.  MINIRECAST_SYSTEM_HIP=1 keeps the system runtime."""
    if os.environ.get("MINIRECAST_SYSTEM_HIP") == "1":
        return None
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
    except Exception:
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    for root in spec.submodule_search_locations:
        cand = os.path.join(root, "lib", "libamdhip64.so")
        if os.path.exists(cand):
            ctypes.CDLL(cand, mode=ctypes.RTLD_GLOBAL)
            return cand
    return None


FORCE_FAILURE_ENV = "MINIRECAST_FORCE_HIP_FAILURE"


def load(path=None):
    """Load (once) and return the ctypes library; RuntimeError if it is missing.

    MINIRECAST_FORCE_HIP_FAILURE=1 makes every load fail as if the library or the GPU
    were missing (the failure path of the reference's warmup, recon.py:477-487, forced
    from the host).  There is no CPU path to fall back to: callers get HipUnavailable."""
    global _lib
    if os.environ.get(FORCE_FAILURE_ENV) == "1":
        raise HipUnavailable(f"HIP back-end disabled by {FORCE_FAILURE_ENV}=1; this build has "
                             f"no CPU back-projection to fall back to")
    with _lib_lock:
        if _lib is not None:
            return _lib
        p = path or os.environ.get("MINIRECAST_HIP_LIB", LIB_PATH)
        if not os.path.exists(p):
            raise HipUnavailable(f"{LIB_NAME} not built ({p}); run tomography-local-demo_amd/build.py")
        _share_torch_hip_runtime()
        lib = ctypes.CDLL(p)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def lib_path():
    return os.environ.get("MINIRECAST_HIP_LIB", LIB_PATH)


def check(rc):
    """Map a C return code to the reference's exception types."""
    if rc == MR_OK:
        return
    msg = load().mr_last_error().decode(errors="replace")
    if rc == MR_EVALUE:
        raise ValueError(msg)
    if rc == MR_ETYPE:
        raise TypeError(msg)
    raise RuntimeError(msg)


def device_count():
    n = ctypes.c_int(0)
    check(load().mr_device_count(ctypes.byref(n)))
    return n.value


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def window_code(window):
    if window not in WINDOWS:
        raise ValueError(f"unknown filter window {window!r}")
    return WINDOWS[window]


def ramp_response(n, tau, window="shepp-logan"):
    """(H float64 [N], N) from the library's float64 host routine."""
    code = window_code(window)
    N = ctypes.c_int(0)
    check(load().mr_ramp_response(int(n), float(tau), code, None, ctypes.byref(N)))
    H = np.empty(N.value, dtype=np.float64)
    check(load().mr_ramp_response(int(n), float(tau), code, ptr(H), ctypes.byref(N)))
    return H, N.value


class Context:
    """Owns one mr_ctx (device ring + streams) for one reconstructor."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.mr_ctx_create(int(device), ctypes.byref(h)))
        self.h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.mr_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- geometry
    def set_cone(self, vectors, rows, cols, window, storage="f32"):
        V = np.ascontiguousarray(vectors, dtype=np.float64).reshape(-1, 12)
        check(self.lib.mr_set_geometry_cone(self.h, len(V), int(rows), int(cols), ptr(V),
                                            window_code(window), STORAGE[storage]))

    def set_parallel(self, angles, rows, cols, pixel_u, pixel_v, window, storage="f32"):
        a = np.ascontiguousarray(angles, dtype=np.float64).ravel()
        check(self.lib.mr_set_geometry_parallel(self.h, len(a), int(rows), int(cols), ptr(a),
                                                float(pixel_u), float(pixel_v),
                                                window_code(window), STORAGE[storage]))

    # ---------------------------------------------------------------- ingest
    def push(self, indices, block):
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        blk = np.ascontiguousarray(block, dtype=np.float32)
        n = ctypes.c_int(0)
        check(self.lib.mr_push_views(self.h, ptr(idx), len(idx), ptr(blk), 0, ctypes.byref(n)))
        return n.value

    def push_device(self, indices, dptr):
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        n = ctypes.c_int(0)
        check(self.lib.mr_push_views(self.h, ptr(idx), len(idx), ctypes.c_void_p(dptr), 1,
                                     ctypes.byref(n)))
        return n.value

    def set_raw_correction(self, dark, scale, op, frame_shape, sample=1, flip=False, row0=0,
                           col0=0):
        d = np.ascontiguousarray(dark, dtype=np.float32)
        g = np.ascontiguousarray(scale, dtype=np.float32)
        check(self.lib.mr_set_raw_correction(self.h, ptr(d), ptr(g), int(op), int(frame_shape[0]),
                                             int(frame_shape[1]), int(sample), int(bool(flip)),
                                             int(row0), int(col0)))

    def push_raw(self, indices, frames):
        """Raw uint16 frames [k, H, W] -> corrected, filtered and stored views."""
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        f = np.ascontiguousarray(frames, dtype=np.uint16)
        n = ctypes.c_int(0)
        check(self.lib.mr_push_raw(self.h, ptr(idx), len(idx), ptr(f), 0, ctypes.byref(n)))
        return n.value

    def push_raw_device(self, indices, dptr):
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        n = ctypes.c_int(0)
        check(self.lib.mr_push_raw(self.h, ptr(idx), len(idx), ctypes.c_void_p(dptr), 1,
                                   ctypes.byref(n)))
        return n.value

    def nonzero_percentiles(self, vol, q, dptr=None, n=None):
        """np.percentile(vol[vol != 0], 100 q) for fractions q (host array or device
        pointer + count); returns (values float64, non-zero count)."""
        qq = np.ascontiguousarray(q, dtype=np.float64).ravel()
        out = np.zeros(len(qq))
        m = ctypes.c_int64(0)
        if dptr is None:
            v = np.ascontiguousarray(vol, dtype=np.float32).ravel()
            check(self.lib.mr_nonzero_percentiles(self.h, ptr(v), v.size, 0, ptr(qq), len(qq),
                                                  ptr(out), ctypes.byref(m)))
        else:
            check(self.lib.mr_nonzero_percentiles(self.h, ctypes.c_void_p(dptr), int(n), 1,
                                                  ptr(qq), len(qq), ptr(out), ctypes.byref(m)))
        return out, int(m.value)

    def clip_histogram(self, vol, lo, hi, edges, dptr=None, n=None):
        e = np.ascontiguousarray(edges, dtype=np.float64)
        hist = np.zeros(len(e) - 1, dtype=np.int64)
        if dptr is None:
            v = np.ascontiguousarray(vol, dtype=np.float32).ravel()
            check(self.lib.mr_clip_histogram(self.h, ptr(v), v.size, 0, float(lo), float(hi),
                                             len(hist), ptr(e), ptr(hist)))
        else:
            check(self.lib.mr_clip_histogram(self.h, ctypes.c_void_p(dptr), int(n), 1, float(lo),
                                             float(hi), len(hist), ptr(e), ptr(hist)))
        return hist

    BP_CONFIG_KEYS = ("kernel", "ppt", "layout", "depth")

    def set_bp_config(self, kernel=4, ppt=2, layout=0, depth=2):
        check(self.lib.mr_set_bp_config(self.h, int(kernel), int(ppt), int(layout), int(depth)))

    def reset_bp_config(self):
        """Back to the default configuration of the geometry's storage form and beam."""
        check(self.lib.mr_set_bp_config(self.h, 0, 0, 0, 0))

    def bp_config(self):
        cfg = np.zeros(4, np.int32)
        check(self.lib.mr_get_bp_config(self.h, ptr(cfg)))
        return dict(zip(self.BP_CONFIG_KEYS, (int(x) for x in cfg)))

    def set_ring_copy_budget(self, nbytes):
        """Cap on the device bytes of the ring copies kernels 6 and 7 read (None or < 0:
        no cap); copies beyond it are freed and requests run without them."""
        check(self.lib.mr_set_ring_copy_budget(self.h, -1 if nbytes is None else int(nbytes)))

    DEVICE_BYTES_KEYS = ("ring", "band_copy", "pair_copy", "filter_scratch")

    def device_bytes(self):
        """Device bytes held now by the ring, its band and row-pair copies and the
        four-step filter scratch."""
        out = np.zeros(4, np.int64)
        check(self.lib.mr_device_bytes(self.h, ptr(out)))
        return dict(zip(self.DEVICE_BYTES_KEYS, (int(x) for x in out)))

    def set_views_per_pass(self, views):
        """Views per part of requests on kernels 4/6/7 (each part summed in view
        order, the parts added in part order; 0: one sequential sum)."""
        check(self.lib.mr_set_bp_views_per_pass(self.h, int(views)))

    def views_per_pass(self):
        v = ctypes.c_int(0)
        check(self.lib.mr_get_bp_views_per_pass(self.h, ctypes.byref(v)))
        return v.value

    def autotune(self, tasks, n_views, out_len, max_tiles=512, reps=2):
        """Time every back-projection configuration on (a sample of) tasks, keep the
        fastest; returns (configuration dict, ms per run)."""
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        cfg = np.zeros(4, np.int32)
        ms = ctypes.c_float(0.0)
        check(self.lib.mr_autotune(self.h, ptr(T), len(T), int(n_views), int(out_len),
                                   int(max_tiles), int(reps), ptr(cfg), ctypes.byref(ms)))
        return dict(zip(self.BP_CONFIG_KEYS, (int(x) for x in cfg))), float(ms.value)

    def store(self, indices, q):
        """Append already-filtered views q [k, rows, cols]."""
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        qq = np.ascontiguousarray(q, dtype=np.float32)
        n = ctypes.c_int(0)
        check(self.lib.mr_store_views(self.h, ptr(idx), len(idx), ptr(qq), 0, ctypes.byref(n)))
        return n.value

    def store_device(self, indices, dptr):
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        n = ctypes.c_int(0)
        check(self.lib.mr_store_views(self.h, ptr(idx), len(idx), ctypes.c_void_p(dptr), 1,
                                      ctypes.byref(n)))
        return n.value

    def set_preprocess(self, gaussian, tilt):
        check(self.lib.mr_set_preprocess(self.h, float(gaussian), float(tilt)))

    def preprocess(self, block, gaussian, tilt):
        b = np.ascontiguousarray(block, dtype=np.float32)
        R, k, C = b.shape
        out = np.empty_like(b)
        check(self.lib.mr_preprocess(self.h, ptr(b), R, k, C, float(gaussian), float(tilt),
                                     ptr(out)))
        return out

    def num_views(self):
        n = ctypes.c_int(0)
        check(self.lib.mr_num_views(self.h, ctypes.byref(n)))
        return n.value

    # ---------------------------------------------------------------- requests
    def bp_points(self, pts, cor_shift, n_views):
        P = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(P), dtype=np.float64)
        if len(P):
            check(self.lib.mr_bp_points(self.h, ptr(P), len(P), float(cor_shift), int(n_views),
                                        ptr(out)))
        return out

    def bp_tasks(self, tasks, n_views, out_len):
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        out = np.zeros(int(out_len), dtype=np.float32)
        if len(T) and out_len:
            check(self.lib.mr_bp_tasks(self.h, ptr(T), len(T), int(n_views), ptr(out),
                                       int(out_len), 0))
        return out

    def bp_tasks_device(self, tasks, n_views, dptr, out_len):
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        check(self.lib.mr_bp_tasks(self.h, ptr(T), len(T), int(n_views), ctypes.c_void_p(dptr),
                                   int(out_len), 1))

    def bp_tasks_adjoint(self, tasks, n_views, g, rows, cols):
        """Host: g [len] float32 -> gq [n_views, rows, cols] float32."""
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        gg = np.ascontiguousarray(g, dtype=np.float32).ravel()
        gq = np.empty((int(n_views), int(rows), int(cols)), dtype=np.float32)
        check(self.lib.mr_bp_tasks_adjoint(self.h, ptr(T), len(T), int(n_views), ptr(gg), len(gg),
                                           0, ptr(gq), 0))
        return gq

    def bp_tasks_adjoint_device(self, tasks, n_views, d_g, g_len, d_gq):
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        check(self.lib.mr_bp_tasks_adjoint(self.h, ptr(T), len(T), int(n_views),
                                           ctypes.c_void_p(d_g), int(g_len), 1,
                                           ctypes.c_void_p(d_gq), 1))

    def bp_tasks_dshift(self, tasks, n_views, out_len):
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        out = np.zeros(int(out_len), dtype=np.float32)
        if len(T) and out_len:
            check(self.lib.mr_bp_tasks_dshift(self.h, ptr(T), len(T), int(n_views), ptr(out),
                                              int(out_len), 0))
        return out

    def bp_tasks_dshift_device(self, tasks, n_views, d_out, out_len):
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        check(self.lib.mr_bp_tasks_dshift(self.h, ptr(T), len(T), int(n_views),
                                          ctypes.c_void_p(d_out), int(out_len), 1))

    def filter_adjoint(self, indices, gq):
        """Host: gq [k, rows, cols] -> graw [rows, k, cols]."""
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        g = np.ascontiguousarray(gq, dtype=np.float32)
        k, R, C = g.shape
        out = np.empty((R, k, C), dtype=np.float32)
        check(self.lib.mr_filter_adjoint(self.h, ptr(idx), len(idx), ptr(g), 0, ptr(out), 0))
        return out

    def filter_adjoint_device(self, indices, d_gq, d_graw):
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        check(self.lib.mr_filter_adjoint(self.h, ptr(idx), len(idx), ctypes.c_void_p(d_gq), 1,
                                         ctypes.c_void_p(d_graw), 1))

    def reset_views(self):
        check(self.lib.mr_reset_views(self.h))

    def cor_scores(self, tasks, n_views):
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        scores = np.zeros(len(T), dtype=np.float64)
        if len(T):
            check(self.lib.mr_cor_scores(self.h, ptr(T), len(T), int(n_views), ptr(scores)))
        return scores

    def filter_rows(self, rows2d, tau, window):
        x = np.ascontiguousarray(rows2d, dtype=np.float32)
        R, n = x.shape
        out = np.empty_like(x)
        check(self.lib.mr_filter_rows(self.h, ptr(x), R, n, float(tau), window_code(window),
                                      ptr(out)))
        return out

    def read_views(self, first, count, rows, cols):
        out = np.empty((int(count), int(rows), int(cols)), dtype=np.float32)
        check(self.lib.mr_read_views(self.h, int(first), int(count), ptr(out)))
        return out

    def last_kernel_ms(self):
        ms = ctypes.c_double(0.0)
        check(self.lib.mr_last_kernel_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def last_tile_ms(self):
        """The tile kernel alone within the last bp_tasks call (HIP events)."""
        ms = ctypes.c_double(0.0)
        check(self.lib.mr_last_tile_kernel_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def l2_gather_bandwidth(self, lines_per_wave, buffer_bytes=1 << 20, lane_group=1):
        """(GB/s, ms) of scattered 16-byte loads served from an L2-resident buffer, with
        lines_per_wave distinct 128-byte lines per load instruction (lane_group > 1: runs
        of that many consecutive lanes on one line)."""
        g = ctypes.c_double(0.0)
        ms = ctypes.c_double(0.0)
        check(self.lib.mr_l2_gather_bandwidth(self.h, int(lines_per_wave), int(lane_group),
                                              int(buffer_bytes), ctypes.byref(g),
                                              ctypes.byref(ms)))
        return g.value, ms.value

    def last_tile_launches(self):
        """Tile-kernel dispatches within the last bp_tasks call (kernels 4 and 6)."""
        n = ctypes.c_int(0)
        check(self.lib.mr_last_tile_kernel_launches(self.h, ctypes.byref(n)))
        return n.value

    def last_bp_kernel(self):
        """Back-projection kernel of the last bp_tasks call: 2, 4-7 (plane tiles), 8 or 9
        (z columns of a volume request, gathered from the ring / staged through LDS), 0
        when nothing was launched."""
        k = ctypes.c_int(0)
        check(self.lib.mr_last_bp_kernel(self.h, ctypes.byref(k)))
        return k.value

    # ---------------------------------------------------------------- device memory
    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        check(self.lib.mr_device_alloc(self.h, int(nbytes), ctypes.byref(p)))
        return p.value

    def free(self, dptr):
        check(self.lib.mr_device_free(self.h, ctypes.c_void_p(dptr)))

    def h2d(self, dptr, arr):
        a = np.ascontiguousarray(arr)
        check(self.lib.mr_memcpy(self.h, ctypes.c_void_p(dptr), ptr(a), a.nbytes, 0))

    def d2h(self, arr, dptr):
        check(self.lib.mr_memcpy(self.h, ptr(arr), ctypes.c_void_p(dptr), arr.nbytes, 1))
        return arr

    def synchronize(self):
        check(self.lib.mr_synchronize(self.h))

    def phantom_cone(self, ellipsoids, vectors, rows, cols, dptr):
        E = np.ascontiguousarray(ellipsoids, dtype=np.float64).reshape(-1, 13)
        V = np.ascontiguousarray(vectors, dtype=np.float64).reshape(-1, 12)
        check(self.lib.mr_phantom_cone(self.h, ptr(E), len(E), ptr(V), len(V), int(rows),
                                       int(cols), ctypes.c_void_p(dptr)))

    def phantom_parallel(self, ellipsoids, angles, rows, cols, pu, pv, far, dptr):
        E = np.ascontiguousarray(ellipsoids, dtype=np.float64).reshape(-1, 13)
        a = np.ascontiguousarray(angles, dtype=np.float64).ravel()
        check(self.lib.mr_phantom_parallel(self.h, ptr(E), len(E), ptr(a), len(a), int(rows),
                                           int(cols), float(pu), float(pv), float(far),
                                           ctypes.c_void_p(dptr)))


class Plan:
    """A compiled batch of tasks on a device work queue (mr_plan), back-projected by a
    persistent launch.  ``run(n)``: a full pass over the first n stored views;
    ``stream(indices, block)``: ingest a batch and bring the output up to all views from
    one replayed HIP graph; ``update()``: bring it up to all views however they were
    ingested; both add only the views not yet in the plan's running FP32 sums (the same
    bits as ``run``).  ``cor_scores(n)``: entropy scores of a plan made of COR trial
    images, one replayed graph."""

    def __init__(self, ctx, tasks, out_len):
        self.ctx = ctx
        self.lib = ctx.lib
        T = np.ascontiguousarray(tasks, dtype=TASK_DTYPE)
        h = ctypes.c_void_p()
        check(self.lib.mr_plan_create(ctx.h, ptr(T), len(T), int(out_len), ctypes.byref(h)))
        self.h = h
        self.out_len = int(out_len)
        self.ntasks = len(T)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.mr_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n_views=-1, copy=True):
        """Output over the first n_views stored views (-1: all), a full pass."""
        out = np.empty(self.out_len, dtype=np.float32) if copy else None
        check(self.lib.mr_plan_run(self.h, int(n_views), ptr(out)))
        return out

    def stream(self, indices, block, copy=True):
        """Ingest views `indices` (block [rows, k, cols] float32) and return the output
        over all stored views."""
        idx = np.ascontiguousarray(indices, dtype=np.int64).ravel()
        blk = np.ascontiguousarray(block, dtype=np.float32)
        out = np.empty(self.out_len, dtype=np.float32) if copy else None
        check(self.lib.mr_plan_stream(self.h, ptr(idx), len(idx), ptr(blk), ptr(out)))
        return out

    def update(self, copy=True):
        """Output over all stored views, adding only the views not yet in the plan's
        running sums (however they were ingested)."""
        out = np.empty(self.out_len, dtype=np.float32) if copy else None
        check(self.lib.mr_plan_update(self.h, ptr(out)))
        return out

    def cor_scores(self, n_views=-1):
        """Entropy scores of the plan's trial images (one graph: persistent launch +
        device scoring), from the first n_views stored views (-1: all)."""
        sc = np.empty(self.ntasks, np.float64)
        check(self.lib.mr_plan_cor_scores(self.h, int(n_views), ptr(sc)))
        return sc

    def output_ptr(self):
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        check(self.lib.mr_plan_output(self.h, ctypes.byref(p), ctypes.byref(n)))
        return p.value

    def last_ms(self):
        ms = ctypes.c_double()
        check(self.lib.mr_plan_last_ms(self.h, ctypes.byref(ms)))
        return ms.value
