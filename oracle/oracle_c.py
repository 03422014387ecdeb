"""ctypes loader of the C oracle builds (TEST INFRASTRUCTURE ONLY; see oracle_c.c)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}
_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib(kind="exact"):
    """kind = 'exact' (float64 oracle) or 'fast' (Numba-equivalent CPU baseline)."""
    if kind in _libs:
        return _libs[kind]
    path = os.path.join(HERE, f"liboracle_{kind}.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.orc_ramp_response.argtypes = [_i, _d, _i, _p]
    L.orc_filter_rows.argtypes = [_p, _i, _i, _d, _i, _p]
    L.orc_fdk_filter_view.argtypes = [_p, _i, _i, _i, _i, _p, _i, _p]
    L.orc_bp_cone.argtypes = [_p, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _d, _p]
    L.orc_bp_parallel.argtypes = [_p, _i, _i, _i, _p, _p, _d, _d, _p, _i64, _d, _d, _p]
    L.orc_bp_cone_acc_view.argtypes = [_p, _i, _i, _p, _d, _p, _i64, _p]
    L.orc_bp_parallel_acc.argtypes = [_p, _i, _i, _i, _i64, _i64, _p, _p, _d, _d, _d, _p, _i64,
                                      _p]
    L.orc_set_num_threads.argtypes = [_i]
    _libs[kind] = L
    return L


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


WINDOW = {"ram-lak": 0, "shepp-logan": 1, "hann": 2, "cosine": 3}


def set_threads(n, kind="exact"):
    lib(kind).orc_set_num_threads(int(n))


def filter_rows(x, tau, window, kind="exact"):
    x = np.ascontiguousarray(x, np.float32)
    rows = x.reshape(-1, x.shape[-1])
    out = np.empty(rows.shape, np.float64)
    rc = lib(kind).orc_filter_rows(_ptr(rows), rows.shape[0], rows.shape[1], float(tau),
                                   WINDOW[window], _ptr(out))
    if rc:
        raise ValueError("bad window")
    return out.reshape(x.shape)


def fdk_filter_view(block, v, vector, window, kind="exact"):
    """View v of a [R][k][C] float32 block -> float64 [R][C]."""
    block = np.ascontiguousarray(block, np.float32)
    R, kb, C = block.shape
    V = np.ascontiguousarray(vector, np.float64)
    out = np.empty((R, C), np.float64)
    rc = lib(kind).orc_fdk_filter_view(_ptr(block), R, kb, C, int(v), _ptr(V), WINDOW[window],
                                       _ptr(out))
    if rc:
        raise ValueError("bad window")
    return out


def bp_cone(q, S, D, UU, VV, N, SDD, SOD, pts, scale, kind="exact"):
    q = np.ascontiguousarray(q, np.float32)
    A, R, C = q.shape
    arr = [np.ascontiguousarray(a, np.float64) for a in (S, D, UU, VV, N, SDD, SOD)]
    P = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.empty(len(P), np.float64)
    lib(kind).orc_bp_cone(_ptr(q), A, R, C, *[_ptr(a) for a in arr], _ptr(P), len(P),
                          float(scale), _ptr(out))
    return out


def bp_parallel(q, cos_t, sin_t, pu, pv, pts, scale, col_shift, kind="exact"):
    q = np.ascontiguousarray(q, np.float32)
    A, R, C = q.shape
    c = np.ascontiguousarray(cos_t, np.float64)
    s = np.ascontiguousarray(sin_t, np.float64)
    P = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.empty(len(P), np.float64)
    lib(kind).orc_bp_parallel(_ptr(q), A, R, C, _ptr(c), _ptr(s), float(pu), float(pv), _ptr(P),
                              len(P), float(scale), float(col_shift), _ptr(out))
    return out


def bp_cone_acc_view(q64, vector, shift, pts, acc, kind="exact"):
    q64 = np.ascontiguousarray(q64, np.float64)
    R, C = q64.shape
    V = np.ascontiguousarray(vector, np.float64)
    P = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    assert acc.dtype == np.float64 and acc.flags.c_contiguous and len(acc) == len(P)
    lib(kind).orc_bp_cone_acc_view(_ptr(q64), R, C, _ptr(V), float(shift), _ptr(P), len(P),
                                   _ptr(acc))
    return acc


def bp_parallel_acc(q64, cos_t, sin_t, pu, pv, col_shift, pts, acc, kind="exact"):
    """acc += parallel-beam back-projection (unscaled) of float64 views q64, a
    [views, rows, cols] array of any strides with unit column stride."""
    assert q64.dtype == np.float64 and q64.ndim == 3 and q64.strides[2] == 8
    A, R, C = q64.shape
    c = np.ascontiguousarray(cos_t, np.float64)
    s = np.ascontiguousarray(sin_t, np.float64)
    assert len(c) == A and len(s) == A
    P = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    assert acc.dtype == np.float64 and acc.flags.c_contiguous and len(acc) == len(P)
    lib(kind).orc_bp_parallel_acc(ctypes.c_void_p(q64.ctypes.data), A, R, C, q64.strides[0] // 8,
                                  q64.strides[1] // 8, _ptr(c), _ptr(s), float(pu), float(pv),
                                  float(col_shift), _ptr(P), len(P), _ptr(acc))
    return acc
