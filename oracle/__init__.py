"""CPU oracle for galp's label-propagation update (ctypes wrapper over lp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the parity reference or the timed CPU baseline.
The product (libgalp, the galp CLI, galp_amd) never imports it.

Each function cites the reference code it restates; see lp_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_BASE_PATH = os.path.join(_HERE, "_build", "libbaseline.so")
_lib = None
_base = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile lp_oracle.c into oracle/_build/liboracle.so."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_lp_iterate.restype = ctypes.c_int64
        L.oracle_lp_iterate.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, _i32p,
                                        ctypes.c_int64, ctypes.c_int]
        L.oracle_lp_run.restype = ctypes.c_int
        L.oracle_lp_run.argtypes = [ctypes.c_int64, _i64p, _i32p, ctypes.c_int, _i32p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_distinct_counts.restype = None
        L.oracle_distinct_counts.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p,
                                             ctypes.c_int64, _i32p]
        L.oracle_vertex_features.restype = None
        L.oracle_vertex_features.argtypes = [ctypes.c_int64, _i64p, _i32p, _f32p, _f32p, _f32p]
        L.oracle_batch_boundaries.restype = ctypes.c_int64
        L.oracle_batch_boundaries.argtypes = [ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,
                                              ctypes.c_int64]
        L.oracle_mlp_num_params.restype = ctypes.c_int64
        L.oracle_mlp_num_params.argtypes = [ctypes.c_int]
        L.oracle_mlp_loss_grad.restype = ctypes.c_double
        L.oracle_mlp_loss_grad.argtypes = [ctypes.c_int, _f64p, ctypes.c_int64, _f64p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def baseline_lib():
    """lp_baseline.cpp: the std::map-structured CPU baseline."""
    global _base
    if _base is None:
        if not os.path.exists(_BASE_PATH):
            build()
        L = ctypes.CDLL(_BASE_PATH)
        L.baseline_map_iterate.restype = ctypes.c_int64
        L.baseline_map_iterate.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, _i32p,
                                           ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        L.baseline_max_threads.restype = ctypes.c_int
        L.baseline_max_threads.argtypes = []
        _base = L
    return _base


def map_iterate(off, nbr, prev, vlist=None, nthreads=0, out=None):
    """One synchronous iteration of the std::map-structured baseline over the
    vertices in vlist (int64; all when None).  Returns (next, changed); next
    holds prev's value at vertices outside vlist."""
    off, nbr = _csr(off, nbr)
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    n = len(off) - 1
    nxt = prev.copy() if out is None else out
    if vlist is not None:
        vlist = np.ascontiguousarray(vlist, dtype=np.int64)
    nv = n if vlist is None else len(vlist)
    c = baseline_lib().baseline_map_iterate(n, off, nbr, prev, nxt,
                                            vlist.ctypes.data if vlist is not None else None,
                                            nv, int(nthreads))
    return nxt, int(c)


def baseline_threads():
    return int(baseline_lib().baseline_max_threads())


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def _csr(off, nbr):
    off = np.ascontiguousarray(off, dtype=np.int64)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    return off, nbr


def lp_iterate(off, nbr, prev, nlabels=None, nthreads=0):
    """One synchronous LP iteration. Returns (next_labels, changed_count)."""
    off, nbr = _csr(off, nbr)
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    n = len(off) - 1
    if nlabels is None:
        nlabels = int(prev.max()) + 1 if n else 1
    nxt = np.empty(n, dtype=np.int32)
    c = lib().oracle_lp_iterate(n, off, nbr, prev, nxt, int(nlabels), int(nthreads))
    return nxt, int(c)


def lp_run(off, nbr, niter=10, nthreads=0, keep_iters=True):
    """The synchronous driver: labels[i]=i, up to niter iterations with the
    n*1e-5 break.  Returns (final_labels, per_iter [k x n] or None, counts[k])."""
    off, nbr = _csr(off, nbr)
    n = len(off) - 1
    out = np.empty(n, dtype=np.int32)
    per = np.empty((niter, n), dtype=np.int32) if keep_iters else None
    counts = np.zeros(niter, dtype=np.int64)
    k = lib().oracle_lp_run(n, off, nbr, int(niter), out,
                            per.ctypes.data if per is not None else None,
                            counts.ctypes.data, int(nthreads))
    return out, (per[:k] if per is not None else None), counts[:k]


def distinct_counts(off, nbr, prev):
    off, nbr = _csr(off, nbr)
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    n = len(off) - 1
    out = np.empty(n, dtype=np.int32)
    lib().oracle_distinct_counts(n, off, nbr, prev, int(prev.max()) + 1 if n else 1, out)
    return out


def vertex_features(off, nbr):
    off, nbr = _csr(off, nbr)
    n = len(off) - 1
    d = np.empty(n, np.float32)
    mean = np.empty(n, np.float32)
    mx = np.empty(n, np.float32)
    lib().oracle_vertex_features(n, off, nbr, d, mean, mx)
    return d, mean, mx


def batch_boundaries(off, B):
    off = np.ascontiguousarray(off, dtype=np.int64)
    n = len(off) - 1
    cap = int(off[-1] // max(B, 1)) * 2 + n + 4 if n else 4
    cap = min(cap, n + 2)
    out = np.empty(max(cap, 2), dtype=np.int64)
    k = lib().oracle_batch_boundaries(n, off, int(B), out, len(out))
    assert k > 0
    return out[:k].copy()


def mlp_num_params(H):
    return int(lib().oracle_mlp_num_params(int(H)))


def mlp_loss_grad(H, params, x, y=None):
    """float64 forward/backward of the 4->H->H->1 MLP; returns (loss, yhat, grad)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 4)
    B = x.shape[0]
    yhat = np.empty(B, np.float64)
    grad = np.empty(len(p), np.float64) if y is not None else None
    yy = np.ascontiguousarray(y, dtype=np.float64) if y is not None else None
    loss = lib().oracle_mlp_loss_grad(int(H), p, B, x,
                                      yy.ctypes.data if yy is not None else None,
                                      yhat.ctypes.data,
                                      grad.ctypes.data if grad is not None else None)
    return float(loss), yhat, grad
