"""ctypes bindings of libgalp (include/galp.h, include/galp_host.h).

Thin plumbing for the tests and the benchmark: every call goes straight to the
C-ABI, every non-zero status raises GalpError with the library's message.
The library must have been built (``__graft_entry__.build()`` or
``make -C galp_amd``); importing this module without it raises, there is no
fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgalp.so")

GALP_OK = 0
STATUS_NAMES = {0: "GALP_OK", -1: "GALP_EINVAL", -2: "GALP_ENOMEM", -3: "GALP_EHIP",
                -4: "GALP_EUNSUPPORTED", -5: "GALP_EIO", -6: "GALP_ESTATE"}

METHOD_SORT, METHOD_HASH = 0, 1
POLICY_INCORE, POLICY_OOC_SYNC, POLICY_OOC_ASYNC, POLICY_LI = 0, 1, 2, 4
DISPATCH_PLAIN, DISPATCH_LEARNED = 0, 1
# galp_graph_dispatch_stats slots
DS_CAND_TILES, DS_OVERFLOW, DS_CAND_VERTS, DS_PLAIN_HUBS = 12, 16, 17, 18


class GalpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class CsrView(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("offsets", ctypes.c_void_p),
                ("neighbors", ctypes.c_void_p), ("offsets_are_64", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Opts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("policy", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("buffer_pow", ctypes.c_int32), ("dispatch", ctypes.c_int32), ("reorder", ctypes.c_int32),
                ("profile", ctypes.c_int32), ("graph_capture", ctypes.c_int32), ("mem_budget", ctypes.c_int64),
                ("offsets64", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class StepStats(ctypes.Structure):
    _fields_ = [("updated", ctypes.c_int64), ("overflowed", ctypes.c_int64), ("device_ms", ctypes.c_double),
                ("iteration", ctypes.c_int32), ("stopped", ctypes.c_int32)]


class RunResult(ctypes.Structure):
    _fields_ = [("f1_s", ctypes.c_double), ("f2_s", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("iterations", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class GraphInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("offsets_are_64", ctypes.c_int32),
                ("nbuckets", ctypes.c_int32), ("bucket_vertices", ctypes.c_int64 * 16),
                ("bucket_edges", ctypes.c_int64 * 16), ("device_bytes", ctypes.c_int64),
                ("preprocess_ms", ctypes.c_double)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32), ("compute_units", ctypes.c_int32),
                ("wavefront_size", ctypes.c_int32), ("lds_per_block", ctypes.c_int64),
                ("total_mem", ctypes.c_int64), ("l2_bytes", ctypes.c_int64)]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_int64), ("total_ms", ctypes.c_double),
                ("alg_bytes", ctypes.c_double)]


class HostCsr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("offsets", ctypes.POINTER(ctypes.c_int64)),
                ("neighbors", ctypes.POINTER(ctypes.c_int32))]


_lib = None


def lib():
    """Loads libgalp.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GalpError(-4, f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "galp_opts_default": (None, [ctypes.POINTER(Opts)]),
        "galp_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "galp_close": (ctypes.c_int, [vp]),
        "galp_device_query": (ctypes.c_int, [vp, ctypes.POINTER(DeviceInfo)]),
        "galp_graph_upload": (ctypes.c_int, [vp, ctypes.POINTER(CsrView), ctypes.POINTER(Opts), ctypes.POINTER(vp)]),
        "galp_graph_free": (ctypes.c_int, [vp]),
        "galp_graph_query": (ctypes.c_int, [vp, ctypes.POINTER(GraphInfo)]),
        "galp_lp_init": (ctypes.c_int, [vp]),
        "galp_lp_step": (ctypes.c_int, [vp, ctypes.POINTER(StepStats)]),
        "galp_lp_run": (ctypes.c_int, [vp, i32, i32, vp, ctypes.POINTER(RunResult)]),
        "galp_lp_get_labels": (ctypes.c_int, [vp, vp]),
        "galp_lp_set_labels": (ctypes.c_int, [vp, vp]),
        "galp_profile_read": (ctypes.c_int, [vp, ctypes.POINTER(KernelStat), i32, ctypes.POINTER(i32)]),
        "galp_profile_reset": (ctypes.c_int, [vp]),
        "galp_graph_gather_probe": (ctypes.c_int, [vp, i32, ctypes.POINTER(dbl), ctypes.POINTER(ctypes.c_uint64)]),
        "galp_last_error": (ctypes.c_char_p, [vp]),
        "galp_load_graph": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(HostCsr))]),
        "galp_save_bin": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(HostCsr)]),
        "galp_gen_planted_partition": (ctypes.c_int, [i64, i64, i32, i32, ctypes.c_uint64,
                                                      ctypes.POINTER(ctypes.POINTER(HostCsr))]),
        "galp_gen_rmat": (ctypes.c_int, [i32, i32, dbl, dbl, dbl, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.POINTER(HostCsr))]),
        "galp_gen_batch_planted": (ctypes.c_int, [i64, i64, i64, i64, i64, i32, i32, dbl, ctypes.c_uint64,
                                                  ctypes.POINTER(ctypes.POINTER(HostCsr)), vp, vp]),
        "galp_gen_lfr": (ctypes.c_int, [i64, dbl, dbl, i32, dbl, i32, i32, dbl, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.POINTER(HostCsr)),
                                        ctypes.POINTER(ctypes.POINTER(ctypes.c_int32))]),
        "galp_host_csr_from_arrays": (ctypes.c_int, [i64, i64, vp, vp, ctypes.POINTER(ctypes.POINTER(HostCsr))]),
        "galp_cpu_lp_run": (ctypes.c_int, [ctypes.POINTER(HostCsr), i32, vp, ctypes.POINTER(dbl),
                                           ctypes.POINTER(i32)]),
        "galp_host_csr_free": (None, [ctypes.POINTER(HostCsr)]),
        "galp_host_free": (None, [vp]),
        "galp_host_last_error": (ctypes.c_char_p, []),
        "galp_accuracy": (ctypes.c_int, [ctypes.c_char_p, vp, i64, ctypes.POINTER(dbl), ctypes.POINTER(dbl),
                                         ctypes.POINTER(dbl)]),
        "galp_mlp_num_params": (i64, [i32]),
        "galp_mlp_init": (ctypes.c_int, [i32, ctypes.c_uint64, vp]),
        "galp_mlp_predict": (ctypes.c_int, [vp, i32, vp, vp, i64, vp]),
        "galp_mlp_loss_grad_noalloc": (ctypes.c_int, [vp, i32, vp, vp, vp, i64, vp, vp, vp]),
        "galp_mlp_loss_grad": (ctypes.c_int, [vp, i32, vp, vp, vp, i64, vp, ctypes.POINTER(ctypes.c_float)]),
        "galp_mlp_train": (ctypes.c_int, [vp, i32, vp, vp, vp, i64, i32, i32, ctypes.c_float, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_float)]),
        "galp_graph_set_mlp": (ctypes.c_int, [vp, i32, vp, ctypes.c_float]),
        "galp_graph_log_enable": (ctypes.c_int, [vp, i64]),
        "galp_graph_log_read": (ctypes.c_int, [vp, vp, vp, i64, ctypes.POINTER(i64)]),
        "galp_graph_features": (ctypes.c_int, [vp, vp, vp, vp]),
        "galp_graph_dispatch_stats": (ctypes.c_int, [vp, vp, i32, ctypes.POINTER(i32)]),
        "galp_graph_set_batch": (ctypes.c_int, [vp, i64, vp]),
        "galp_graph_autotune": (ctypes.c_int, [vp, i32, ctypes.c_char_p, vp]),
        "galp_graph_set_learned_iterations": (ctypes.c_int, [vp, ctypes.c_uint64]),
        "galp_graph_set_tuning": (ctypes.c_int, [vp, vp]),
        "galp_batch_iterations": (ctypes.c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check_host(rc):
    if rc != GALP_OK:
        raise GalpError(rc, lib().galp_host_last_error().decode())


# ------------------------------------------------------------------ host ---
def _take_csr(ptr):
    c = ptr.contents
    n, m = int(c.n), int(c.m)
    off = np.ctypeslib.as_array(c.offsets, shape=(n + 1,)).copy()
    nbr = np.ctypeslib.as_array(c.neighbors, shape=(max(m, 1),))[:m].copy() if m else np.zeros(0, np.int32)
    lib().galp_host_csr_free(ptr)
    return off, nbr


def gen_planted_partition(n, community, k_intra, k_inter, seed):
    p = ctypes.POINTER(HostCsr)()
    _check_host(lib().galp_gen_planted_partition(n, community, k_intra, k_inter, seed, ctypes.byref(p)))
    return _take_csr(p)


def gen_batch_planted(count, n_range=(1000, 200000), comm_range=(32, 1024), deg_range=(8, 48), inter_frac=0.25,
                      seed=4096):
    """`count` planted-partition graphs as one disjoint-union CSR.  Returns
    (off, nbr, vbase, spec): graph k owns vertices [vbase[k], vbase[k+1]) and
    is gen_planted_partition(*spec[k]) shifted by vbase[k]."""
    p = ctypes.POINTER(HostCsr)()
    vbase = np.zeros(count + 1, np.int64)
    spec = np.zeros((count, 5), np.int64)
    _check_host(lib().galp_gen_batch_planted(count, n_range[0], n_range[1], comm_range[0], comm_range[1],
                                             deg_range[0], deg_range[1], inter_frac, seed, ctypes.byref(p),
                                             vbase.ctypes.data, spec.ctypes.data))
    off, nbr = _take_csr(p)
    return off, nbr, vbase, spec


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    p = ctypes.POINTER(HostCsr)()
    _check_host(lib().galp_gen_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(p)))
    return _take_csr(p)


def gen_lfr(n, gamma=2.5, avg_deg=20.0, max_deg=2000, beta=1.5, cmin=50, cmax=50000, mu=0.3, seed=8,
            with_communities=False):
    p = ctypes.POINTER(HostCsr)()
    cp = ctypes.POINTER(ctypes.c_int32)()
    _check_host(lib().galp_gen_lfr(n, gamma, avg_deg, max_deg, beta, cmin, cmax, mu, seed, ctypes.byref(p),
                                   ctypes.byref(cp) if with_communities else None))
    off, nbr = _take_csr(p)
    if not with_communities:
        return off, nbr
    comm = np.ctypeslib.as_array(cp, shape=(n,)).copy()
    lib().galp_host_free(ctypes.cast(cp, ctypes.c_void_p))
    return off, nbr, comm


def load_graph(path):
    p = ctypes.POINTER(HostCsr)()
    _check_host(lib().galp_load_graph(os.fsencode(path), ctypes.byref(p)))
    return _take_csr(p)


def save_bin(path, off, nbr):
    off = np.ascontiguousarray(off, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    c = HostCsr(len(off) - 1, len(nbr), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    _check_host(lib().galp_save_bin(os.fsencode(path), ctypes.byref(c)))


def cpu_lp_run(off, nbr, niter=10):
    """The serial CPU path (galp_cpu_lp_run): returns (labels, iterations, seconds)."""
    off = np.ascontiguousarray(off, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    c = HostCsr(len(off) - 1, len(nbr), off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                nbr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    out = np.empty(len(off) - 1, np.int32)
    secs, iters = ctypes.c_double(), ctypes.c_int32()
    _check_host(lib().galp_cpu_lp_run(ctypes.byref(c), int(niter), out.ctypes.data, ctypes.byref(secs),
                                      ctypes.byref(iters)))
    return out, iters.value, secs.value


def accuracy(truth_file, labels):
    labels = np.ascontiguousarray(labels, np.int32)
    nmi, fm, ari = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _check_host(lib().galp_accuracy(os.fsencode(truth_file), labels.ctypes.data, len(labels),
                                    ctypes.byref(nmi), ctypes.byref(fm), ctypes.byref(ari)))
    return nmi.value, fm.value, ari.value


# ---------------------------------------------------------------- device ---
class Context:
    """A galp_ctx on one device (galp_open / galp_close)."""

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        rc = lib().galp_open(device, ctypes.byref(self.h))
        if rc != GALP_OK:
            raise GalpError(rc, f"galp_open({device}) failed")

    def check(self, rc):
        if rc != GALP_OK:
            raise GalpError(rc, lib().galp_last_error(self.h).decode())

    def device_info(self):
        d = DeviceInfo()
        self.check(lib().galp_device_query(self.h, ctypes.byref(d)))
        return {"name": d.name.decode(), "arch": d.arch.decode(), "compute_units": d.compute_units,
                "wavefront_size": d.wavefront_size, "lds_per_block": d.lds_per_block,
                "total_mem": d.total_mem, "l2_bytes": d.l2_bytes}

    def close(self):
        if self.h:
            lib().galp_close(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def make_opts(method=METHOD_HASH, policy=POLICY_INCORE, variant=1, buffer_pow=24, dispatch=DISPATCH_PLAIN,
              reorder=1, profile=0, mem_budget=0, graph_capture=0, offsets64=0):
    o = Opts()
    lib().galp_opts_default(ctypes.byref(o))
    o.method, o.policy, o.variant, o.buffer_pow = method, policy, variant, buffer_pow
    o.dispatch, o.reorder, o.profile, o.mem_budget = dispatch, reorder, profile, mem_budget
    o.graph_capture = graph_capture
    o.offsets64 = offsets64
    return o


class Graph:
    """A galp_graph: a CSR resident on the device plus its dispatch plan."""

    def __init__(self, ctx, off, nbr, **opts):
        self.ctx = ctx
        self.off = np.ascontiguousarray(off, np.int64)
        self.nbr = np.ascontiguousarray(nbr, np.int32)
        self.n = len(self.off) - 1
        self.m = len(self.nbr)
        view = CsrView(self.n, self.m, self.off.ctypes.data, self.nbr.ctypes.data if self.m else None, 1, 0)
        o = make_opts(**opts)
        self.h = ctypes.c_void_p()
        ctx.check(lib().galp_graph_upload(ctx.h, ctypes.byref(view), ctypes.byref(o), ctypes.byref(self.h)))

    def free(self):
        if self.h:
            lib().galp_graph_free(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def info(self):
        i = GraphInfo()
        self.ctx.check(lib().galp_graph_query(self.h, ctypes.byref(i)))
        nb = i.nbuckets
        return {"n": i.n, "m": i.m, "offsets_are_64": i.offsets_are_64,
                "bucket_vertices": list(i.bucket_vertices[:nb]), "bucket_edges": list(i.bucket_edges[:nb]),
                "device_bytes": i.device_bytes, "preprocess_ms": i.preprocess_ms}

    def init(self):
        self.ctx.check(lib().galp_lp_init(self.h))

    def step(self):
        st = StepStats()
        self.ctx.check(lib().galp_lp_step(self.h, ctypes.byref(st)))
        return {"updated": st.updated, "overflowed": st.overflowed, "device_ms": st.device_ms,
                "iteration": st.iteration, "stopped": bool(st.stopped)}

    def set_labels(self, labels):
        labels = np.ascontiguousarray(labels, np.int32)
        assert labels.shape == (self.n,)
        self.ctx.check(lib().galp_lp_set_labels(self.h, labels.ctypes.data))

    def labels(self):
        out = np.empty(self.n, np.int32)
        self.ctx.check(lib().galp_lp_get_labels(self.h, out.ctypes.data))
        return out

    def run(self, niter=10, verbose=0, want_labels=True):
        out = np.empty(self.n, np.int32) if want_labels else None
        r = RunResult()
        self.ctx.check(lib().galp_lp_run(self.h, niter, verbose, out.ctypes.data if out is not None else None,
                                         ctypes.byref(r)))
        return out, {"f1_s": r.f1_s, "f2_s": r.f2_s, "device_ms": r.device_ms, "iterations": r.iterations}

    def profile(self):
        buf = (KernelStat * 64)()
        cnt = ctypes.c_int32()
        self.ctx.check(lib().galp_profile_read(self.h, buf, 64, ctypes.byref(cnt)))
        return [{"name": buf[i].name.decode(), "launches": buf[i].launches, "total_ms": buf[i].total_ms,
                 "alg_bytes": buf[i].alg_bytes} for i in range(min(cnt.value, 64))]

    def profile_reset(self):
        self.ctx.check(lib().galp_profile_reset(self.h))

    def gather_probe(self, reps=5):
        """(ms per pass, checksum) of the gather-only probe over this graph."""
        ms = ctypes.c_double()
        cs = ctypes.c_uint64()
        self.ctx.check(lib().galp_graph_gather_probe(self.h, reps, ctypes.byref(ms), ctypes.byref(cs)))
        return ms.value, cs.value


    # ---- autotuner ---------------------------------------------------------
    def autotune(self, reps=3, cache_path=None):
        out = np.zeros(16, np.int32)
        self.ctx.check(lib().galp_graph_autotune(self.h, reps, os.fsencode(cache_path) if cache_path else None,
                                                 out.ctypes.data))
        self.tuned_streams = int(out[12])   # 1 or 3 streams per iteration (0: not timed)
        # learned dispatch: iterations the judge left to the network
        self.learned_iterations = int(np.uint32(out[13])) | (int(np.uint32(out[14])) << 32)
        return out[:12]

    def set_learned_iterations(self, mask):
        self.ctx.check(lib().galp_graph_set_learned_iterations(self.h, ctypes.c_uint64(mask & (2 ** 64 - 1))))

    def set_tuning(self, options):
        o = np.ascontiguousarray(options, np.int32)
        assert len(o) == 12
        self.ctx.check(lib().galp_graph_set_tuning(self.h, o.ctypes.data))

    # ---- batched runs ------------------------------------------------------
    def set_batch(self, vbase):
        vb = np.ascontiguousarray(vbase, np.int64)
        self.ngraphs = len(vb) - 1
        self.ctx.check(lib().galp_graph_set_batch(self.h, self.ngraphs, vb.ctypes.data))

    def batch_iterations(self):
        out = np.zeros(self.ngraphs, np.int32)
        self.ctx.check(lib().galp_batch_iterations(self.h, out.ctypes.data))
        return out

    # ---- learned dispatch --------------------------------------------------
    def set_mlp(self, hidden, params, cand_thresh):
        p = np.ascontiguousarray(params, np.float32)
        assert len(p) == mlp_num_params(hidden)
        self.ctx.check(lib().galp_graph_set_mlp(self.h, hidden, p.ctypes.data, float(cand_thresh)))

    def log_enable(self, capacity):
        self.ctx.check(lib().galp_graph_log_enable(self.h, int(capacity)))

    def log_read(self):
        cnt = ctypes.c_int64()
        self.ctx.check(lib().galp_graph_log_read(self.h, None, None, 0, ctypes.byref(cnt)))
        k = cnt.value
        x = np.empty((max(k, 1), 4), np.float32)
        y = np.empty(max(k, 1), np.float32)
        self.ctx.check(lib().galp_graph_log_read(self.h, x.ctypes.data, y.ctypes.data, k, ctypes.byref(cnt)))
        return x[:k], y[:k]

    def features(self):
        d = np.empty(max(self.n, 1), np.float32)
        mean = np.empty_like(d)
        mx = np.empty_like(d)
        self.ctx.check(lib().galp_graph_features(self.h, d.ctypes.data, mean.ctypes.data, mx.ctypes.data))
        return d[:self.n], mean[:self.n], mx[:self.n]

    def dispatch_stats(self):
        out = np.zeros(32, np.int64)
        cnt = ctypes.c_int32()
        self.ctx.check(lib().galp_graph_dispatch_stats(self.h, out.ctypes.data, 32, ctypes.byref(cnt)))
        return out[:cnt.value]


# ------------------------------------------------------------------- MLP ---
def mlp_num_params(hidden):
    return int(lib().galp_mlp_num_params(hidden))


def mlp_init(hidden, seed):
    p = np.empty(mlp_num_params(hidden), np.float32)
    rc = lib().galp_mlp_init(hidden, seed, p.ctypes.data)
    if rc != GALP_OK:
        raise GalpError(rc, "galp_mlp_init")
    return p


def mlp_predict(ctx, hidden, params, x):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
    y = np.empty(max(len(x), 1), np.float32)
    ctx.check(lib().galp_mlp_predict(ctx.h, hidden, p.ctypes.data, x.ctypes.data, len(x), y.ctypes.data))
    return y[:len(x)]


def mlp_loss_grad(ctx, hidden, params, x, y):
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
    y = np.ascontiguousarray(y, np.float32)
    g = np.empty(len(p), np.float32)
    loss = ctypes.c_float()
    ctx.check(lib().galp_mlp_loss_grad(ctx.h, hidden, p.ctypes.data, x.ctypes.data, y.ctypes.data, len(x),
                                       g.ctypes.data, ctypes.byref(loss)))
    return loss.value, g


def mlp_train(ctx, hidden, params, x, y, epochs=20, batch=256, lr=1e-2, seed=1):
    p = np.array(params, np.float32, copy=True)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
    y = np.ascontiguousarray(y, np.float32)
    loss = ctypes.c_float()
    ctx.check(lib().galp_mlp_train(ctx.h, hidden, p.ctypes.data, x.ctypes.data, y.ctypes.data, len(x), epochs,
                                   batch, lr, seed, ctypes.byref(loss)))
    return p, loss.value
