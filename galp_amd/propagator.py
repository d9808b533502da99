"""Python mirror of the reference's propagator interface over the C-ABI.

LabelPropagator<V,E> (label_propagation/include/label_propagator.h:11-34)
exposes `run(niter) -> (f1, f2)` and `get_labels()`; main.cu:139-207 picks
the concrete class from method*5+policy.  `make_propagator` keeps that
mapping (same case numbers, same rejections for the hybrid and multi-GPU
cases) and every object runs on libgalp's HIP kernels.
"""
from __future__ import annotations

import numpy as np

from . import _abi

PROC_NAMES = {0: "dpp\tincore", 1: "dpp\tsync", 2: "dpp\tasync", 4: "dpp\tli",
              5: "lfht{l}\tincore", 6: "lfht{l}\tsync", 7: "lfht{l}\tasync"}


class LabelPropagator:
    """One graph on one device; `run` mirrors LFHTBase::run / DPPBase::run
    (lfht.cuh:73-112, dpp.cuh:64-102) with the synchronous update."""

    def __init__(self, ctx, offsets, neighbors, method=1, policy=0, lfht_policy=1, buffer_pow=24,
                 dispatch=_abi.DISPATCH_PLAIN, reorder=1, profile=0, mem_budget=0):
        self.ctx = ctx
        self.graph = _abi.Graph(ctx, offsets, neighbors, method=method, policy=policy, variant=lfht_policy,
                                buffer_pow=buffer_pow, dispatch=dispatch, reorder=reorder, profile=profile,
                                mem_budget=mem_budget)
        self.n = self.graph.n
        self._labels = None

    def run(self, niter=10, verbose=False):
        labels, res = self.graph.run(niter, 1 if verbose else 0, want_labels=True)
        self._labels = labels
        self.last_run = res
        return res["f1_s"], res["f2_s"]

    def get_labels(self):
        if self._labels is None:
            return self.graph.labels()
        return self._labels.copy()

    def close(self):
        self.graph.free()


def make_propagator(ctx, offsets, neighbors, method, policy, lfht_policy=1, buffer_pow=24, ngpus=1, **kw):
    """main.cu:139-207: method*5+policy selects the propagator; returns
    (proc_name, propagator).  Raises GalpError(EUNSUPPORTED) for the cases the
    one-GPU build leaves out (3, 8 hybrid; 9, 10 multi-GPU; serial CPU)."""
    key = method * 5 + policy
    if key in (3, 8):
        raise _abi.GalpError(-4, "policy 3 (CPU-GPU hybrid) is not part of the one-GPU build")
    if key in (9, 10) or ngpus != 1:
        raise _abi.GalpError(-4, "multi-GPU policies are not part of the one-GPU build")
    if key not in PROC_NAMES:
        raise _abi.GalpError(-4, "the serial CPU path is not part of the MI355X build")
    name = PROC_NAMES[key].format(l=lfht_policy)
    return name, LabelPropagator(ctx, offsets, neighbors, method=method, policy=policy,
                                 lfht_policy=lfht_policy, buffer_pow=buffer_pow, **kw)
