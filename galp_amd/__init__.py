"""galp_amd -- MI355X (gfx950) label-propagation update behind galp's interface.

The product is native: libgalp.so (HIP kernels + C-ABI, include/galp.h) and
the `galp` CLI (galp_amd/bin/galp).  This package holds the ctypes bindings
(`_abi`) and a Python mirror of the reference's propagator interface
(`propagator`), used by the tests and the benchmark.
"""
from . import _abi
from ._abi import Context, GalpError, Graph
from .propagator import LabelPropagator, make_propagator

__all__ = ["_abi", "Context", "GalpError", "Graph", "LabelPropagator", "make_propagator"]
