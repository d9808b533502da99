"""galp_amd -- MI355X (gfx950) label-propagation update behind galp's interface.

The product is native: libgalp.so (HIP kernels + C-ABI, include/galp.h), the
`galp` CLI (galp_amd/bin/galp) and the graph converter
(galp_amd/bin/galp_convert).  This package holds the ctypes bindings (`_abi`)
the tests and the benchmark call the C-ABI through.
"""
from . import _abi
from ._abi import Context, GalpError, Graph

__all__ = ["_abi", "Context", "GalpError", "Graph"]
