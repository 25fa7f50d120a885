"""MI355X implementation of pypercolate's Newman-Ziff bond-percolation hot
path: a drop-in for ``percolate`` (andsor/pypercolate 0.4.6,
percolate/__init__.py:88-97) and ``percolate.hpc``.

The sweep, the reduction over runs and the canonical convolution run as
hand-written HIP kernels for gfx950 in libpercolate_hip.so; there is no CPU
fallback.
"""
from pypercolate_amd.percolate import (  # noqa: F401
    sample_states,
    single_run_arrays,
    microcanonical_averages,
    microcanonical_averages_arrays,
    canonical_averages,
    spanning_1d_chain,
    spanning_2d_grid,
    statistics,
    percolation_graph,
)
from pypercolate_amd import hpc  # noqa: F401

__version__ = '0.1.0'
