"""grappolo_amd -- an MI355X-native engine for Grappolo's colored multi-phase Louvain.

The engine is native code (``lib/libgrappolo_hip.so``: HIP kernels for gfx950
behind a C ABI, plus Grappolo's own C++ entry points).  This package is the
Python side of that boundary: graph containers, the seeded generators and
ctypes bindings whose names and argument meanings follow the reference's C++
API (``runMultiPhaseColoring``, ``algoLouvainWithDistOneColoring`` ...).
"""
from .graph import Graph, gen_planted, gen_lfr, gen_rmat, read_binary, write_binary, load_metis

__all__ = ["Graph", "gen_planted", "gen_lfr", "gen_rmat", "read_binary", "write_binary",
           "load_metis"]
