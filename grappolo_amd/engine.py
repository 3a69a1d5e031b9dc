"""ctypes bindings of libgrappolo_hip.so (include/grappolo_hip.h).

Names and argument meanings follow the reference's C++ API
(DefineStructure/color_comm.h, basic_comm.h, basic_util.h, coloring.h); the
reference returns through out-pointers, these return tuples.  Every call runs
on the GPU; a missing library or device raises, there is no CPU fallback.

Device-resident forms (``*_dev``) take a :class:`DeviceGraph` whose arrays are
torch CUDA tensors (torch is used only to hold device memory and streams).
"""
import ctypes

import numpy as np

from ._native import hip_lib
from .graph import EDGE_DTYPE, Graph

_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_F64P = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p

STATUS = {0: "GLV_OK", 1: "GLV_E_INVALID_GRAPH", 2: "GLV_E_OOM", 3: "GLV_E_DEVICE",
          4: "GLV_E_TOO_MANY_COLORS", 5: "GLV_E_UNSUPPORTED", 6: "GLV_E_INVALID_ARG", 7: "GLV_E_IO"}


class GlvError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn}: {STATUS.get(code, code)}: {msg}")
        self.code = code


class Tuning(ctypes.Structure):
    """glv_tuning: launch shapes of the coloring rounds, the sweeps and the coloring core."""
    _fields_ = [("jp_blocks", ctypes.c_int32), ("persist_blocks", ctypes.c_int32),
                ("wide_blocks", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("core_min_deg", ctypes.c_int64)]

    def as_dict(self):
        return {"jp_blocks": self.jp_blocks, "persist_blocks": self.persist_blocks,
                "wide_blocks": self.wide_blocks, "core_min_deg": self.core_min_deg}


class TuneTrial(ctypes.Structure):
    _fields_ = [("t", Tuning), ("stage", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("ms", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("phases", ctypes.c_int32), ("total_iterations", ctypes.c_int32),
                ("final_clusters", ctypes.c_int64), ("final_modularity", ctypes.c_double),
                ("last_modularity", ctypes.c_double), ("time_clustering", ctypes.c_double),
                ("time_building", ctypes.c_double), ("time_coloring", ctypes.c_double),
                ("time_renumber", ctypes.c_double), ("time_total", ctypes.c_double),
                ("edges_visited", ctypes.c_int64), ("edges_visited_mod", ctypes.c_int64),
                ("colors_first", ctypes.c_int32), ("n_phase", ctypes.c_int32),
                ("phase_cap", ctypes.c_int32), ("trace_cap", ctypes.c_int32),
                ("n_trace", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("phase_nv", _I64P), ("phase_nnz", _I64P), ("phase_iters", _I32P),
                ("phase_colors", _I32P), ("phase_colored", _I32P), ("phase_mod", _F64P),
                ("phase_clusters", _I64P), ("trace_q", _F64P), ("phase_seconds", _F64P),
                ("moves", ctypes.c_int64)]


_SIG = {
    "glv_set_device": [ctypes.c_int],
    "glv_get_device_count": [_I32P],
    "glv_runMultiPhaseColoring": [ctypes.c_int64, _I64P, _VP, _I64P, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, ctypes.POINTER(Stats)],
    "glv_runMultiPhaseBasic": [ctypes.c_int64, _I64P, _VP, _I64P, ctypes.c_int, ctypes.c_int64,
                               ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(Stats)],
    "glv_algoDistanceOneVertexColoringOpt": [ctypes.c_int64, _I64P, _VP, _I32P, _I32P, _F64P],
    "glv_algoLouvainWithDistOneColoring": [ctypes.c_int64, _I64P, _VP, _I64P, _I32P, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_double, _F64P, _F64P, _I32P],
    "glv_parallelLouvianMethod": [ctypes.c_int64, _I64P, _VP, _I64P, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, _F64P, _F64P, _I32P],
    "glv_renumberClustersContiguously": [_I64P, ctypes.c_int64, _I64P],
    "glv_buildNextLevelGraphOpt": [ctypes.c_int64, _I64P, _VP, _I64P, ctypes.c_int64,
                                   ctypes.POINTER(_I64P), ctypes.POINTER(_VP), _I64P],
    "glv_buildNewGraphVF": [ctypes.c_int64, _I64P, _VP, _I64P, ctypes.c_int64,
                            ctypes.POINTER(_I64P), ctypes.POINTER(_VP), _I64P],
    "glv_vertexFollowing": [ctypes.c_int64, _I64P, _VP, _I64P, _I64P],
    "glv_modularity": [ctypes.c_int64, _I64P, _VP, _I64P, _F64P],
    "glv_runMultiPhaseColoring_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                      ctypes.c_double, _VP, ctypes.POINTER(Stats)],
    "glv_runMultiPhaseBasic_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP,
                                   ctypes.c_int, ctypes.c_double, _VP, ctypes.POINTER(Stats)],
    "glv_color_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _I32P, _I32P, _VP],
    "glv_colored_phase_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, _VP,
                              ctypes.c_int, ctypes.c_double, ctypes.c_double, _F64P, _I32P, _VP],
    "glv_plain_phase_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, ctypes.c_double,
                            ctypes.c_double, ctypes.c_int, _F64P, _I32P, _VP],
    "glv_color_step_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, _VP, ctypes.c_int,
                           ctypes.c_int, _I64P, _VP],
    "glv_modularity_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, _F64P, _VP],
    "glv_renumber_dev": [ctypes.c_int64, _VP, _I64P, _VP],
    "glv_coarsen_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, ctypes.c_int64,
                        ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP), _I64P,
                        _I64P, _VP],
    "glv_vertexFollowing_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _I64P, _VP],
    "glv_buildNewGraphVF_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, ctypes.c_int64,
                                ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP), _I64P,
                                _I64P, _VP],
    "glv_vertexFollowingPrepass_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, ctypes.POINTER(_VP),
                                       ctypes.POINTER(_VP), ctypes.POINTER(_VP), _I64P, _I64P, _I64P, _I64P,
                                       _VP],
    "glv_computeCommunityComparisons": [_I64P, ctypes.c_int64, _I64P, ctypes.c_int64, _VP],
    "glv_computeCommunityComparisons_dev": [_VP, ctypes.c_int64, _VP, ctypes.c_int64, _VP, _VP],
    "glv_computeGiniCoefficient": [_I64P, ctypes.c_int, _F64P],
    "glv_computeGiniCoefficient_dev": [_VP, ctypes.c_int, _F64P, _VP],
    "glv_algoColoringMultiHashMaxMin": [ctypes.c_int64, _I64P, _VP, _I32P, ctypes.c_int, ctypes.c_int, _I32P],
    "glv_algoColoringMultiHashMaxMin_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, ctypes.c_int,
                                            ctypes.c_int, _I32P, _VP],
    "glv_vBaseRedistribution": [ctypes.c_int64, _I64P, _VP, _I32P, ctypes.c_int, ctypes.c_int, _I32P],
    "glv_vBaseRedistribution_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int,
                                    _I32P, _VP],
    "glv_dev_free": [_VP],
    "glv_count_conflicts_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _I64P, _VP],
    "glv_profile_enable": [ctypes.c_int],
    "glv_profile_reset": [],
    "glv_profile_read": [ctypes.c_int, _F64P, _F64P, _I64P],
    "glv_runMultiPhaseBasic_batch_dev": [ctypes.c_int64, _I64P, ctypes.c_int64, ctypes.c_int64, _VP, _VP,
                                         _VP, _VP, ctypes.c_int, ctypes.c_double, _VP, ctypes.c_void_p],
    "glv_get_tuning": [_VP],
    "glv_set_tuning": [_VP],
    "glv_autotune_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int64,
                         ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP,
                         _VP, ctypes.c_int, _VP, _VP],
    "glv_get_bins": [_I64P],
    "glv_set_bins": [_I64P],
    "glv_autotune_bins_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, ctypes.c_int, _I64P, _F64P,
                              _F64P, _VP],
    "glv_mlp_param_count": [],
    "glv_mlp_forward_dev": [_VP, _VP, ctypes.c_int64, _VP, _VP],
    "glv_mlp_grad_dev": [_VP, _VP, _VP, ctypes.c_int64, ctypes.c_float, _VP, _VP],
    "glv_mlp_train_dev": [_VP, _VP, _VP, _VP, _VP, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                          ctypes.c_float, ctypes.c_float, ctypes.c_uint64, ctypes.c_int64, _VP, _VP],
    "glv_las_trace_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, ctypes.c_int64,
                          ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _VP, _VP,
                          _VP, ctypes.c_int64, _VP, ctypes.c_void_p, ctypes.c_void_p],
    "glv_runMultiPhaseColoring_las_dev": [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP,
                                          ctypes.c_int64, ctypes.c_double, ctypes.c_double, _VP,
                                          ctypes.c_float, ctypes.c_int, _VP, ctypes.c_void_p,
                                          ctypes.c_void_p],
}

KERNEL_CLASSES = ["sweep_bin1_deg4", "sweep_bin2_deg8", "sweep_bin3_deg16", "sweep_bin4_deg32",
                  "sweep_bin5_deg64", "sweep_bin6_wavehash", "sweep_bin7_blockhash",
                  "sweep_bin8_hbmhash", "modularity", "commit", "coloring_round", "coarsen",
                  "renumber", "sweep_persistent_iteration", "coloring_tail", "sweep_bin9_giant",
                  "sweep_wide_iteration"]

_fns = {}


def _f(name):
    if name not in _fns:
        lib = hip_lib()
        f = getattr(lib, name)
        f.argtypes = _SIG[name]
        f.restype = ctypes.c_int
        _fns[name] = f
        lib.glv_last_error.restype = ctypes.c_char_p
        lib.glv_last_error.argtypes = []
    return _fns[name]


def _call(name, *args):
    rc = _f(name)(*args)
    if rc != 0:
        msg = hip_lib().glv_last_error()
        raise GlvError(name, rc, msg.decode() if msg else "")


def _p(a, t):
    return a.ctypes.data_as(t)


def _host_graph(G):
    edges = G.edge_list()
    return [G.numVertices, _p(G.edgeListPtrs, _I64P), edges.ctypes.data_as(_VP)], edges


def new_stats(phase_cap=256, trace_cap=4096):
    st = Stats()
    arrs = dict(phase_nv=np.zeros(phase_cap, np.int64), phase_nnz=np.zeros(phase_cap, np.int64),
                phase_iters=np.zeros(phase_cap, np.int32),
                phase_colors=np.zeros(phase_cap, np.int32),
                phase_colored=np.zeros(phase_cap, np.int32), phase_mod=np.zeros(phase_cap),
                phase_clusters=np.zeros(phase_cap, np.int64), trace_q=np.zeros(trace_cap),
                phase_seconds=np.zeros(phase_cap))
    st.phase_cap = phase_cap
    st.trace_cap = trace_cap
    types = dict(Stats._fields_)
    for k, v in arrs.items():
        setattr(st, k, v.ctypes.data_as(types[k]))
    st._arrays = arrs  # keep alive
    return st


def stats_dict(st):
    d = {name: getattr(st, name) for name, _ in Stats._fields_[:18]}
    d["moves"] = st.moves
    n = min(st.n_phase, st.phase_cap)
    d["per_phase"] = {k: v[:n].copy() for k, v in st._arrays.items() if k != "trace_q"}
    d["trace"] = st._arrays["trace_q"][:min(st.n_trace, st.trace_cap)].copy()
    return d


def profile_enable(on=True):
    _call("glv_profile_enable", int(bool(on)))


def profile_reset():
    _call("glv_profile_reset")


def profile_read():
    """{class name: (ms, algorithmic bytes, launches)} accumulated since reset."""
    out = {}
    for i, name in enumerate(KERNEL_CLASSES):
        ms, by, n = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int64(0)
        _call("glv_profile_read", i, ctypes.byref(ms), ctypes.byref(by), ctypes.byref(n))
        out[name] = (ms.value, by.value, n.value)
    return out


def device_count():
    c = ctypes.c_int32(0)
    _call("glv_get_device_count", ctypes.byref(c))
    return c.value


def set_device(d):
    _call("glv_set_device", int(d))


# ------------------------------------------------------------------ host API
def runMultiPhaseColoring(G, coloring=1, numColors=16, replaceMap=0, minGraphSize=1000,
                          threshold=1e-9, C_threshold=1e-6, numThreads=1, threadsOpt=1):
    """runMultiPhaseColoring (color_comm.h:7).  Returns (C_orig, stats dict)."""
    args, keep = _host_graph(G)
    C_orig = np.full(G.numVertices, -1, dtype=np.int64)
    st = new_stats()
    _call("glv_runMultiPhaseColoring", *args, _p(C_orig, _I64P), coloring, numColors, replaceMap,
          minGraphSize, threshold, C_threshold, numThreads, threadsOpt, ctypes.byref(st))
    return C_orig, stats_dict(st)


def runMultiPhaseBasic(G, basicOpt=0, minGraphSize=100000, threshold=1e-9, C_threshold=0.01,
                       numThreads=1, threadsOpt=1):
    """runMultiPhaseBasic (basic_comm.h:5).  Returns (C_orig, stats dict)."""
    args, keep = _host_graph(G)
    C_orig = np.full(G.numVertices, -1, dtype=np.int64)
    st = new_stats()
    _call("glv_runMultiPhaseBasic", *args, _p(C_orig, _I64P), basicOpt, minGraphSize, threshold,
          C_threshold, numThreads, threadsOpt, ctypes.byref(st))
    return C_orig, stats_dict(st)


def algoDistanceOneVertexColoringOpt(G):
    """coloring.h:8.  Returns (colors int32[NV], largest color)."""
    args, keep = _host_graph(G)
    colors = np.empty(G.numVertices, dtype=np.int32)
    mc = ctypes.c_int32(-1)
    t = ctypes.c_double(0)
    _call("glv_algoDistanceOneVertexColoringOpt", *args, _p(colors, _I32P), ctypes.byref(mc),
          ctypes.byref(t))
    return colors, mc.value


def algoColoringMultiHashMaxMin(G, nHash=2, nItrs=4):
    """coloring.h:12.  Returns (colors, maxColor)."""
    args, keep = _host_graph(G)
    c = np.empty(G.numVertices, np.int32)
    mc = ctypes.c_int32(0)
    _call("glv_algoColoringMultiHashMaxMin", *args, _p(c, _I32P), nHash, nItrs, ctypes.byref(mc))
    return c, mc.value


def vBaseRedistribution(G, colors, ncolors, type=0):
    """coloring.h:15.  Returns (recolored copy, rounds)."""
    args, keep = _host_graph(G)
    c = np.array(colors, dtype=np.int32)
    r = ctypes.c_int32(0)
    _call("glv_vBaseRedistribution", *args, _p(c, _I32P), int(ncolors), int(type), ctypes.byref(r))
    return c, r.value


def multihash_dev(dg, colors, nHash=2, nItrs=4):
    mc = ctypes.c_int32(0)
    _call("glv_algoColoringMultiHashMaxMin_dev", dg.numVertices, dg.nnz, _ptr(dg.off), _ptr(dg.tails),
          _ptr(colors), nHash, nItrs, ctypes.byref(mc), _stream())
    return mc.value


def balance_dev(dg, colors, ncolors, type=0):
    r = ctypes.c_int32(0)
    _call("glv_vBaseRedistribution_dev", dg.numVertices, dg.nnz, _ptr(dg.off), _ptr(dg.tails), _ptr(colors),
          int(ncolors), int(type), ctypes.byref(r), _stream())
    return r.value


def algoLouvainWithDistOneColoring(G, color, numColor, Lower=-1.0, thresh=1e-6):
    """color_comm.h:10.  Returns (modularity, C, numItr)."""
    args, keep = _host_graph(G)
    color = np.ascontiguousarray(color, dtype=np.int32)
    C = np.empty(G.numVertices, dtype=np.int64)
    q, t, it = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int32(0)
    _call("glv_algoLouvainWithDistOneColoring", *args, _p(C, _I64P), _p(color, _I32P),
          int(numColor), Lower, thresh, ctypes.byref(q), ctypes.byref(t), ctypes.byref(it))
    return q.value, C, it.value


def parallelLouvianMethod(G, Lower=-1.0, thresh=1e-6, noMap=False):
    """basic_comm.h:21 (noMap: basic_comm.h:28).  Returns (modularity, C, numItr)."""
    args, keep = _host_graph(G)
    C = np.empty(G.numVertices, dtype=np.int64)
    q, t, it = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int32(0)
    _call("glv_parallelLouvianMethod", *args, _p(C, _I64P), Lower, thresh, int(bool(noMap)),
          ctypes.byref(q), ctypes.byref(t), ctypes.byref(it))
    return q.value, C, it.value


def renumberClustersContiguously(C):
    """basic_util.h:5.  Returns (renumbered copy of C, number of clusters)."""
    C = np.array(C, dtype=np.int64)
    k = ctypes.c_int64(0)
    _call("glv_renumberClustersContiguously", _p(C, _I64P), C.shape[0], ctypes.byref(k))
    return C, k.value


def buildNextLevelGraphOpt(G, C, numUniqueClusters):
    """basic_util.h:6.  Returns the coarsened Graph (rows sorted by tail)."""
    return _host_contract("glv_buildNextLevelGraphOpt", G, C, numUniqueClusters)


def buildNewGraphVF(G, C, numUniqueClusters):
    """basic_util.h:15.  Contracts without zero self-loops, weights truncated to
    integers (vertexFollowing.cpp:93-258).  Rows sorted by tail."""
    return _host_contract("glv_buildNewGraphVF", G, C, numUniqueClusters)


def vertexFollowing(G):
    """basic_util.h:14.  Returns (C, numNode): C[i] = -1 for isolated vertices,
    the neighbour a degree-one vertex follows, else i."""
    args, keep = _host_graph(G)
    C = np.empty(G.numVertices, np.int64)
    n = ctypes.c_int64(0)
    _call("glv_vertexFollowing", *args, _p(C, _I64P), ctypes.byref(n))
    return C, n.value


def _host_contract(fn, G, C, numUniqueClusters):
    args, keep = _host_graph(G)
    C = np.ascontiguousarray(C, dtype=np.int64)
    po, pe, ne = _I64P(), _VP(), ctypes.c_int64(0)
    _call(fn, *args, _p(C, _I64P), int(numUniqueClusters),
          ctypes.byref(po), ctypes.byref(pe), ctypes.byref(ne))
    k = int(numUniqueClusters)
    ptrs = np.ctypeslib.as_array(po, shape=(k + 1,)).copy()
    nnz = int(ptrs[-1])
    buf = (ctypes.c_char * max(nnz * 24, 1)).from_address(pe.value)
    edges = np.frombuffer(buf, dtype=EDGE_DTYPE, count=nnz).copy()
    free = hip_lib().glv_free
    free.argtypes = [_VP]
    free.restype = None
    free(ctypes.cast(po, _VP))
    free(pe)
    return Graph.from_edge_list(ptrs, edges, numEdges=ne.value)


def modularity(G, C):
    """The in-sweep modularity (parallelLouvainWithColoring.cpp:226-249) of labels C."""
    args, keep = _host_graph(G)
    C = np.ascontiguousarray(C, dtype=np.int64)
    q = ctypes.c_double(0)
    _call("glv_modularity", *args, _p(C, _I64P), ctypes.byref(q))
    return q.value


# ---------------------------------------------------------------- device API
class DeviceGraph:
    """A Graph resident in device memory (torch tensors on cuda)."""

    def __init__(self, G, device="cuda"):
        import torch
        self.numVertices = G.numVertices
        self.nnz = G.nnz
        self.numEdges = G.numEdges
        self.off = torch.from_numpy(G.edgeListPtrs).to(device)
        self.tails = torch.from_numpy(G.tails).to(device)
        self.w = torch.from_numpy(G.weights).to(device)

    @classmethod
    def from_tensors(cls, off, tails, w, numEdges):
        """Wraps CSR tensors already on the device."""
        dg = cls.__new__(cls)
        dg.numVertices = off.shape[0] - 1
        dg.nnz = tails.shape[0]
        dg.numEdges = numEdges
        dg.off, dg.tails, dg.w = off, tails, w
        return dg

    def ptrs(self):
        return [self.numVertices, self.nnz, _VP(self.off.data_ptr()), _VP(self.tails.data_ptr()),
                _VP(self.w.data_ptr())]

    def to_host(self):
        return Graph(self.off.cpu().numpy(), self.tails.cpu().numpy(), self.w.cpu().numpy(),
                     numEdges=self.numEdges)


def _stream():
    import torch
    return _VP(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return _VP(t.data_ptr())


def run_colored_dev(dg, C_orig, coloring=1, replaceMap=0, minGraphSize=1000, threshold=1e-9,
                    C_threshold=1e-6, stats=None, numColors=16):
    st = stats if stats is not None else new_stats()
    _call("glv_runMultiPhaseColoring_dev", *dg.ptrs(), _ptr(C_orig), coloring, numColors, replaceMap,
          minGraphSize, threshold, C_threshold, _stream(), ctypes.byref(st))
    return stats_dict(st)


def run_basic_dev(dg, C_orig, basicOpt=0, threshold=1e-9):
    st = new_stats()
    _call("glv_runMultiPhaseBasic_dev", *dg.ptrs(), _ptr(C_orig), basicOpt, threshold,
          _stream(), ctypes.byref(st))
    return stats_dict(st)


def color_dev(dg, colors):
    mc, rounds = ctypes.c_int32(0), ctypes.c_int32(0)
    _call("glv_color_dev", dg.numVertices, dg.nnz, _ptr(dg.off), _ptr(dg.tails), _ptr(colors),
          ctypes.byref(mc), ctypes.byref(rounds), _stream())
    return mc.value, rounds.value


def count_conflicts_dev(dg, colors):
    c = ctypes.c_int64(0)
    _call("glv_count_conflicts_dev", dg.numVertices, dg.nnz, _ptr(dg.off), _ptr(dg.tails),
          _ptr(colors), ctypes.byref(c), _stream())
    return c.value


def colored_phase_dev(dg, C, colors, numColor, Lower=-1.0, thresh=1e-6):
    q, it = ctypes.c_double(0), ctypes.c_int32(0)
    _call("glv_colored_phase_dev", *dg.ptrs(), _ptr(C), _ptr(colors), int(numColor), Lower,
          thresh, ctypes.byref(q), ctypes.byref(it), _stream())
    return q.value, it.value


def plain_phase_dev(dg, C, Lower=-1.0, thresh=1e-6, noMap=False):
    q, it = ctypes.c_double(0), ctypes.c_int32(0)
    _call("glv_plain_phase_dev", *dg.ptrs(), _ptr(C), Lower, thresh, int(bool(noMap)),
          ctypes.byref(q), ctypes.byref(it), _stream())
    return q.value, it.value


def color_step_dev(dg, C, colors, numColor, which):
    m = ctypes.c_int64(0)
    _call("glv_color_step_dev", *dg.ptrs(), _ptr(C), _ptr(colors), int(numColor), int(which),
          ctypes.byref(m), _stream())
    return m.value


def modularity_dev(dg, C):
    q = ctypes.c_double(0)
    _call("glv_modularity_dev", *dg.ptrs(), _ptr(C), ctypes.byref(q), _stream())
    return q.value


def renumber_dev(C):
    k = ctypes.c_int64(0)
    _call("glv_renumber_dev", C.shape[0], _ptr(C), ctypes.byref(k), _stream())
    return k.value


def coarsen_dev(dg, C, k, vf=False):
    """Returns a host Graph copied from the engine's coarsened device graph
    (vf: the buildNewGraphVF contraction)."""
    import torch
    po, pt, pw = _VP(), _VP(), _VP()
    nnz, ne = ctypes.c_int64(0), ctypes.c_int64(0)
    _call("glv_buildNewGraphVF_dev" if vf else "glv_coarsen_dev", *dg.ptrs(), _ptr(C), int(k),
          ctypes.byref(po), ctypes.byref(pt), ctypes.byref(pw), ctypes.byref(nnz), ctypes.byref(ne),
          _stream())
    torch.cuda.synchronize()
    off = np.empty(int(k) + 1, np.int64)
    tails = np.empty(max(nnz.value, 1), np.int64)
    w = np.empty(max(nnz.value, 1))
    hip_copy_d2h(off.ctypes.data, po.value, off.nbytes)
    if nnz.value:
        hip_copy_d2h(tails.ctypes.data, pt.value, nnz.value * 8)
        hip_copy_d2h(w.ctypes.data, pw.value, nnz.value * 8)
    for p in (po, pt, pw):
        _call("glv_dev_free", p)
    return Graph(off, tails[:nnz.value], w[:nnz.value], numEdges=ne.value)


def vertex_following_prepass_dev(dg):
    """The driver's vertex-following hook on the device (vertexFollowing,
    renumber, buildNewGraphVF).  Returns (graph, numNode); the graph is `dg`
    itself when nothing was folded."""
    import torch
    po, pt, pw = _VP(), _VP(), _VP()
    nv, nnz, ne, n = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    _call("glv_vertexFollowingPrepass_dev", *dg.ptrs(), ctypes.byref(po), ctypes.byref(pt), ctypes.byref(pw),
          ctypes.byref(nv), ctypes.byref(nnz), ctypes.byref(ne), ctypes.byref(n), _stream())
    if n.value == 0:
        return dg, 0
    torch.cuda.synchronize()
    dev = dg.off.device
    off = torch.empty(nv.value + 1, dtype=torch.int64, device=dev)
    tails = torch.empty(nnz.value, dtype=torch.int64, device=dev)
    w = torch.empty(nnz.value, dtype=torch.float64, device=dev)
    for t, p in ((off, po), (tails, pt), (w, pw)):
        hip_copy(t.data_ptr(), p.value, t.numel() * 8, 3)  # hipMemcpyDeviceToDevice
    for p in (po, pt, pw):
        if p.value:
            _call("glv_dev_free", p)
    return DeviceGraph.from_tensors(off, tails, w, ne.value), n.value


def vertex_following_dev(dg, C):
    """Fills device labels C (vertexFollowing); returns numNode."""
    n = ctypes.c_int64(0)
    nv, nnz, off, tails, _w = dg.ptrs()
    _call("glv_vertexFollowing_dev", nv, nnz, off, tails, _ptr(C), ctypes.byref(n), _stream())
    return n.value


class ClusterComparison(ctypes.Structure):
    """glv_cluster_comparison"""
    _fields_ = [("nC1", ctypes.c_int64), ("nC2", ctypes.c_int64), ("SameSame", ctypes.c_int64),
                ("SameDiff", ctypes.c_int64), ("DiffSame", ctypes.c_int64), ("Agree", ctypes.c_int64),
                ("precision", ctypes.c_double), ("recall", ctypes.c_double), ("fScore", ctypes.c_double),
                ("gini1", ctypes.c_double), ("gini2", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def computeCommunityComparisons(C1, C2):
    """utilityClusteringFunctions.h:71 (C1 = truth).  Returns (metrics dict,
    relabelled C1, relabelled C2) -- the reference relabels -1 in place."""
    a = np.array(C1, dtype=np.int64)
    b = np.array(C2, dtype=np.int64)
    out = ClusterComparison()
    _call("glv_computeCommunityComparisons", _p(a, _I64P), a.shape[0], _p(b, _I64P), b.shape[0],
          ctypes.byref(out))
    return out.as_dict(), a, b


def compare_dev(C1, C2):
    """computeCommunityComparisons on device label tensors (relabelled in place)."""
    out = ClusterComparison()
    _call("glv_computeCommunityComparisons_dev", _ptr(C1), C1.shape[0], _ptr(C2), C2.shape[0],
          ctypes.byref(out), _stream())
    return out.as_dict()


def computeGiniCoefficient(sizes):
    """basic_util.h:18.  Returns (gini, sizes sorted as the reference leaves them)."""
    s = np.array(sizes, dtype=np.int64)
    g = ctypes.c_double(0)
    _call("glv_computeGiniCoefficient", _p(s, _I64P), s.shape[0], ctypes.byref(g))
    return g.value, s


def gini_dev(sizes):
    """computeGiniCoefficient of a device int64 tensor (sorted in place)."""
    g = ctypes.c_double(0)
    _call("glv_computeGiniCoefficient_dev", _ptr(sizes), sizes.shape[0], ctypes.byref(g), _stream())
    return g.value


class BatchStats(ctypes.Structure):
    _fields_ = [("final_modularity", _F64P), ("last_modularity", _F64P), ("phases", _I32P),
                ("total_iterations", _I32P), ("final_clusters", _I64P), ("levels", ctypes.c_int32),
                ("iterations", ctypes.c_int32), ("time_total", ctypes.c_double),
                ("launches", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def run_basic_batch_dev(dg, vbase, C_orig, basicOpt=0, threshold=1e-9):
    """runMultiPhaseBasic on every graph of a block-diagonal union at once
    (``graph.union``); C_orig receives graph-local labels.  Returns per-graph
    arrays and the batch's level and iteration counts."""
    vb = np.ascontiguousarray(vbase, dtype=np.int64)
    ng = len(vb) - 1
    out = dict(final_modularity=np.zeros(ng), last_modularity=np.zeros(ng),
               phases=np.zeros(ng, np.int32), total_iterations=np.zeros(ng, np.int32),
               final_clusters=np.zeros(ng, np.int64))
    st = BatchStats(_p(out["final_modularity"], _F64P), _p(out["last_modularity"], _F64P),
                    _p(out["phases"], _I32P), _p(out["total_iterations"], _I32P),
                    _p(out["final_clusters"], _I64P), 0, 0, 0.0)
    _call("glv_runMultiPhaseBasic_batch_dev", ng, _p(vb, _I64P), *dg.ptrs(), _ptr(C_orig),
          int(basicOpt), float(threshold), _stream(), ctypes.byref(st))
    out.update(levels=st.levels, iterations=st.iterations, time_total=st.time_total, launches=st.launches)
    return out


def get_bins():
    """Upper degree of each sweep bin (10 values) used by later calls."""
    b = np.zeros(10, np.int64)
    _call("glv_get_bins", _p(b, _I64P))
    return b


def set_bins(max_deg):
    b = np.ascontiguousarray(max_deg, dtype=np.int64)
    if b.shape != (10,):
        raise ValueError("10 bin bounds expected")
    _call("glv_set_bins", _p(b, _I64P))


def autotune_bins_dev(dg, reps=3):
    """(bins, ms with the default bins, ms with the returned bins) of one plain sweep of dg."""
    b = np.zeros(10, np.int64)
    t0, t1 = ctypes.c_double(0), ctypes.c_double(0)
    _call("glv_autotune_bins_dev", *dg.ptrs(), int(reps), _p(b, _I64P), ctypes.byref(t0),
          ctypes.byref(t1), _stream())
    return b, t0.value, t1.value


def get_tuning():
    """The launch shapes later calls use (dict)."""
    t = Tuning()
    _call("glv_get_tuning", ctypes.byref(t))
    return t.as_dict()


def set_tuning(jp_blocks=64, persist_blocks=64, wide_blocks=0, core_min_deg=1024):
    t = Tuning(int(jp_blocks), int(persist_blocks), int(wide_blocks), 0, int(core_min_deg))
    _call("glv_set_tuning", ctypes.byref(t))


def autotune_dev(dg, runs=True, apply=True, reps=1, coloring=1, minGraphSize=1000, threshold=1e-9,
                 C_threshold=1e-6, cap=64):
    """Times launch-shape candidates on dg (glv_autotune_dev): returns
    (best shapes, [(shapes, stage, ms), ...]); apply makes them current."""
    best = Tuning()
    trials = (TuneTrial * cap)()
    n = ctypes.c_int(0)
    _call("glv_autotune_dev", *dg.ptrs(), int(coloring), int(minGraphSize), float(threshold),
          float(C_threshold), int(reps), int(bool(runs)), int(bool(apply)), ctypes.byref(best), trials,
          int(cap), ctypes.byref(n), _stream())
    log = [(trials[i].t.as_dict(), ("coloring", "run")[trials[i].stage], trials[i].ms)
           for i in range(min(n.value, cap))]
    return best.as_dict(), log


_hip_rt = None


def hip_copy_d2h(dst, src, nbytes):
    """Blocking device-to-host copy of engine-owned memory (HIP runtime)."""
    hip_copy(dst, src, nbytes, 2)  # hipMemcpyDeviceToHost


def hip_copy(dst, src, nbytes, kind):
    """Blocking hipMemcpy of `kind` (2: device to host, 3: device to device)."""
    global _hip_rt
    if nbytes == 0:
        return
    if _hip_rt is None:
        _hip_rt = ctypes.CDLL("libamdhip64.so")
        _hip_rt.hipMemcpy.argtypes = [_VP, _VP, ctypes.c_size_t, ctypes.c_int]
        _hip_rt.hipMemcpy.restype = ctypes.c_int
    rc = _hip_rt.hipMemcpy(_VP(dst), _VP(src), nbytes, kind)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
