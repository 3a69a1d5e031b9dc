"""glv_modularity_noalloc_dev (no allocation of any kind: its scratch is a
fixed device array per GPU) against glv_modularity_dev, the entry point it is
a variant of, through the C ABI: two calls in a row, the second on a smaller
graph (the array keeps the first call's contents), then a graph beyond the
array's capacity, which is rejected before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_VP = ctypes.c_void_p


def _noalloc(dg, C):
    import torch
    from grappolo_amd import engine
    lib = engine.hip_lib()
    f = lib.glv_modularity_noalloc_dev
    f.argtypes = [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, ctypes.POINTER(ctypes.c_double), _VP]
    f.restype = ctypes.c_int
    q = ctypes.c_double(0)
    rc = f(*dg.ptrs(), _VP(C.data_ptr()), ctypes.byref(q), _VP(torch.cuda.current_stream().cuda_stream))
    return rc, q.value


def _labels(g):
    import torch
    from grappolo_amd import engine
    dg = engine.DeviceGraph(g, "cuda")
    C = torch.empty(g.numVertices, dtype=torch.int64, device="cuda")
    engine.run_colored_dev(dg, C)
    return dg, C


def test_noalloc_modularity_equals_modularity_dev_on_two_calls():
    from grappolo_amd import engine
    from grappolo_amd.graph import gen_lfr, gen_planted
    # first: FP64 weights U[0.5, 2] (sums in another order: 1e-12 relative)
    dg1, C1 = _labels(gen_lfr(300000, seed=11))
    print("noalloc call 1: LFR 300000 vertices", flush=True)
    rc, q1 = _noalloc(dg1, C1)
    assert rc == 0, engine.hip_lib().glv_last_error()
    ref1 = engine.modularity_dev(dg1, C1)
    assert abs(q1 - ref1) <= 1e-12 * abs(ref1), (q1, ref1)
    # second, smaller, unit weights: exact; the scratch still holds call 1's data
    dg2, C2 = _labels(gen_planted(20000, 40, 12, 0.3, seed=12))
    print("noalloc call 2: planted 20000 vertices", flush=True)
    rc, q2 = _noalloc(dg2, C2)
    assert rc == 0, engine.hip_lib().glv_last_error()
    assert q2 == engine.modularity_dev(dg2, C2)
    # singletons: the closed form sum_i l_i / 2m - sum_i k_i^2 / (2m)^2 (no self-loops: l_i = 0)
    import torch
    g3 = gen_planted(5000, 10, 8, 0.2, seed=13)
    dg3 = engine.DeviceGraph(g3, "cuda")
    S = torch.arange(g3.numVertices, dtype=torch.int64, device="cuda")
    rc, q3 = _noalloc(dg3, S)
    assert rc == 0
    k = np.diff(g3.edgeListPtrs).astype(np.float64)
    assert q3 == pytest.approx(-np.sum(k * k) / k.sum() ** 2, rel=1e-12)


def test_noalloc_rejects_graphs_beyond_its_scratch():
    from grappolo_amd import engine
    lib = engine.hip_lib()
    m = ctypes.c_int64(0)
    assert lib.glv_noalloc_max_vertices(ctypes.byref(m)) == 0
    assert m.value == 4194304
    f = lib.glv_modularity_noalloc_dev
    f.argtypes = [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, ctypes.POINTER(ctypes.c_double), _VP]
    q = ctypes.c_double(0)
    # rejected on the sizes alone: no pointer is read
    rc = f(m.value + 1, 2 * (m.value + 1), _VP(8), _VP(8), _VP(8), _VP(8), ctypes.byref(q), None)
    assert engine.STATUS.get(rc) == "GLV_E_INVALID_ARG"
    lib.glv_last_error.restype = ctypes.c_char_p
    assert b"no-allocation scratch" in lib.glv_last_error()
