"""Launch shapes and their tuner (glv_get_tuning / glv_set_tuning /
glv_autotune_dev; no counterpart in the reference): the grids of the
persistent coloring rounds, the persistent sweep and the whole-GPU sweep,
and the degree where the coloring's core starts.  None may change a result,
so every shape is checked against the oracle's golden run of RMAT scale 19
(the whole-GPU sweep, coarse persistent phases and a hub core in the
coloring) and the oracle's coloring."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wide_rmat19.json")
DEFAULT = {"jp_blocks": 64, "persist_blocks": 64, "wide_blocks": 0, "core_min_deg": 1024}


def _sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_tuning_limits_and_round_trip():
    """Shapes outside the kernels' limits are refused; valid ones read back."""
    from grappolo_amd import engine
    assert engine.get_tuning() == DEFAULT
    try:
        for bad in ({"jp_blocks": 0}, {"jp_blocks": 257}, {"persist_blocks": 257}, {"persist_blocks": 0},
                    {"wide_blocks": -1}, {"core_min_deg": 0}):
            with pytest.raises(engine.GlvError):
                engine.set_tuning(**{**DEFAULT, **bad})
            assert engine.get_tuning() == DEFAULT
        engine.set_tuning(jp_blocks=256, persist_blocks=16, wide_blocks=128, core_min_deg=256)
        assert engine.get_tuning() == {"jp_blocks": 256, "persist_blocks": 16, "wide_blocks": 128,
                                       "core_min_deg": 256}
        # the dense persistent sweep takes up to one workgroup per CU
        engine.set_tuning(**{**DEFAULT, "persist_blocks": 256})
        assert engine.get_tuning()["persist_blocks"] == 256
    finally:
        engine.set_tuning(**DEFAULT)


@pytest.fixture(scope="module")
def rmat19():
    from grappolo_amd import engine
    from grappolo_amd.graph import gen_rmat
    with open(GOLDEN) as f:
        gold = json.load(f)
    g = gen_rmat(19, seed=3)
    assert _sha(g.edgeListPtrs, g.tails, g.weights) == gold["graph_sha256"]
    return g, engine.DeviceGraph(g, "cuda"), gold


@pytest.mark.gpu
@pytest.mark.parametrize("shapes", [
    {"jp_blocks": 256, "persist_blocks": 16, "wide_blocks": 128, "core_min_deg": 256},
    {"jp_blocks": 16, "persist_blocks": 2, "wide_blocks": 64, "core_min_deg": 1 << 20},
])
def test_shapes_keep_the_results(rmat19, shapes):
    """Other grids and core degrees: the coloring equals the oracle's and the
    run's labels, modularity and iterations equal the oracle's golden run."""
    import torch
    from grappolo_amd import engine
    g, dg, gold = rmat19
    engine.set_tuning(**shapes)
    try:
        colors, mc = engine.algoDistanceOneVertexColoringOpt(g)
        ref, rmc, _ = oracle.color_jp(g)
        assert mc == rmc
        np.testing.assert_array_equal(colors, ref)
        C = torch.empty(g.numVertices, dtype=torch.int64, device="cuda")
        st = engine.run_colored_dev(dg, C, minGraphSize=1000, threshold=1e-9, C_threshold=1e-6)
        assert _sha(C.cpu().numpy()) == gold["labels_sha256"]
        assert st["final_modularity"] == gold["final_modularity"]
        assert st["per_phase"]["phase_iters"].tolist() == gold["phase_iters"]
    finally:
        engine.set_tuning(**DEFAULT)


@pytest.mark.gpu
def test_autotune_picks_measured_shapes(rmat19):
    """The tuner times the coloring candidates and whole runs, returns the
    fastest it measured, applies it on request, and the run with it keeps the
    golden result."""
    import torch
    from grappolo_amd import engine
    g, dg, gold = rmat19
    try:
        best, log = engine.autotune_dev(dg, runs=True, apply=True, reps=1)
        stages = [s for _, s, _ in log]
        assert stages.count("coloring") >= 3 and stages.count("run") >= 3
        assert all(ms > 0 for _, _, ms in log)
        assert best in [t for t, _, _ in log]
        assert engine.get_tuning() == best
        # the chosen shapes were the fastest of their stage within 2 %
        run_ms = {json.dumps(t, sort_keys=True): ms for t, s, ms in log if s == "run"}
        if run_ms:
            assert run_ms[json.dumps(best, sort_keys=True)] <= min(run_ms.values()) / 0.98 + 1e-9
        C = torch.empty(g.numVertices, dtype=torch.int64, device="cuda")
        st = engine.run_colored_dev(dg, C, minGraphSize=1000, threshold=1e-9, C_threshold=1e-6)
        assert _sha(C.cpu().numpy()) == gold["labels_sha256"]
        assert st["final_modularity"] == gold["final_modularity"]
    finally:
        engine.set_tuning(**DEFAULT)
