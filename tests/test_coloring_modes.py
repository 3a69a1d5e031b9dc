"""coloring = 2 (balanced, vBase.cpp:6-168) and coloring = 3 (incomplete
multi-hash, coloringMultiHashMaxMin.cpp:60-226) on the device against the
oracle, through the C ABI: colors bit for bit, then whole runMultiPhaseColoring
runs with those colorings (runMultiPhaseColoring.cpp:68-87, :183-197)."""
import numpy as np
import pytest

import oracle
from grappolo_amd import engine

from graphs import catalog, pendants, unit_weights

pytestmark = pytest.mark.gpu

GRAPHS = dict(catalog(), pendants=pendants(seed=4))
NAMES = list(GRAPHS)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("nItrs", [1, 2, 4])
def test_multihash_equals_oracle(name, nItrs):
    import torch
    g = GRAPHS[name]
    exp, emc = oracle.color_multihash(g, 2, nItrs)
    got, mc = engine.algoColoringMultiHashMaxMin(g, 2, nItrs)
    assert mc == emc
    np.testing.assert_array_equal(got, exp)
    dg = engine.DeviceGraph(g)
    c = torch.empty(max(g.numVertices, 1), dtype=torch.int32, device="cuda")[:g.numVertices]
    assert engine.multihash_dev(dg, c, 2, nItrs) == emc
    np.testing.assert_array_equal(c.cpu().numpy(), exp)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("type", [0, 1])
def test_balance_equals_oracle(name, type):
    import torch
    g = GRAPHS[name]
    for base, mc in (oracle.color_greedy(g), oracle.color_jp(g)[:2]):
        exp, rounds = oracle.color_balance(g, base, mc + 1, type)
        got, grounds = engine.vBaseRedistribution(g, base, mc + 1, type)
        assert grounds == rounds
        np.testing.assert_array_equal(got, exp)
        dg = engine.DeviceGraph(g)
        c = torch.from_numpy(base.copy()).cuda()
        assert engine.balance_dev(dg, c, mc + 1, type) == rounds
        np.testing.assert_array_equal(c.cpu().numpy(), exp)


@pytest.mark.parametrize("coloring", ["balanced", "multihash"])
@pytest.mark.parametrize("name", ["karate", "planted_unit", "planted_w", "lfr_w", "rmat12", "hubs_big"])
def test_pipeline_other_colorings(name, coloring):
    import torch
    g = GRAPHS[name]
    mgs = 100 if g.numVertices > 100 else 10
    mode = {"balanced": 2, "multihash": 3}[coloring]
    dg = engine.DeviceGraph(g)
    C = torch.empty(g.numVertices, dtype=torch.int64, device="cuda")
    st = engine.run_colored_dev(dg, C, coloring=mode, numColors=12, minGraphSize=mgs)
    Co, sto = oracle.run_colored(g, coloring=coloring, frozen=True, minGraphSize=mgs, numColors=12)
    if unit_weights(g):
        np.testing.assert_array_equal(C.cpu().numpy(), Co)
        assert st["final_modularity"] == sto["final_modularity"]
        assert st["phases"] == sto["phases"]
        assert st["total_iterations"] == sto["total_iterations"]
    else:
        assert abs(st["final_modularity"] - sto["final_modularity"]) < 1e-3
    assert st["colors_first"] == sto["colors_first"]
