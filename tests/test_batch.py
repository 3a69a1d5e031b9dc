"""Batched runMultiPhaseBasic: every graph of a block-diagonal union follows
exactly the run it would have alone (labels, modularity, phases, iterations
on integer weights), through the C ABI."""
import numpy as np
import pytest

from grappolo_amd.graph import Graph, gen_planted, gen_planted_many, union

from graphs import hubs, karate, with_loops_and_isolated

pytestmark = pytest.mark.gpu


def _batch_graphs():
    gs = [karate()]
    rng = np.random.default_rng(4)
    for i in range(14):
        nv = int(rng.integers(40, 3000))
        gs.append(gen_planted(nv, int(rng.integers(2, 30)), float(rng.uniform(4, 20)),
                              float(rng.uniform(0.1, 0.5)), seed=100 + i))
    gs.append(Graph(np.zeros(6, np.int64), np.zeros(0, np.int64), np.zeros(0)))  # edgeless
    gs.append(hubs(seed=3, nv=9000, hub_degrees=(2500, 400)))  # HBM-hash bin
    gs.append(hubs(seed=5, nv=30000, hub_degrees=(17000, 900)))  # giant bin
    gs.append(gen_planted(5, 1, 2, 0.0, seed=9))
    return gs


def _standalone(g, basic_opt):
    import torch
    from grappolo_amd import engine
    dg = engine.DeviceGraph(g, "cuda")
    C = torch.empty(max(g.numVertices, 1), dtype=torch.int64, device="cuda")[:g.numVertices]
    st = engine.run_basic_dev(dg, C, basicOpt=basic_opt)
    return C.cpu().numpy(), st


@pytest.mark.parametrize("basic_opt", [0, 1])
def test_batch_equals_separate_runs(basic_opt):
    import torch
    from grappolo_amd import engine
    gs = _batch_graphs()
    U, vb = union(gs)
    dg = engine.DeviceGraph(U, "cuda")
    C = torch.empty(U.numVertices, dtype=torch.int64, device="cuda")
    res = engine.run_basic_batch_dev(dg, vb, C, basicOpt=basic_opt)
    Cb = C.cpu().numpy()
    assert res["levels"] >= 2
    # each level's iterations run as captured runs of 4 (one launch each;
    # finished graphs leave the sweep lists between launches)
    assert res["iterations"] >= max(res["total_iterations"])
    assert 1 <= res["launches"] <= res["iterations"]
    for i, g in enumerate(gs):
        Ci, st = _standalone(g, basic_opt)
        np.testing.assert_array_equal(Cb[vb[i]:vb[i + 1]], Ci, err_msg=f"graph {i}")
        assert res["final_modularity"][i] == st["final_modularity"], i
        assert res["last_modularity"][i] == st["last_modularity"], i
        assert res["phases"][i] == st["phases"], i
        assert res["total_iterations"][i] == st["total_iterations"], i
        assert res["final_clusters"][i] == st["final_clusters"], i


def test_batch_weighted_within_tolerance():
    import torch
    from grappolo_amd import engine
    gs = [with_loops_and_isolated(seed=s, nv=1500) for s in (1, 2, 3)]
    gs += [gen_planted(2000, 8, 12, 0.3, seed=s, wlo=0.5, whi=2.0) for s in (7, 8)]
    U, vb = union(gs)
    dg = engine.DeviceGraph(U, "cuda")
    C = torch.empty(U.numVertices, dtype=torch.int64, device="cuda")
    res = engine.run_basic_batch_dev(dg, vb, C)
    for i, g in enumerate(gs):
        _, st = _standalone(g, 0)
        assert abs(res["final_modularity"][i] - st["final_modularity"]) < 1e-3
        q = engine.modularity(g, C[vb[i]:vb[i + 1]].cpu().numpy())
        assert abs(q - res["final_modularity"][i]) < 1e-3


def test_batch_rejects_cross_graph_entries():
    import torch
    from grappolo_amd import engine
    g = gen_planted(200, 4, 6, 0.2, seed=1)
    U, vb = union([g, g])
    vb_bad = np.array([0, 150, 400], np.int64)  # splits the first graph
    dg = engine.DeviceGraph(U, "cuda")
    C = torch.empty(U.numVertices, dtype=torch.int64, device="cuda")
    with pytest.raises(engine.GlvError):
        engine.run_basic_batch_dev(dg, vb_bad, C)
    with pytest.raises(engine.GlvError):
        engine.run_basic_batch_dev(dg, np.array([0, 200, 300], np.int64), C)


def test_autotuned_bins_keep_results():
    """Bin ranges change only which kernel serves a vertex: with the tuned
    ranges a unit-weight run is identical to the default one."""
    import torch
    from grappolo_amd import engine
    gs = [gen_planted(3000, 12, float(d), 0.3, seed=int(d)) for d in (6, 14, 30)]
    U, vb = union(gs)
    dg = engine.DeviceGraph(U, "cuda")
    default = engine.get_bins()
    tuned, t_def, t_tuned = engine.autotune_bins_dev(dg, reps=2)
    assert t_tuned <= t_def * 1.0 + 1e-9
    assert tuned[0] == 0 and np.all(np.diff(tuned) >= 0) and np.all(tuned <= default)
    assert np.all(tuned[7:] == default[7:])
    C0 = torch.empty(U.numVertices, dtype=torch.int64, device="cuda")
    C1 = torch.empty_like(C0)
    r0 = engine.run_basic_batch_dev(dg, vb, C0)
    st0 = engine.run_colored_dev(dg, torch.empty_like(C0))
    try:
        # every bin but the hash ones emptied into the next: the widest groups serve all degrees <= 64
        engine.set_bins(np.array([0, 0, 0, 0, 0, 64, 256, 2048, 16384, default[9]], np.int64))
        r1 = engine.run_basic_batch_dev(dg, vb, C1)
        st1 = engine.run_colored_dev(dg, torch.empty_like(C0))
        engine.set_bins(tuned)
        r2 = engine.run_basic_batch_dev(dg, vb, torch.empty_like(C0))
    finally:
        engine.set_bins(default)
    assert torch.equal(C0, C1)
    np.testing.assert_array_equal(r0["final_modularity"], r1["final_modularity"])
    np.testing.assert_array_equal(r0["final_modularity"], r2["final_modularity"])
    assert st0["final_modularity"] == st1["final_modularity"]
    with pytest.raises(engine.GlvError):
        engine.set_bins(np.array([0, 8, 8, 16, 32, 64, 256, 2048, 16384, default[9]], np.int64))


def test_batched_4096_graph_workload():
    """The batched workload itself: the 4096 planted-partition graphs of
    bench.batch_specs() (vertices log-uniform on [100, 50000], blocks
    log-uniform on [4, 256], mean degree U[8, 32], mixing U[0.1, 0.5], unit
    weights, seeds 1000 + g) in one batched runMultiPhaseBasic.  Every 128th
    graph is run alone and must match bit for bit (labels, modularity,
    phases, iterations); the smallest graphs of that sample are checked
    against the oracle's runMultiPhaseBasic (BasicCommunitiesDetection/
    runMultiPhaseBasic.cpp:52-170)."""
    import time

    import torch

    import bench
    import oracle
    from grappolo_amd import engine
    t = time.perf_counter()
    specs = bench.batch_specs(4096)
    graphs = gen_planted_many(specs)
    U, vb = union(graphs)
    print(f"batch workload: {U} in {time.perf_counter() - t:.1f}s", flush=True)
    dg = engine.DeviceGraph(U, "cuda")
    C = torch.empty(U.numVertices, dtype=torch.int64, device="cuda")
    res = engine.run_basic_batch_dev(dg, vb, C)
    Cb = C.cpu().numpy()
    del dg
    sample = list(range(0, len(graphs), 128))
    for i in sample:
        Ci, st = _standalone(graphs[i], 0)
        np.testing.assert_array_equal(Cb[vb[i]:vb[i + 1]], Ci, err_msg=f"graph {i}")
        assert res["final_modularity"][i] == st["final_modularity"], i
        assert res["phases"][i] == st["phases"], i
        assert res["total_iterations"][i] == st["total_iterations"], i
    small = sorted(sample, key=lambda i: graphs[i].numVertices)[:6]
    for i in small:
        Co, so = oracle.run_basic(graphs[i], basicOpt=0)
        np.testing.assert_array_equal(Cb[vb[i]:vb[i + 1]], Co, err_msg=f"graph {i} vs oracle")
        assert res["final_modularity"][i] == so["final_modularity"], i
