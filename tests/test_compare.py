"""Cluster-comparison metrics on the device (utilityClusterComparisonMetrics
.cpp:50-330) against the oracle, through the C ABI: pair counts, community
counts and the relabelled vectors bit for bit; precision, recall, F-score and
the Gini coefficients equal as doubles (the same expressions on the same
integers)."""
import numpy as np
import pytest

import oracle
from grappolo_amd import driver, engine
from grappolo_amd.graph import gen_planted

pytestmark = pytest.mark.gpu


def _same(got, exp):
    keys = {"nC1": "nC1", "nC2": "nC2", "SameSame": "same_same", "SameDiff": "same_diff",
            "DiffSame": "diff_same", "Agree": "agree", "precision": "precision", "recall": "recall",
            "fScore": "f_score", "gini1": "gini1", "gini2": "gini2"}
    for g, e in keys.items():
        a, b = got[g], exp[e]
        if isinstance(b, float) and np.isnan(b):
            assert np.isnan(a), g
        else:
            assert a == b, (g, a, b)


def _cases():
    rng = np.random.default_rng(5)
    out = []
    for case in range(10):
        n = int(rng.integers(1, 50000))
        C1 = rng.integers(case % 2, int(rng.integers(2, 3000)), n)
        C2 = rng.integers((case // 2) % 2, int(rng.integers(2, 5000)), n)
        if case >= 4:
            C1[rng.random(n) < 0.05] = -1
        if case >= 6:
            C2[rng.random(n) < 0.05] = -1
        if (C1 == -1).all():
            C1[0] = 1
        out.append((C1, C2))
    out.append((np.array([0]), np.array([0])))        # one vertex: no pairs, NaN ratios
    out.append((np.array([3, 3, 3]), np.array([-1, -1, -1])))
    out.append((np.zeros(1000, np.int64), np.arange(1000)))  # one truth community, all split
    return out


@pytest.mark.parametrize("case", range(13))
def test_compare_equals_oracle(case):
    import torch
    C1, C2 = _cases()[case]
    exp, ea, eb = oracle.compare(C1, C2)
    got, a, b = engine.computeCommunityComparisons(C1, C2)
    _same(got, exp)
    np.testing.assert_array_equal(a, ea)
    np.testing.assert_array_equal(b, eb)
    d1, d2 = torch.from_numpy(C1.astype(np.int64)).cuda(), torch.from_numpy(C2.astype(np.int64)).cuda()
    _same(engine.compare_dev(d1, d2), exp)
    np.testing.assert_array_equal(d1.cpu().numpy(), ea)
    np.testing.assert_array_equal(d2.cpu().numpy(), eb)


def test_compare_louvain_against_planted_truth(tmp_path):
    """The use the metrics exist for: a planted partition as truth against the
    clustering the engine finds, through the clustInfo files."""
    g, blocks = gen_planted(200000, 400, 12, 0.25, seed=3, return_blocks=True)
    opts = driver.ClusteringParameters(coloring=1, output=True, minGraphSize=1000)
    C, st, _ = driver.cluster(g, opts, inFile=str(tmp_path / "planted"))
    C_read = driver.read_communities(tmp_path / "planted_clustInfo")
    np.testing.assert_array_equal(C_read, C)
    exp, _, _ = oracle.compare(blocks, C_read)
    got, _, _ = engine.computeCommunityComparisons(blocks, C_read)
    _same(got, exp)
    assert 0 < got["precision"] <= 1 and 0 < got["recall"] <= 1
    assert got["SameSame"] + got["SameDiff"] == sum(int(c) * (int(c) - 1) // 2 for c in np.bincount(blocks))


def test_compare_rejects_bad_labels():
    with pytest.raises(engine.GlvError):
        engine.computeCommunityComparisons(np.array([0, -2, 1]), np.array([0, 1, 1]))
    with pytest.raises(engine.GlvError):
        engine.computeCommunityComparisons(np.array([0, 1]), np.array([0, 1, 1]))
    with pytest.raises(engine.GlvError):
        engine.computeCommunityComparisons(np.array([-1, -1]), np.array([0, 1]))


@pytest.mark.parametrize("n", [1, 2, 17, 5000, 300000])
def test_gini_equals_oracle(n):
    import torch
    rng = np.random.default_rng(n)
    sz = rng.integers(-5, 100000, n)
    exp, es = oracle.gini(sz)
    got, s = engine.computeGiniCoefficient(sz)
    assert got == exp
    np.testing.assert_array_equal(s, es)
    d = torch.from_numpy(sz).cuda()
    assert engine.gini_dev(d) == exp
    np.testing.assert_array_equal(d.cpu().numpy(), es)
