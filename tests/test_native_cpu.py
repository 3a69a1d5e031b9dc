"""CPU-side checks of the native libraries: every symbol the headers declare
is exported, the generators produce the graphs they promise, and the file
formats are byte-compatible with the reference's own readers and writer."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from grappolo_amd import _native
from grappolo_amd.graph import (Graph, gen_lfr, gen_planted, gen_rmat, load_metis, read_binary,
                                write_binary)

from graphs import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
need_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference checker not built")


def _declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(glv_\w+)\s*\(", text)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_hip_library_exports_every_declared_symbol():
    syms = _exported(_native.HIP_LIB)
    missing = [f for f in _declared("grappolo_hip.h") if f not in syms]
    assert not missing, missing
    lib = ctypes.CDLL(_native.HIP_LIB)  # loads without a GPU
    for f in _declared("grappolo_hip.h"):
        getattr(lib, f)


def test_hip_library_exports_reference_cpp_api():
    """The C++ entry points keep the reference's mangled names, so callers
    built against DefineStructure/*.h relink unchanged."""
    syms = _exported(_native.HIP_LIB)
    for m in ["_Z21runMultiPhaseColoringP5graphPliiilddii",
              "_Z18runMultiPhaseBasicP5graphPlilddii",
              "_Z30algoLouvainWithDistOneColoringP5graphPliPiiddPdS2_",
              "_Z35algoLouvainWithDistOneColoringNoMapP5graphPliPiiddPdS2_",
              "_Z32algoDistanceOneVertexColoringOptP5graphPiiPd",
              "_Z21parallelLouvianMethodP5graphPliddPdPi",
              "_Z26parallelLouvianMethodNoMapP5graphPliddPdPi",
              "_Z28renumberClustersContiguouslyPll",
              "_Z22buildNextLevelGraphOptP5graphS0_Plli",
              "_Z15vertexFollowingP5graphPl",
              "_Z15buildNewGraphVFP5graphS0_Pll",
              "_Z22computeGiniCoefficientPli",
              "_Z27algoColoringMultiHashMaxMinP5graphPiiPdii",
              "_Z19vBaseRedistributionP5graphPiii",
              "_Z27computeCommunityComparisonsRSt6vectorIlSaIlEElS2_l"]:
        assert m in syms, m


def test_host_library_exports_every_declared_symbol():
    syms = _exported(_native.HOST_LIB)
    missing = [f for f in _declared("glv_gen.h") if f not in syms]
    assert not missing, missing


def test_reports_no_device_on_cpu_host():
    from grappolo_amd import engine
    assert engine.device_count() >= 0


def _check_simple_symmetric(g):
    rows = np.repeat(np.arange(g.numVertices), g.degrees())
    assert np.all(rows != g.tails)  # no self-loops
    key = rows * g.numVertices + g.tails
    assert np.all(np.diff(key) > 0)  # rows sorted, no repeats
    order = np.lexsort((rows, g.tails))
    np.testing.assert_array_equal(g.tails[order], rows)  # symmetric
    np.testing.assert_array_equal(g.weights[order], g.weights)
    assert g.numEdges * 2 == g.nnz


def test_planted_generator():
    g, blocks = gen_planted(65536, 64, 16, 0.3, seed=1, return_blocks=True)
    assert g.numVertices == 65536 and g.nnz == 65536 * 16  # exactly nv*deg/2 edges
    _check_simple_symmetric(g)
    assert np.all(np.bincount(blocks) == 1024)
    rows = np.repeat(np.arange(g.numVertices), g.degrees())
    inter = np.mean(blocks[rows] != blocks[g.tails])
    assert abs(inter - 0.3) < 0.01
    assert np.all(g.weights == 1.0)
    g2 = gen_planted(65536, 64, 16, 0.3, seed=1)
    np.testing.assert_array_equal(g.tails, g2.tails)  # seeded


def test_lfr_generator():
    g, comm = gen_lfr(200000, seed=2, return_communities=True)
    _check_simple_symmetric(g)
    d = g.degrees()
    assert d.max() <= 1000 and np.percentile(d, 5) >= 8
    assert 22 < g.nnz / g.numVertices < 27  # discrete power law 2.5 on [10,1000]: mean ~25.8
    sizes = np.bincount(comm)
    assert sizes.min() >= 20 and sizes.max() <= 5000 + 20
    rows = np.repeat(np.arange(g.numVertices), d)
    inter = np.mean(comm[rows] != comm[g.tails])
    assert abs(inter - 0.4) < 0.03
    assert g.weights.min() >= 0.5 and g.weights.max() <= 2.0


def test_rmat_generator():
    g = gen_rmat(14, seed=3)
    _check_simple_symmetric(g)
    d = g.degrees()
    assert np.mean(d == 0) > 0.2  # skew leaves many vertices isolated
    assert d.max() > 50 * np.mean(d[d > 0])
    np.testing.assert_array_equal(gen_rmat(14, seed=3).tails, g.tails)


def test_binary_round_trip(tmp_path):
    g = gen_planted(2048, 8, 10, 0.3, seed=4, wlo=0.5, whi=2.0)
    p = tmp_path / "g.bin"
    write_binary(p, g)
    assert os.path.getsize(p) == 24 + 8 * (g.numVertices + 1) + 24 * g.nnz
    h = read_binary(p)
    for a, b in ((g.edgeListPtrs, h.edgeListPtrs), (g.tails, h.tails), (g.weights, h.weights)):
        np.testing.assert_array_equal(a, b)
    assert h.numEdges == g.numEdges


@need_ref
def test_binary_format_matches_reference(tmp_path):
    g = gen_planted(2048, 8, 10, 0.3, seed=4, wlo=0.5, whi=2.0)
    ours, theirs = tmp_path / "ours.bin", tmp_path / "ref.bin"
    write_binary(ours, g)
    oracle.ref_write_binary(theirs, g)
    assert ours.read_bytes() == theirs.read_bytes()
    r = oracle.ref_load_binary(ours)
    np.testing.assert_array_equal(r.tails, g.tails)
    np.testing.assert_array_equal(r.weights, g.weights)


def test_metis_karate():
    g = load_metis(os.path.join(GOLDEN, "karate.graph"))
    assert g.numVertices == 34 and g.nnz == 156 and g.numEdges == 78
    _check_simple_symmetric(g)


@need_ref
def test_metis_matches_reference_loader():
    path = os.path.join(GOLDEN, "karate.graph")
    g, r = load_metis(path), oracle.ref_load_metis(path)
    for a, b in ((g.edgeListPtrs, r.edgeListPtrs), (g.tails, r.tails), (g.weights, r.weights)):
        np.testing.assert_array_equal(a, b)


def test_invalid_binary_rejected(tmp_path):
    p = tmp_path / "bad.bin"
    np.array([3, 1, 1, 0, 1, 2, 99], dtype=np.int64).tofile(p)  # offsets end past nnz
    with pytest.raises(RuntimeError):
        read_binary(p)


def test_graph_container_edge_list_round_trip():
    g = gen_planted(512, 4, 8, 0.2, seed=5)
    e = g.edge_list()
    assert e.dtype.itemsize == 24
    h = Graph.from_edge_list(g.edgeListPtrs, e, g.numEdges)
    np.testing.assert_array_equal(h.tails, g.tails)
    np.testing.assert_array_equal(e["head"], np.repeat(np.arange(512), g.degrees()))


# ------------------------------------------------ driver output and input files
def test_clust_info_writer_format(tmp_path):
    """driverForGraphClustering.cpp:195-202: `<inFile>_clustInfo`, one "%ld\n"
    per vertex."""
    from grappolo_amd import driver
    C = np.array([3, -1, 0, 12345678901, 7], np.int64)
    path = driver.write_clust_info(str(tmp_path / "g.bin"), C)
    assert path == str(tmp_path / "g.bin_clustInfo")
    assert open(path, "rb").read() == b"3\n-1\n0\n12345678901\n7\n"
    np.testing.assert_array_equal(driver.read_communities(path), C)


def test_read_communities_keeps_the_reference_parse(tmp_path):
    """driverForClusterComparison.cpp:59-83: tokens until end of file; a file
    without trailing whitespace loses its last value, an empty file gives none."""
    from grappolo_amd import driver
    cases = {b"1 2\n3\n": [1, 2, 3], b"1\n2\n3": [1, 2], b"": [], b"\n": [], b"5": [],
             b"  4\t-1\r\n": [4, -1]}
    for i, (text, exp) in enumerate(cases.items()):
        p = tmp_path / f"c{i}"
        p.write_bytes(text)
        np.testing.assert_array_equal(driver.read_communities(p), np.array(exp, np.int64))
