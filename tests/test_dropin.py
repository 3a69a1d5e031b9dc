"""The C++ drop-in (include/grappolo/*.h) exercised by a compiled caller.

tests/dropin/dropin_caller.cpp is written against the drop-in headers with
the call shape of the reference's clustering driver
(driverForGraphClustering.cpp:135-203: malloc()ed graph handed over,
C_orig set to -1, the -o file written afterwards), linked to
libgrappolo_hip.so and run as a fresh process per case.  Its C_orig must
equal the oracle's, the library must free G exactly once (glibc aborts on a
double free; the caller allocates again afterwards), and a call the
reference rejects with an assert must end the process with status 1 and a
message."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from grappolo_amd.graph import gen_planted, write_binary

from graphs import hubs, karate

HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "dropin", "dropin_caller")


def test_caller_is_built_and_links():
    """build() compiles the caller against the headers and the library."""
    assert os.path.exists(CALLER), "run __graft_entry__.build() first"
    out = subprocess.run(["ldd", CALLER], capture_output=True, text=True, check=True).stdout
    assert "libgrappolo_hip.so" in out and "not found" not in out.split("libgrappolo_hip.so")[1].split("\n")[0]


def _run(tmp_path, g, mode, mgs):
    path = tmp_path / "g.bin"
    write_binary(path, g)
    out = tmp_path / "out.txt"
    env = dict(os.environ, MALLOC_CHECK_="3")
    r = subprocess.run([CALLER, str(path), str(out), mode, str(mgs)], capture_output=True, text=True,
                       timeout=300, env=env)
    return r, out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["karate", "planted", "hubs"])
def test_runMultiPhaseColoring_from_compiled_caller(tmp_path, name):
    g = {"karate": karate, "planted": lambda: gen_planted(4096, 16, 16, 0.3, seed=5), "hubs": hubs}[name]()
    mgs = 10 if g.numVertices < 100 else 100
    print(f"dropin color {name} nv={g.numVertices}", flush=True)
    r, out = _run(tmp_path, g, "color", mgs)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "dropin ok" in r.stdout
    got = np.loadtxt(out, dtype=np.int64)
    exp, _ = oracle.run_colored(g, coloring="jp", frozen=True, minGraphSize=mgs)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.gpu
def test_runMultiPhaseBasic_from_compiled_caller(tmp_path):
    g = gen_planted(4096, 16, 16, 0.3, seed=5)
    print("dropin basic planted", flush=True)
    r, out = _run(tmp_path, g, "basic", 1000)
    assert r.returncode == 0, r.stderr + r.stdout
    got = np.loadtxt(out, dtype=np.int64)
    exp, _ = oracle.run_basic(g, basicOpt=0)
    np.testing.assert_array_equal(got, exp)


@pytest.mark.gpu
def test_stage_entry_points_from_compiled_caller(tmp_path):
    """Coloring, one colored phase, renumbering and coarsening through the
    stage drop-ins; the caller frees the returned coarse graph with free()."""
    g = gen_planted(4096, 16, 16, 0.3, seed=5)
    print("dropin stages planted", flush=True)
    r, out = _run(tmp_path, g, "stages", 0)
    assert r.returncode == 0, r.stderr + r.stdout
    lines = out.read_text().split("\n")
    ncol, itr, q, k, nv2, ne2 = lines[0].split()
    C = np.array(lines[1:1 + g.numVertices], dtype=np.int64)
    off = np.array(lines[1 + g.numVertices:2 + g.numVertices + int(nv2)], dtype=np.int64)
    colors, mc, _ = oracle.color_jp(g)
    assert int(ncol) == mc + 1
    qo, Co, ito, _ = oracle.louvain_colored(g, colors, mc + 1, -1.0, 1e-6, frozen=True)
    assert int(itr) == ito and float(q) == qo
    Cr, kr = oracle.renumber(Co)
    np.testing.assert_array_equal(C, Cr)
    assert int(k) == kr and int(nv2) == kr
    exp = oracle.coarsen(g, Cr, kr)
    np.testing.assert_array_equal(off, exp.edgeListPtrs)
    assert int(ne2) == exp.numEdges


@pytest.mark.gpu
def test_rejected_call_exits_with_status_1(tmp_path):
    """runMultiPhaseColoring.cpp:55 asserts 0 < coloring < 4; the drop-in
    prints the reason and exits with status 1 instead of aborting."""
    g = karate()
    print("dropin badcolor karate", flush=True)
    r, _ = _run(tmp_path, g, "badcolor", 10)
    assert r.returncode == 1
    assert "coloring" in (r.stderr + r.stdout).lower()
