"""The clustering driver's flow on the device engine.

driverForGraphClustering.cpp:60-245 reads a graph, optionally folds
degree-0/degree-1 vertices (vertex following, :112-132), runs the clustering
entry point its options select (:179-189) and optionally writes the labels to
`<inFile>_clustInfo`, one "%ld" per line (:192-203).  driverForClusterComparison
.cpp:47-91 reads two such files and compares them (computeCommunityComparisons).
Here the graph stays resident on the device from the fold to the last phase.
"""
import dataclasses
import os

import numpy as np

from . import engine


@dataclasses.dataclass
class ClusteringParameters:
    """clustering_parameters (DefineStructure/defs.h:105-126) with the defaults
    of parseInputParameters.cpp:46-49.  The thread and strong-scaling options
    of the reference do not apply to one device."""
    VF: bool = False           # -v
    output: bool = False       # -o
    coloring: int = 0          # -c: 0 none, 1 distance-1, 2 balanced, 3 incomplete multi-hash
    numColors: int = 16        # -n
    basicOpt: int = 0          # -b: 1 = replaceMap / NoMap sweeps
    syncType: int = 0          # -y (runMultiPhaseSyncType is not provided)
    C_thresh: float = 0.01     # -d
    minGraphSize: int = 100000  # -m
    threshold: float = 1e-6    # -t


def cluster(G, opts=None, inFile=None, device="cuda"):
    """Runs the driver's flow on Graph G.  Returns (C_orig, stats, graph):
    C_orig as numpy labels of the graph actually clustered (the folded graph
    when VF is on, driverForGraphClustering.cpp:123-128), the engine's run
    statistics and that graph's DeviceGraph.  With opts.output the labels are
    written to `<inFile>_clustInfo`."""
    import torch
    opts = opts or ClusteringParameters()
    if opts.syncType != 0:
        raise engine.GlvError("cluster", 5, "runMultiPhaseSyncType is not provided")
    dg = engine.DeviceGraph(G, device)
    folded = 0
    if opts.VF:
        dg, folded = engine.vertex_following_prepass_dev(dg)
    replaceMap = 1 if opts.basicOpt == 1 else 0  # :106-108
    C = torch.full((dg.numVertices,), -1, dtype=torch.int64, device=device)
    if dg.numVertices == 0:
        st = {}
    elif opts.coloring != 0:
        st = engine.run_colored_dev(dg, C, coloring=opts.coloring, numColors=opts.numColors,
                                    replaceMap=replaceMap, minGraphSize=opts.minGraphSize,
                                    threshold=opts.threshold, C_threshold=opts.C_thresh)
    else:
        st = engine.run_basic_dev(dg, C, basicOpt=opts.basicOpt, threshold=opts.threshold)
    st["vertices_folded"] = folded
    C_orig = C.cpu().numpy()
    if opts.output:
        if inFile is None:
            raise ValueError("output needs inFile")
        write_clust_info(inFile, C_orig)
    return C_orig, st, dg


def clust_info_path(inFile):
    """driverForGraphClustering.cpp:195: `<inFile>_clustInfo`."""
    return f"{os.fspath(inFile)}_clustInfo"


def write_clust_info(inFile, C):
    """One "%ld\\n" line per vertex (driverForGraphClustering.cpp:198-202).
    Returns the path written."""
    path = clust_info_path(inFile)
    C = np.ascontiguousarray(C, dtype=np.int64)
    with open(path, "wb") as f:
        if C.size:
            f.write(("\n".join(map(str, C.tolist())) + "\n").encode())
    return path


def read_communities(path):
    """The community-file parse of driverForClusterComparison.cpp:59-83:
    whitespace-separated integers read until end of file.  Its loop pushes
    one value past the last successful read and then drops the final
    element, so a file that does not end in whitespace loses its last
    label; that is kept."""
    with open(path, "rb") as f:
        text = f.read()
    vals = np.array(text.split(), dtype=np.int64) if text.strip() else np.zeros(0, np.int64)
    if vals.size and not text[-1:].isspace():
        vals = vals[:-1]
    return vals
