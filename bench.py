#!/usr/bin/env python3
"""Benchmark of the colored multi-phase Louvain (runMultiPhaseColoring) on one MI355X.

One step = one complete run to convergence (coloring, colored sweeps,
modularity, renumbering, coarsening, the final plain phase) of the seeded
LFR-style graph with 4,194,304 vertices (degrees power law 2.5 on [10, 1000],
community sizes power law 1.5 on [20, 5000], mixing 0.4, weights U[0.5, 2],
seed 2), resident in HBM before the timed region starts, with coloring 1,
threshold 1e-9, C_threshold 1e-6, minGraphSize 1000.

value = edges visited per second over the K timed steps, where edges visited
= sum over phases of iterations x directed entries of the phase's graph.
The JSON line also carries the HBM roofline of the dominant kernel (HIP
events around every launch, algorithmic bytes per launch) and the CPU
baseline: the reference's own OpenMP code (oracle/_ref/ref_bench) on a
bounded sample, or the C restatement when the reference build is absent.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Louvain edges-visited/s + time-to-converge vs plain GPU at ΔQ≤1e-3"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "lfr4m": dict(kind="lfr", nv=4194304, seed=2,
                  desc="LFR-style 4,194,304 vertices, degree exp 2.5 on [10,1000], community "
                       "exp 1.5 on [20,5000], mixing 0.4, weights U[0.5,2], seed 2"),
    "planted65k": dict(kind="planted", nv=65536, seed=1,
                       desc="planted partition 65,536 vertices, 64 blocks, mean degree 16, "
                            "mixing 0.3, unit weights, seed 1"),
    "rmat24": dict(kind="rmat", scale=24, seed=3,
                   desc="RMAT scale 24, edge factor 16, a/b/c/d 0.57/0.19/0.19/0.05, "
                        "symmetrised, unit weights, seed 3"),
}


def make_graph(spec, nv=None):
    from grappolo_amd.graph import gen_lfr, gen_planted, gen_rmat
    if spec["kind"] == "lfr":
        return gen_lfr(nv or spec["nv"], seed=spec["seed"])
    if spec["kind"] == "planted":
        return gen_planted(nv or spec["nv"], 64, 16, 0.3, seed=spec["seed"])
    return gen_rmat(spec["scale"], seed=spec["seed"])


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(spec):
    """The reference's OpenMP colored Louvain on a bounded sample of the
    workload (same generator and parameters, fewer vertices)."""
    from grappolo_amd.graph import write_binary
    sample_nv = {"lfr": 1048576, "planted": 65536, "rmat": None}[spec["kind"]]
    if spec["kind"] == "rmat":
        from grappolo_amd.graph import gen_rmat
        g = gen_rmat(20, seed=spec["seed"])
        sample = "RMAT scale 20 (same a/b/c/d, edge factor 16, seed 3)"
    else:
        g = make_graph(spec, sample_nv)
        sample = f"{spec['kind']} graph with {sample_nv:,} vertices, same parameters and seed"
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    cores = host_cores()
    if os.path.exists(ref_bin):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "g.bin")
            write_binary(path, g)
            env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close",
                       OMP_PLACES="cores")
            out = subprocess.run([ref_bin, path, str(cores), "1000", "1e-9", "1e-6"], env=env,
                                 capture_output=True, text=True, timeout=600, check=True)
            r = json.loads(out.stdout.strip().splitlines()[-1])
        return {"value": r["ev_per_s"], "unit": "edges/s", "cores": cores, "kind": "reference",
                "sample": f"{sample}; reference runMultiPhaseColoring loop, -Ofast -fopenmp, "
                          f"{r['seconds']:.2f} s, final Q {r['final_modularity']:.6f}",
                "final_modularity": r["final_modularity"], "seconds": r["seconds"]}
    import oracle
    t = time.perf_counter()
    _, st = oracle.run_colored(g, coloring="greedy", frozen=False, minGraphSize=1000)
    dt = time.perf_counter() - t
    return {"value": st["edges_visited"] / dt, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"{sample}; C restatement, one thread, {dt:.2f} s",
            "final_modularity": st["final_modularity"], "seconds": dt}


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and parts[0] == kernel:
                try:
                    return float(parts[1])
                except ValueError:
                    return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lfr4m", choices=list(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true",
                    help="do not time kernels with HIP events in the timed region")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the engine runs on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    from grappolo_amd import engine

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    spec = WORKLOADS[args.workload]
    t = time.perf_counter()
    g = make_graph(spec)
    t_gen = time.perf_counter() - t
    print(f"[bench] {args.workload}: {g} generated in {t_gen:.1f}s", file=sys.stderr, flush=True)
    dg = engine.DeviceGraph(g, device="cuda:0")
    C = torch.empty(g.numVertices, dtype=torch.int64, device="cuda:0")
    params = dict(minGraphSize=1000, threshold=1e-9, C_threshold=1e-6)

    for i in range(args.warmup):
        st = engine.run_colored_dev(dg, C, **params)
        print(f"[bench] warmup {i}: {st['time_total'] * 1e3:.1f} ms Q={st['final_modularity']:.6f}",
              file=sys.stderr, flush=True)

    profile = not args.no_profile
    if profile:
        engine.profile_reset()
        engine.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runs = []
    for i in range(args.steps):
        runs.append(engine.run_colored_dev(dg, C, **params))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern = {}
    if profile:
        engine.profile_enable(False)
        kern = engine.profile_read()
    elapsed = t1 - t0
    ev = sum(r["edges_visited"] for r in runs)
    last = runs[-1]
    pp = last["per_phase"]
    value = ev / elapsed

    roofline = None
    if kern:
        sweeps = {k: v for k, v in kern.items() if v[2] > 0 and k.startswith("sweep")}
        dom = max(sweeps, key=lambda k: sweeps[k][0])
        ms, by, n = sweeps[dom]
        achieved = by / (ms * 1e-3) / 1e9
        traffic = read_traffic(dom)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "kernel": dom, "launches": n,
                    "avg_launch_us": round(ms * 1e3 / n, 3),
                    "algorithmic_bytes_per_launch": round(by / n, 1)}

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(spec)

    out = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": spec["desc"], "nv": g.numVertices, "nnz": g.nnz,
                   "coloring": 1, "threshold": 1e-9, "C_threshold": 1e-6,
                   "minGraphSize": 1000, "parallelism": "single GPU"},
        "time_to_converge_ms": elapsed / args.steps * 1e3,
        "final_modularity": last["final_modularity"], "phases": last["phases"],
        "iterations": last["total_iterations"], "colors": last["colors_first"],
        "clusters": last["final_clusters"], "edges_visited_per_step": ev / args.steps,
        "per_phase": [{"nv": int(pp["phase_nv"][i]), "nnz": int(pp["phase_nnz"][i]),
                       "colors": int(pp["phase_colors"][i]), "iterations": int(pp["phase_iters"][i]),
                       "clusters": int(pp["phase_clusters"][i]),
                       "ms": round(float(pp["phase_seconds"][i]) * 1e3, 2)}
                      for i in range(len(pp["phase_nv"]))],
        "roofline": roofline, "cpu_baseline": cpu,
        "kernels": {k: {"ms": round(v[0], 3), "GB/s": round(v[1] / (v[0] * 1e-3) / 1e9, 1)
                        if v[0] > 0 else None, "launches": v[2]} for k, v in kern.items() if v[2]},
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
