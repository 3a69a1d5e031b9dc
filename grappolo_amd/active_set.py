"""Learned active-set pruning of the colored Louvain sweeps.

An 8 -> 64 -> 64 -> 1 MLP (BF16 weights on the MI355X matrix cores, FP32
accumulation) is trained on traces of the unpruned colored method -- per
vertex: its features after one sweep and whether the next sweep moves it --
and then predicts, at the start of every iteration after the first of a
colored phase, which vertices will not move; those are not swept.  A phase
ends only when an iteration over every vertex gains less than the threshold,
so the unpruned method is both the fallback and the arbiter of the
modularity reached.  Phases of at most 15360 vertices (the dense persistent
sweep, where a color step costs its heaviest vertex and a grid barrier
whatever is pruned) run unpruned.  There is no counterpart in the reference.

Features of a vertex (8 BF16 values, computed on the device):
  0 log2(1 + entries)/10, 1 weight to own community / degree,
  2 slog(best gain / 2k), 3 slog(second-best gain / 2k),
  4 min(sweeps since last move, 16)/16, 5 color / colors,
  6 log2(1 + own community size)/16, 7 min(phase, 15)/16
with slog(g) = sign(g) log2(1 + 1024|g|)/11 and -1 where no other community
is adjacent.

All computation runs in libgrappolo_hip.so on the GPU; torch only holds the
device buffers.
"""
import ctypes
import math

import numpy as np

from . import engine
from .engine import _VP, _call, _ptr, _stream


class LasStats(ctypes.Structure):
    _fields_ = [("swept_vertices", ctypes.c_int64), ("swept_entries", ctypes.c_int64),
                ("full_sweeps", ctypes.c_int32), ("pruned_sweeps", ctypes.c_int32),
                ("samples", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def param_count():
    return engine._f("glv_mlp_param_count")()


def init_params(seed=0, device="cuda"):
    """He-normal weights and zero biases, as a float32 device tensor."""
    import torch
    n = param_count()
    rng = np.random.default_rng(seed)
    P = np.zeros(n, np.float32)
    P[0:512] = rng.normal(0, math.sqrt(2.0 / 8), 512)
    P[576:4672] = rng.normal(0, math.sqrt(2.0 / 64), 4096)
    P[4736:4800] = rng.normal(0, math.sqrt(1.0 / 64), 64)
    return torch.from_numpy(P).to(device)


def forward_dev(P, feats):
    """Logits (float32 tensor) of the BF16 feature rows feats (int16 tensor n x 8)."""
    import torch
    n = feats.shape[0]
    out = torch.empty(n, dtype=torch.float32, device=feats.device)
    _call("glv_mlp_forward_dev", _ptr(P), _ptr(feats), n, _ptr(out), _stream())
    return out


def grad_dev(P, feats, labels, pos_weight=1.0):
    """(loss, gradient tensor) of the mean weighted logistic loss over all rows."""
    import torch
    n = feats.shape[0]
    g = torch.empty(param_count() + 1, dtype=torch.float32, device=feats.device)
    _call("glv_mlp_grad_dev", _ptr(P), _ptr(feats), _ptr(labels), n, float(pos_weight), _ptr(g),
          _stream())
    return float(g[-1].item()), g[:-1]


def train_dev(P, m, v, feats, labels, steps, batch=4096, lr=1e-3, pos_weight=1.0, seed=0, step0=0):
    """Adam steps in place on P (moments m, v); returns the per-step losses."""
    import torch
    losses = torch.empty(max(int(steps), 1), dtype=torch.float32, device=feats.device)
    _call("glv_mlp_train_dev", _ptr(P), _ptr(m), _ptr(v), _ptr(feats), _ptr(labels),
          feats.shape[0], int(steps), int(batch), float(lr), float(pos_weight),
          ctypes.c_uint64(seed), int(step0), _ptr(losses), _stream())
    return losses[:int(steps)]


def trace_dev(dg, C_orig, rate, seed, cap, minGraphSize=1000, threshold=1e-9, C_threshold=1e-6):
    """The unpruned colored run with trace capture.  Returns (feats, labels,
    verts, stats, las) with the first min(samples, cap) rows."""
    import torch
    dev = C_orig.device
    feats = torch.empty((max(cap, 1), 8), dtype=torch.int16, device=dev)
    verts = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
    labels = torch.empty(max(cap, 1), dtype=torch.uint8, device=dev)
    st = engine.new_stats()
    ls = LasStats()
    _call("glv_las_trace_dev", *dg.ptrs(), _ptr(C_orig), int(minGraphSize), threshold, C_threshold,
          float(rate), ctypes.c_uint64(seed), _ptr(feats), _ptr(verts), _ptr(labels), int(cap),
          _stream(), ctypes.byref(st), ctypes.byref(ls))
    n = min(ls.samples, cap)
    return feats[:n], labels[:n], verts[:n], engine.stats_dict(st), ls.as_dict()


def run_colored_las_dev(dg, C_orig, P, tau, minGraphSize=1000, threshold=1e-9, C_threshold=1e-6,
                        max_rechecks=0):
    """runMultiPhaseColoring with learned pruning; returns (stats, las stats)."""
    st = engine.new_stats()
    ls = LasStats()
    _call("glv_runMultiPhaseColoring_las_dev", *dg.ptrs(), _ptr(C_orig), int(minGraphSize),
          threshold, C_threshold, _ptr(P), float(tau), int(max_rechecks), _stream(),
          ctypes.byref(st), ctypes.byref(ls))
    return engine.stats_dict(st), ls.as_dict()


def choose_tau(logits, labels, recall=0.98):
    """Largest probability threshold that keeps `recall` of the movers
    (label-1 rows) active; returns (tau, fraction of rows kept)."""
    import torch
    p = torch.sigmoid(logits.double())
    pos = p[labels.bool()]
    if pos.numel() == 0:
        return 0.0, 1.0
    q = torch.quantile(pos, 1.0 - recall).item()
    tau = min(max(q, 0.0), 0.999)
    return tau, float((p >= tau).double().mean().item())


class ActiveSetPredictor:
    """Parameters, Adam state and threshold of a trained predictor."""

    def __init__(self, seed=0, device="cuda"):
        import torch
        self.P = init_params(seed, device)
        self.m = torch.zeros_like(self.P)
        self.v = torch.zeros_like(self.P)
        self.step = 0
        self.tau = 0.0

    def fit(self, feats, labels, steps=2000, batch=4096, lr=2e-3, pos_weight=None, seed=1):
        """Adam on the trace rows; pos_weight defaults to the class ratio (capped at 20)."""
        if pos_weight is None:
            pos = float(labels.double().mean().item())
            pos_weight = min(20.0, (1.0 - pos) / max(pos, 1e-6)) if pos > 0 else 1.0
        losses = train_dev(self.P, self.m, self.v, feats, labels, steps, batch, lr, pos_weight,
                           seed, self.step)
        self.step += steps
        return losses

    def calibrate(self, feats, labels, recall=0.98):
        self.tau, kept = choose_tau(forward_dev(self.P, feats), labels, recall)
        return self.tau, kept

    def run(self, dg, C_orig, max_rechecks=0, **params):
        return run_colored_las_dev(dg, C_orig, self.P, self.tau, max_rechecks=max_rechecks, **params)
