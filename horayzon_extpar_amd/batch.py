"""Batched domains: one persistent trace launch over many resident domains,
replayed from a captured HIP graph, and a grid-search autotuner of its
launch parameters (no counterpart in the reference, which handles one
domain per call).

A batch holds domains that share the search parameters (num_hori,
dist_search, refine_factor, svf_type); ``group_by_nhori`` splits a mixed
workload.  Results of a batched run are bit-identical to separate runs of
each domain (tests/test_gpu_batch.py).
"""

from __future__ import annotations

import ctypes
import itertools
import time

from . import _lib
from ._lib import HzStats, check, default_options
from .horizon_svf import ResidentDomain


class Batch:
    def __init__(self, domains):
        self.domains = list(domains)
        arr = (ctypes.c_void_p * len(self.domains))(*[d._h for d in self.domains])
        h = ctypes.c_void_p()
        check(_lib.lib().hz_batch_create(arr, len(self.domains), ctypes.byref(h)))
        self._h = h
        self.ncell = sum(d.ncell for d in self.domains)

    def run(self, num_hori=24, dist_search=40000.0, refine_factor=1, svf_type=2,
            use_graph=True, chunk=None, **opts):
        opt = default_options(**opts)
        if chunk:
            opt.queue_chunk = int(chunk)
        st = HzStats()
        check(_lib.lib().hz_batch_run(self._h, int(num_hori), float(dist_search),
                                      int(refine_factor), int(svf_type), ctypes.byref(opt),
                                      int(bool(use_graph)), ctypes.byref(st)))
        for d in self.domains:
            d.num_hori = int(num_hori)
        return st.as_dict()

    def fetch(self):
        return [d.fetch() for d in self.domains]

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hz_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def group_by_nhori(items):
    """items: iterable of (Domain, nhori) -> {nhori: [Domain, ...]}"""
    groups = {}
    for dom, nh in items:
        groups.setdefault(int(nh), []).append(dom)
    return groups


DEFAULT_SPACE = dict(leaf_size=(1, 2), refill=(32, 48), chunk=(32, 64), block_size=(128,))


def autotune(batch: Batch, num_hori=24, dist_search=40000.0, refine_factor=1, svf_type=2,
             space=None, repeats=1):
    """Time every combination of the search space (graph replay of the trace
    launch, HIP events) and return (best options dict, best trace ms, table).
    The first run of a configuration captures its graph and is not timed."""
    space = dict(DEFAULT_SPACE if space is None else space)
    keys = list(space)
    table = []
    best = None
    for values in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, values))
        batch.run(num_hori, dist_search, refine_factor, svf_type, use_graph=True, **cfg)
        t = min(batch.run(num_hori, dist_search, refine_factor, svf_type, use_graph=True,
                          **cfg)["t_trace_ms"] for _ in range(max(1, repeats)))
        table.append((cfg, t))
        if best is None or t < best[1]:
            best = (cfg, t)
    return best[0], best[1], table


def separate_trace_ms(domains, num_hori=24, dist_search=40000.0, refine_factor=1, svf_type=2,
                      **opts):
    """Sum of the trace times of one launch per domain (the unbatched form)."""
    t = 0.0
    for d in domains:
        t += d.run(num_hori, dist_search, refine_factor, svf_type, **opts)["t_trace_ms"]
    return t


def make_resident(items, grid_type=1):
    t0 = time.perf_counter()
    out = [(ResidentDomain(*dom.args(), grid_type=grid_type), nh) for dom, nh in items]
    return out, time.perf_counter() - t0
