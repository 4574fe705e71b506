"""Benchmark: horizon + SVF throughput on one MI355X.

Workload (default): 1 km ICON-like grid over a 600 x 400 km box (553,200
cells, 1.66M grid_type-1 triangles), alpine-like synthetic terrain, nhori=24,
dist_search=40 km, ray_org_elev=0.2 m, refine_factor=1, svf_type=2.

One step = one full device pass of the hot path with the FP32 scene and cell
frames already resident in HBM: LBVH build, Morton cell ordering, the
elevation search of every cell and the fused sector mean / SVF, outputs left
in HBM.  value = cells per second over the timed steps of the plain GPU
method (the reference walk, search_mode 0).

The same run times the learned form (predictor + certified lattice search,
results equal to the walk's) and the certified climb (the lattice search
without the predictor), each against the walk's results, the C-ABI call end
to end, and the batched 64-domain workload.

Prints ONE JSON line (rank 0).  The host-CPU baseline (the oracle's C++ port
of the reference walk, every host CPU this process may use, time-budgeted
sample) is timed in the same run unless --no-cpu.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--mode", default="walk", choices=["walk", "climb"])
    ap.add_argument("--leaf", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--kernel", type=int, default=0, help="0 persistent, 1 one lane per cell")
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--morton", type=int, default=0, help="LBVH codes: 0 3D, 1 2D")
    ap.add_argument("--width", type=int, default=0,
                    help="traversal tree: 2 binary (default), 4 four-wide quantized")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 64-domain leg (config 4)")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the end-to-end C-ABI call legs")
    ap.add_argument("--weights", default="")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a plain copy "
                         "kernel of known traffic, for scaling rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE)")
    ap.add_argument("--calibrate-gather", type=int, default=0, metavar="MIB",
                    help="first run one gather of the 64-byte rows of a MIB MiB table in "
                         "random order, each row once (scattered reads of known traffic)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the "
              "horizon search shards by cell ranges with a replicated scene and is not "
              "run across GPUs here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from horayzon_extpar_amd import synthetic
    from horayzon_extpar_amd.horizon_svf import ResidentDomain

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
    if args.calibrate_gather:
        calibrate_gather(args.calibrate_gather)
    mode = {"walk": 0, "climb": 1}[args.mode]
    dom_in = synthetic.CONFIGS[args.config]()
    nhori, dist, rf, svf = 24, 40000.0, 1, 2
    print(f"[bench] workload {args.config}: ncell={dom_in.ncell} nvert={dom_in.nvert}",
          file=sys.stderr, flush=True)
    opts = dict(search_mode=mode, leaf_size=args.leaf or None, block_size=args.block or None,
                trace_kernel=args.kernel, refill=args.refill or None, bvh_morton=args.morton,
                bvh_width=args.width or None)
    with ResidentDomain(*dom_in.args(), grid_type=1, ray_org_elev=0.2) as dom:
        cst = dom.create_stats
        for _ in range(args.warmup):
            dom.run(nhori, dist, rf, svf, **opts)
        # instrumented pass (untimed): node visits and triangle tests per ray
        cnt = dom.run(nhori, dist, rf, svf, collect_counters=1, **opts)
        print(f"[bench] counters: rays={cnt['num_rays']} nodes={cnt['node_visits']} "
              f"tris={cnt['tri_tests']} depth={cnt['bvh_depth']}", file=sys.stderr, flush=True)
        t_trace, t_bvh, t_ord, rays = [], [], [], 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = dom.run(nhori, dist, rf, svf, **opts)
            t_trace.append(st["t_trace_ms"])
            t_bvh.append(st["t_bvh_ms"])
            t_ord.append(st["t_order_ms"])
            rays = st["num_rays"]
        t1 = time.perf_counter()
        h, s = dom.fetch()
        learned_out = None
        if not args.no_learned and mode == 0:
            learned_out = learned_leg(dom, args, h, s, dom_in, (t1 - t0) / args.steps)
    ncell = dom_in.ncell
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = ncell / (ms_step * 1e-3)
    tr = float(np.mean(t_trace))
    n_node = cnt["node_visits"] / cnt["num_rays"]
    n_tri = cnt["tri_tests"] / cnt["num_rays"]
    b_ray = 64.0 * n_node + 36.0 * n_tri
    bytes_launch = rays * b_ray + ncell * (64 + 8 * (nhori + 1))
    achieved = bytes_launch / (tr * 1e-3) / 1e9
    traffic = profiled_traffic(args)
    node_rate = cnt["node_visits"] / (tr * 1e-3)
    gather = gather_ceiling(cnt["bvh_nodes"], node_rate) if args.kernel == 0 else None
    out = {
        "metric": "horizon+SVF cells/s @nhori=24, \u22640.1\u00b0 err; rays/s; learned vs plain "
                  "speedup",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 ray tests, f64 walk state",
        "data": "synthetic (seeded ICON-like grid and terrain)",
        "config": {"workload": f"{args.config}: 600x400 km @1 km ICON-like grid, "
                               f"{ncell} cells, {cnt['num_triangles']} triangles (grid_type 1)",
                   "num_hori": nhori, "dist_search_m": dist, "refine_factor": rf,
                   "svf_type": svf, "search": args.mode,
                   "trace_kernel": ["persistent", "lane-per-cell"][args.kernel],
                   "bvh": "four-wide, quantized child boxes" if (args.width == 4 and
                                                                 args.kernel == 0)
                   else "binary"},
        "rays_per_s": rays / (tr * 1e-3),
        "rays_per_cell_sector": rays / (ncell * nhori),
        "ms_trace": tr, "ms_bvh": float(np.mean(t_bvh)), "ms_order": float(np.mean(t_ord)),
        "ms_prologue_host": cst.t_prologue_ms, "ms_upload": cst.t_upload_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic[0] if traffic else None,
                     "traffic_gather_scaled": traffic[1] if traffic else None,
                     "traffic_source": traffic[2] if traffic else None,
                     "frac_measured_traffic": (traffic[0] / (tr * 1e-3) / 1e9 / PEAK_HBM_GBS
                                               if traffic else None),
                     "frac_measured_traffic_gather_scaled": (
                         traffic[1] / (tr * 1e-3) / 1e9 / PEAK_HBM_GBS
                         if traffic and traffic[1] else None),
                     "limiter": "the node fetches run at the GPU's rate for dependent "
                                "random 64-byte gathers (gather_ceiling, measured in this "
                                "run; it does not rise from 2 to 8 wavefronts per SIMD); "
                                "symptoms: VALU 64-65% busy at 24% lane use, the vector "
                                "data path 81% busy with 44% waiting on L1 for node fetches "
                                "that missed L2, HBM below peak; the HBM roofline is the "
                                "bound the kernel is priced against",
                     "gather_ceiling": gather,
                     "limiter_source": "profiles/k_trace_c2_counters.md, "
                                       "profiles/k_trace_c2_memory_path.md",
                     "kernel": "k_trace", "bytes_per_ray": b_ray,
                     "nodes_per_ray": n_node, "tris_per_ray": n_tri},
        "trace_counters": {"refill_frac": cnt["refill_frac"],
                           "lanes_per_iter": cnt["lanes_per_iter"],
                           "bvh_nodes": cnt["bvh_nodes"]},
        "horizon_checksum": float(np.sum(h)), "skyview_mean": float(np.mean(s)),
    }
    if learned_out:
        out["learned"] = learned_out
        out["learned_vs_plain_speedup"] = learned_out["speedup_vs_plain"]
        out["learned_vs_climb_speedup"] = learned_out["learned_vs_climb_speedup"]
    if not args.no_e2e:
        out["end_to_end"] = end_to_end_leg(dom_in, opts)
        out["end_to_end_device_prologue"] = end_to_end_leg(dom_in, dict(opts, prologue="device"))
    if not args.no_batch and mode == 0:
        out["batched_c4"] = batch_leg(args)
    if not args.no_cpu:
        out["cpu_baseline"], h_orc, s_orc = cpu_baseline(dom_in, args.cpu_budget,
                                                        args.cpu_threads)
        if mode == 0:
            out["accuracy"] = accuracy_leg(dom_in, h, s, h_orc, s_orc, args.cpu_threads)
    print(json.dumps(out), flush=True)


def gather_ceiling(num_nodes, node_rate):
    """The GPU's rate for what bounds the walk's node fetches: dependent
    random reads of 64 bytes from 128-byte records (hz_calibrate_gather), on
    an array of the trace's tree size, at the trace kernel's occupancy, fully
    random and with every other step in an 8 MB resident set (the walk's
    upper tree levels stay in L2); the trace kernel's node fetches per
    second against both."""
    from horayzon_extpar_amd import _lib
    full = _lib.calibrate_gather(num_nodes, num_nodes)
    half = _lib.calibrate_gather(num_nodes, min(65536, num_nodes))
    return {"unit": "node fetches/s", "achieved": node_rate, "random": full,
            "half_in_8mb": half, "frac_random": node_rate / full,
            "frac_half_in_8mb": node_rate / half,
            "source": "hz_calibrate_gather: dependent random 64-byte reads of 128-byte "
                      "records, four 16-byte loads per lane, 4 wavefronts per SIMD, 60% of "
                      "lanes per step"}


def end_to_end_leg(dom_in, opts, reps=3):
    """The drop-in call as a caller makes it: horizon_svf_comp on host
    arrays (FP64 prologue, upload, LBVH, order, trace, download), wall time
    of the C-ABI call in a process whose HIP runtime is already initialised.
    The first call allocates the device buffers (and, with the default
    option cache = 1, keeps them); value is the median of the reps calls
    after it, which reuse them.  With opts["prologue"] == "device" the
    prologue runs on the GPU (its time then includes the upload of the raw
    ICON arrays) and nothing is kept."""
    import numpy as np
    from horayzon_extpar_amd import _lib
    from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py
    _lib.lib().hz_cache_release()
    walls, stats = [], []
    for _ in range(reps + 1):
        st = {}
        t0 = time.perf_counter()
        horizon_svf_comp_py(*dom_in.args(), 24, 1, 40000.0, 0.2, 1, 2, stats=st, **opts)
        walls.append(time.perf_counter() - t0)
        stats.append(st)
    _lib.lib().hz_cache_release()
    first = walls[0]
    walls, stats = walls[1:], stats[1:]
    k = int(np.argsort(walls)[len(walls) // 2])
    st = stats[k]
    return {"value": dom_in.ncell / walls[k], "unit": "cells/s", "wall_ms": walls[k] * 1e3,
            "wall_ms_first_call": first * 1e3,
            "ms_prologue_host": st["t_prologue_ms"], "ms_alloc_upload": st["t_upload_ms"],
            "ms_bvh": st["t_bvh_ms"], "ms_order": st["t_order_ms"], "ms_trace": st["t_trace_ms"],
            "ms_download": st["t_download_ms"], "ms_total_library": st["t_total_ms"]}


def _timed_runs(dom, args, opts):
    """warmup + args.steps runs of dom with opts; (wall seconds per step,
    last stats, mean stats of the timed runs)."""
    import numpy as np
    for _ in range(max(1, args.warmup)):
        dom.run(24, 40000.0, 1, 2, **opts)
    sts = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sts.append(dom.run(24, 40000.0, 1, 2, **opts))
    step = (time.perf_counter() - t0) / args.steps
    mean = {k: float(np.mean([x[k] for x in sts])) for k in sts[0]}
    return step, sts[-1], mean


def _vs_walk(dom, h_walk, s_walk):
    import numpy as np
    h, s = dom.fetch()
    dh = np.abs(h - h_walk)
    return {"angles_differing_from_walk": int((dh > 1e-9).sum()),
            "cells_differing_from_walk": int(np.any(dh > 1e-9, axis=0).sum()),
            "max_abs_dh_deg": float(dh.max()), "max_abs_dsvf": float(np.abs(s - s_walk).max())}


def learned_leg(dom, args, h_walk, s_walk, dom_in, plain_step_s):
    """The lattice searches on the same resident domain, against the walk's
    results of the timed plain run:
    * learned (search_mode 2, certified): each step runs the LBVH build, the
      cell classification, cell ordering, raster + descriptors + predictor
      and the search (near cells walk, far cells start the lattice search at
      the prediction); its results equal the walk's;
    * climb (search_mode 1, certified): the same without the predictor (the
      lattice search climbs from 0 deg), so learned_vs_climb is what the
      network gains over the lattice search alone;
    * learned_uncertified: the plain lattice search everywhere, which misses
      the walk where rays pass through mesh holes (reported, not the form
      the speedup is quoted on)."""
    import numpy as np
    from horayzon_extpar_amd import learned

    path = args.weights or learned.DEFAULT_WEIGHTS
    if not os.path.exists(path):
        print(f"[bench] no predictor weights at {path}; learned leg skipped", file=sys.stderr)
        return None
    pred = learned.Predictor.from_file(path)
    dom.set_predictor(pred)
    n = dom_in.ncell
    step, st, m = _timed_runs(dom, args, dict(search_mode=2))
    acc = _vs_walk(dom, h_walk, s_walk)
    P = dom.predictions()
    err = np.abs(P - h_walk.T)
    out = {"value": n / step, "unit": "cells/s", "ms_per_step": step * 1e3,
           "certified": True, "ms_predict": m["t_predict_ms"], "ms_trace": m["t_trace_ms"],
           "ms_classify": m["t_certify_ms"], "ms_near_walk": m["t_cert_walk_ms"],
           "ms_border_team_walk": m["t_border_walk_ms"],
           "ms_far_lattice": m["t_cert_lattice_ms"],
           "cells_walked": int(st["cert_walk_cells"]), "rays": st["num_rays"],
           "rays_cast_ahead_unused": st["spec_rays"],
           "rays_per_cell_sector": st["num_rays"] / (n * 24),
           "speedup_vs_plain": plain_step_s / step, **acc,
           "prediction_mae_deg": float(err.mean()),
           "weights": os.path.relpath(path, ROOT)}
    print(f"[bench] learned (certified): {step * 1e3:.2f} ms/step, trace "
          f"{m['t_trace_ms']:.2f} ms, {acc['angles_differing_from_walk']} angles differ",
          file=sys.stderr, flush=True)
    step_u, st_u, m_u = _timed_runs(dom, args, dict(search_mode=2, certify=0))
    out["learned_uncertified"] = {"ms_per_step": step_u * 1e3, "ms_trace": m_u["t_trace_ms"],
                                  "rays": st_u["num_rays"],
                                  "speedup_vs_plain": plain_step_s / step_u,
                                  **_vs_walk(dom, h_walk, s_walk)}
    dom.set_predictor(None)
    step_c, st_c, m_c = _timed_runs(dom, args, dict(search_mode=1))
    out["climb"] = {"value": n / step_c, "unit": "cells/s", "ms_per_step": step_c * 1e3,
                    "certified": True, "ms_trace": m_c["t_trace_ms"],
                    "ms_near_walk": m_c["t_cert_walk_ms"],
                    "ms_far_lattice": m_c["t_cert_lattice_ms"],
                    "cells_walked": int(st_c["cert_walk_cells"]), "rays": st_c["num_rays"],
                    "speedup_vs_plain": plain_step_s / step_c,
                    **_vs_walk(dom, h_walk, s_walk)}
    out["learned_vs_climb_speedup"] = step_c / step
    print(f"[bench] climb (certified): {step_c * 1e3:.2f} ms/step; learned/climb "
          f"{step_c / step:.3f}x", file=sys.stderr, flush=True)
    return out


def batch_leg(args):
    """Config 4: 64 seeded domains (5k..400k cells, 2 / 1 / 0.5 km, nhori
    24 or 48).  Per nhori group: one launch per domain (unbatched) against
    one persistent launch over the group, replayed from a HIP graph with
    autotuned leaf size, refill threshold and queue chunk.  Trace time only
    (HIP events); the per-domain LBVH builds run before the launch."""
    import numpy as np
    from horayzon_extpar_amd import batch as hb
    from horayzon_extpar_amd import synthetic

    items = synthetic.config4()
    res, t_prep = hb.make_resident(items)
    out = {"domains": len(items), "cells": int(sum(d.ncell for d, _ in items)),
           "ms_prologue_upload_all": t_prep * 1e3, "groups": []}
    t_sep = t_bat = 0.0
    equal = True
    for nh, doms in sorted(hb.group_by_nhori(res).items()):
        for d in doms:
            d.run(nh)
        sep_ms = hb.separate_trace_ms(doms, nh)
        sep_out = [d.fetch() for d in doms]
        with hb.Batch(doms) as b:
            best, _, _ = hb.autotune(b, nh)
            st = [b.run(nh, use_graph=True, **best) for _ in range(max(1, args.steps))]
            got = b.fetch()
        bat_ms = float(np.mean([x["t_trace_ms"] for x in st]))
        for (h0, s0), (h1, s1) in zip(sep_out, got):
            equal &= bool(np.array_equal(h0, h1) and np.array_equal(s0, s1))
        ncell = int(sum(d.ncell for d in doms))
        out["groups"].append({"num_hori": nh, "domains": len(doms), "cells": ncell,
                              "ms_separate_launches": sep_ms, "ms_batched_graph": bat_ms,
                              "ms_lbvh_builds": float(np.mean([x["t_bvh_ms"] for x in st])),
                              "autotuned": best})
        t_sep += sep_ms
        t_bat += bat_ms
    for d, _ in res:
        d.close()
    out["cells_per_s_batched"] = out["cells"] / (t_bat * 1e-3)
    out["cells_per_s_separate"] = out["cells"] / (t_sep * 1e-3)
    out["speedup_batched_vs_separate"] = t_sep / t_bat
    out["bitwise_equal_to_separate"] = equal
    return out


def calibrate_copy(mib):
    """One device-to-device copy of mib MiB through torch (reads and writes
    exactly mib MiB each)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"[bench] calibration copy: {mib} MiB read + {mib} MiB written", file=sys.stderr,
          flush=True)
    del a, b


def calibrate_gather(mib):
    """One index_select of the 64-byte rows of a mib MiB float32 table in a
    random permutation: reads every row once (plus 8 bytes of index per
    row) and writes mib MiB."""
    import torch
    rows = mib * (1 << 20) // 64
    table = torch.ones((rows, 16), dtype=torch.float32, device="cuda")
    idx = torch.randperm(rows, device="cuda")
    torch.cuda.synchronize()
    out = table.index_select(0, idx)
    torch.cuda.synchronize()
    print(f"[bench] calibration gather: {rows} rows of 64 B in random order ({mib} MiB + "
          f"{rows * 8 >> 20} MiB of index read, {mib} MiB written)", file=sys.stderr, flush=True)
    del table, idx, out


def profiled_traffic(args):
    """(HBM bytes per launch on the copy calibration, on the gather
    calibration, source) of the default walk kernel on C2, from the
    committed rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE, FETCH scaled
    by a streaming 1 GiB copy or by a random 64-byte gather of 1 GiB); None
    for any other configuration."""
    default = (args.config == "c2" and args.mode == "walk" and args.kernel == 0
               and args.width in (0, 2) and args.leaf in (0, 1) and args.block in (0, 128)
               and args.morton == 0)
    path = os.path.join(ROOT, "profiles", "k_trace_c2_pmc.md")
    if not default or not os.path.exists(path):
        return None
    copy = gather = None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            copy = int(line.split()[1])
        elif line.startswith("traffic_bytes_per_launch_gather_scaled:"):
            gather = int(line.split()[1])
    if copy is None:
        return None
    return (copy, gather,
            "profiles/k_trace_c2_pmc.md: rocprofv3 FETCH_SIZE (x 2.0 on a 1 GiB streaming "
            "copy; x 0.563 on a random 64-byte gather) + WRITE_SIZE x 1.0, "
            "k_trace_pt<0, false, 128, 0, false, false, false>")


def accuracy_leg(dom_in, h, s, h_orc, s_orc, threads, nsample=48):
    """The accuracy gate of the run (walk mode): the device horizons against
    the oracle's on every cell the CPU baseline traced (bit-for-bit parity
    expected), and against the FP64 half-plane true horizon (every triangle
    that can meet the azimuth's half-plane) on nsample seeded cells: on angles
    that are neither grazing (within 1e-3 deg of a lattice step) nor on a
    search segment passing within two edges (2 km) of a mesh-boundary vertex
    (where the walk can see through a hole), the walk must return the lattice
    value below the true horizon, 0.25 + 0.5 floor(h / 0.5) deg, or 0;
    grazing and border angles are counted apart."""
    import numpy as np
    from horayzon_extpar_amd import diagnostics
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    done = ~np.isnan(s_orc)
    dh = np.abs(h[:, done] - h_orc[:, done])
    out = {"oracle_cells_compared": int(done.sum()),
           "oracle_angles_differing": int((dh > 0).sum()),
           "oracle_max_abs_dh_deg": float(dh.max()) if dh.size else None,
           "oracle_max_abs_dsvf": float(np.abs(s[done] - s_orc[done]).max()) if done.any()
           else None}
    rng = np.random.default_rng(11)
    cells = np.sort(rng.choice(dom_in.ncell, nsample, replace=False))
    t0 = time.perf_counter()
    th = np.degrees(orc.true_horizon(dom_in, cells, num_hori=24, grid_type=1,
                                     nthreads=threads or host_cpus()["threads"]))
    exp = np.where(th > 0, 0.25 + 0.5 * np.floor(th / 0.5), 0.0)
    grazing = np.abs(th / 0.5 - np.round(th / 0.5)) * 0.5 <= 1e-3
    border = diagnostics.border_azimuths(dom_in.clon, dom_in.clat, dom_in.vlon, dom_in.vlat,
                                         dom_in.cells_of_vertex, cells, np.arange(24) * 15.0,
                                         40000.0, 2000.0) & ~grazing
    clear = ~grazing & ~border
    got = h[:, cells].T
    agree = np.abs(got - exp) < 1e-9
    err = np.abs(got - np.maximum(th, 0.0))[clear]
    out["fp64_brute_force"] = {
        "cells": int(nsample), "clear_angles": int(clear.sum()),
        "clear_angles_agreeing": int(agree[clear].sum()),
        "grazing_angles": int(grazing.sum()),
        "grazing_angles_agreeing": int(agree[grazing].sum()),
        "border_angles": int(border.sum()),
        "border_angles_agreeing": int(agree[border].sum()),
        "max_abs_err_vs_true_deg": float(err.max()) if err.size else None,
        "seconds": time.perf_counter() - t0}
    return out


def host_cpus():
    """The host CPUs this process may run on: os.cpu_count(), the affinity
    mask, the cgroup CPU quota (cpu.max, in CPUs) and the model name; the
    thread count is the affinity count, capped by the quota where one is
    set (threads beyond the quota only time-slice the same CPU time)."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "cgroup_quota_cpus": None, "model": None}
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            info["cgroup_quota_cpus"] = int(q) / int(per)
    except (OSError, ValueError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    t = info["affinity"]
    if info["cgroup_quota_cpus"]:
        t = max(1, min(t, int(info["cgroup_quota_cpus"])))
    info["threads"] = t
    return info


def cpu_baseline(dom_in, budget_s, threads):
    """The oracle (C++ port of the reference walk, FP32 CPU BVH, std::thread
    pool with dynamic chunks like TBB's auto-partitioning): the whole C2
    domain on every host CPU this process may use (time-budgeted), plus one
    thread on a block of cells at the domain centre.  The reference's
    Embree/TBB build is not available on the host (Embree is not
    installed), so the port stands in for it.  Returns the JSON object and
    the oracle's horizon / skyview of the whole-domain run (NaN where not
    reached)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    cpus = host_cpus()
    nt = threads or cpus["threads"]
    n = dom_in.ncell
    h_orc, s_orc, st = orc.horizon_svf(dom_in, grid_type=1, nthreads=nt, time_budget_s=budget_s)
    cells = st.cells_done
    b = n // 2
    _, _, s1 = orc.horizon_svf(dom_in, grid_type=1, nthreads=1, cell_begin=b, cell_end=b + 20000,
                               time_budget_s=budget_s / 3)
    return {"value": cells / st.t_trace, "unit": "cells/s", "cores": nt, "kind": "port",
            "sample": f"{cells} of {n} cells (whole domain in index order, {st.t_trace:.1f} s "
                      f"of tracing on {nt} threads); BVH build {st.t_setup:.1f} s excluded",
            "rays_per_s": st.num_rays / st.t_trace,
            "host": {"nproc": cpus["nproc"], "affinity_cpus": cpus["affinity"],
                     "cgroup_quota_cpus": cpus["cgroup_quota_cpus"], "cpu_model": cpus["model"],
                     "embree": "unavailable"},
            "single_thread": {"value": s1.cells_done / s1.t_trace, "unit": "cells/s",
                              "cores": 1, "sample": f"{s1.cells_done} cells at the domain "
                                                    f"centre, {s1.t_trace:.1f} s"}}, h_orc, s_orc


if __name__ == "__main__":
    main()
