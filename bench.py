#!/usr/bin/env python3
"""Throughput of the Lagrange + ADI-remap hydro cycle on one MI355X.

One step = one iteration of the time loop (Mahyco.config:23-36) over the whole mesh of
the 3D bi-material shock tube: [0,1]^3, 256x128x128 hex cells, env A (x<0.5) rho=1 p=1
gamma=1.4, env B rho=0.125 p=0.1 gamma=1.67, superbee limiter, order 2, CFL 0.3, FP64.
The state is resident in HBM before the timed region starts; the time step stays on the
device, so the timed steps run without host round trips.

Prints one JSON line: metric, value (FP64 cell-cycles/s), roofline of the dominant kernel
(algorithmic bytes per launch / its average HIP-event duration on the library stream),
the cycle-level HBM fraction, and the CPU baseline (the oracle on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_words(nc, nn, m, d):
    """Compulsory FP64 words per cycle of the minimal fused schedule:
    cells W_c = 2(6m+4) + 2(d+1)(3m+1), nodes W_n = 2(2d+1) + 2(d+1)(d+2) + d(d+2)."""
    wc = 2 * (6 * m + 4) + 2 * (d + 1) * (3 * m + 1)
    wn = 2 * (2 * d + 1) + 2 * (d + 1) * (d + 2) + d * (d + 2)
    return nc * wc + nn * wn


def kernel_bytes(kind, nc, nn, m, d):
    """Algorithmic bytes of one launch of each kernel (every field read or written once;
    neighbour reuse is free; the Eulerian grid is rebuilt from indices)."""
    nv = 3 * m + 1
    if kind.startswith("k_sweep_"):
        kind = kind[:-2]  # same bytes for every direction
    words = {
        # read u, x_prev (d each) and M, p, q per cell; write u, x_new (d each) and m_node
        "k_node_lag": nn * (4 * d + 1) + nc * 3,
        # read x_new (d) and M, rho, c, p, e per cell; write U (3m+1), V, L
        "k_cell_lag": nn * d + nc * (5 + nv + 2),
        # read U (3m+1) and the node coordinates / velocity along the sweep; write U, mass fluxes (m)
        "k_sweep_cell": nc * (2 * nv + m) + nn * (d + 1),
        # read and write the dual state (d+1) and phi (d); read x along the sweep and the mass fluxes
        "k_sweep_node": nn * (2 * (d + 1) + 2 * d + 1) + nc * m,
        # read U (3m+1), V_E, L, u (d per node); write rho, e, p, c, q, M, f
        "k_finalize": nc * (nv + 2 + 7) + nn * d,
        "k_node_velocity": nn * (d + 1 + d),
        "k_dt": 0,
    }
    return 8.0 * words[kind]


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` at this mesh from the FETCH_SIZE + WRITE_SIZE profile
    summary kept in profiles/ (scaled by a plain copy's counters), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{args.workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("mesh") != [args.nx, args.ny, args.nz]:
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k["bytes_per_launch"]


def cpu_baseline(case_fn, cycles, threads, single_cycles=1):
    """The oracle (C++ restatement of the reference, OpenMP) on the host cores: `cycles`
    iterations on `threads` threads, then `single_cycles` on one thread."""
    from oracle.pyoracle import OracleCase
    case = case_fn()
    nc = case.nb_cells
    o = OracleCase(case, pow_mode=0, threads=threads)
    o.run(1)  # first iteration outside the timed region (page-in)
    t0 = time.perf_counter()
    o.run(cycles)
    dt = time.perf_counter() - t0
    o.close()
    out = {"value": nc * cycles / dt, "unit": "cell-cycles/s", "cores": threads, "kind": "port",
           "sample": f"{case.name}: {cycles} iterations after 1 untimed, {dt:.1f} s, OMP threads={threads}"}
    if single_cycles > 0:
        o = OracleCase(case, pow_mode=0, threads=1)
        o.run(1)
        t0 = time.perf_counter()
        o.run(single_cycles)
        d1 = time.perf_counter() - t0
        o.close()
        out["one_thread"] = {"value": nc * single_cycles / d1, "unit": "cell-cycles/s", "cores": 1,
                             "sample": f"{single_cycles} iteration(s) after 1 untimed, {d1:.1f} s"}
    return out


def cpu_baseline_runs(runs, threads, budget_s, sample_iters=10):
    """The oracle on the host cores solving the runs of an ensemble workload one after the
    other, on a bounded sample: runs = [(case, iterations to its end)] in the workload's order;
    for each, `sample_iters` iterations are timed after one untimed and scaled by its iteration
    count (the oracle and the GPU give the same bits, so the same count), until budget_s seconds
    have been timed.  Per-run solve times exclude case creation."""
    from oracle.pyoracle import OracleCase
    est, spent = [], 0.
    for case, total in runs:
        if total < 2:
            continue
        o = OracleCase(case, pow_mode=0, threads=threads)
        o.run(1)
        n = min(sample_iters, total - 1)
        t0 = time.perf_counter()
        o.run(n)
        dt = time.perf_counter() - t0
        o.close()
        est.append(dt / n * total)
        spent += dt
        if spent >= budget_s:
            break
    if not est:
        return None
    return {"value": len(est) / sum(est), "unit": "runs/s", "cores": threads, "kind": "port",
            "p50_s": _pct(est, 50), "p99_s": _pct(est, 99),
            "sample": f"the first {len(est)} runs back to back: {sample_iters} iterations of each timed after 1 "
                      f"untimed, scaled by its iteration count ({spent:.1f} s timed, OMP threads={threads}); "
                      "iterations only, no case creation"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["bisod3d", "sedov3d"], default="bisod3d",
                    help="bisod3d: 3D bi-material tube 256x128x128, superbee (default); sedov3d: 3D Sedov "
                         "blast 192^3 on [0,1.2]^3, one material, minmod")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--graph", type=int, default=6, help="cycles per hipGraph (0: eager launches)")
    ap.add_argument("--profile-steps", type=int, default=12)
    ap.add_argument("--cpu-cycles", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-single-cycles", type=int, default=4, help="iterations timed on one host thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="batched / serve: seconds of oracle iterations timed for the CPU baseline")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default kernel variants (the autotuner otherwise runs in the warmup)")
    ap.add_argument("--tune-cycles", type=int, default=12, help="timed iterations per autotuner candidate")
    ap.add_argument("--variant", type=int, default=None,
                    help="apply this kernel-variant code (as the autotuner reports it) instead of tuning")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (scale reference for HBM counters)")
    ap.add_argument("--mode", choices=["plain", "batched", "serve"], default="plain",
                    help="plain: single-mesh cycle (default); batched: 256 uneven 2D meshes, 300 "
                         "iterations each; serve: Poisson stream with continuous batching")
    ap.add_argument("--runs", type=int, default=256)
    ap.add_argument("--batched-only", action="store_true",
                    help="batched / serve: skip the back-to-back comparison (for profiling the ensemble's kernels alone)")
    ap.add_argument("--rate", type=float, default=16.0, help="serve: arrivals per second")
    ap.add_argument("--seed", type=int, default=20261202)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is not implemented)", file=sys.stderr)
        sys.exit(2)

    from mahyco_amd import cases
    from mahyco_amd import types as T
    from mahyco_amd.lib import GpuCase

    if args.mode == "batched":
        return bench_batched(args)
    if args.mode == "serve":
        return bench_serve(args)

    if args.workload == "sedov3d":
        n = args.nx or 192
        args.nx = args.ny = args.nz = n
        make = lambda: cases.config3_sedov3d(n)
    else:
        args.nx, args.ny, args.nz = args.nx or 256, args.ny or 128, args.nz or 128
        make = lambda: cases.config2_bisod3d(args.nx, args.ny, args.nz)
    case = make()
    nc, nn, m, d = case.nb_cells, case.nb_nodes, len(case.envs), case.dim
    print(f"[bench] {case.name}: cells={nc} nodes={nn} envs={m}", file=sys.stderr, flush=True)
    g = GpuCase(case)
    # The warmup is rounded up to whole graphs, so the graph the timed region replays is
    # captured and instantiated before t0 (a graph returns the coordinate parity and the
    # sweep rotation to where they started: its length is a multiple of lcm(2, d)).
    warm = args.warmup
    if args.graph > 0:
        period = 6 if d == 3 else 2
        gp = -(-args.graph // period) * period
        warm = max(gp, -(-args.warmup // gp) * gp)
    variant, tuned_ms = None, None
    if args.variant is not None:  # a kernel-variant code chosen by an earlier autotuner run
        g.set_variant(args.variant)
    elif not args.no_autotune:  # part of the warmup: the candidates run the case's own iterations
        variant, tuned_ms = g.autotune(args.tune_cycles)
        print(f"[bench] autotuned variant {variant} ({tuned_ms:.3f} ms/iteration)", file=sys.stderr, flush=True)
    g.queue(warm, graph=args.graph)
    st = g.synchronize()
    assert st == T.MH_OK, st
    captures0 = g.graph_captures
    t0 = time.perf_counter()
    g.queue(args.steps, graph=args.graph)
    st = g.synchronize()
    t1 = time.perf_counter()
    assert st == T.MH_OK, T.STATUS_NAMES.get(st, st)
    captures_timed = g.graph_captures - captures0
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = nc * args.steps / wall
    print(f"[bench] timed {args.steps} iterations: {ms_per_step:.3f} ms/iteration", file=sys.stderr, flush=True)

    # per-kernel HIP-event durations on the library stream over a further profiled region
    kern, roof = {}, None
    if args.profile_steps > 0:
        g.profile(True)
        g.queue(args.profile_steps)
        g.synchronize()
        prof = g.profile_read()
        g.profile(False)
        kern = {k: {"ms_avg": v[0] / v[1], "launches": v[1], "ms_total": v[0]} for k, v in prof.items() if v[1]}
        for k, v in kern.items():
            v["alg_GBs"] = kernel_bytes(k, nc, nn, m, d) / (v["ms_avg"] * 1e-3) / 1e9 if v["ms_avg"] > 0 else None
        dominant = max(kern, key=lambda k: kern[k]["ms_total"] * (0 if k == "k_dt" else 1))
        b = kernel_bytes(dominant, nc, nn, m, d)
        achieved = b / (kern[dominant]["ms_avg"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant, args), "kernel": dominant,
                "alg_bytes_per_launch": b}
    alg_cycle = 8.0 * algorithmic_words(nc, nn, m, d)
    cyc_gbs = alg_cycle / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": "FP64 cell-cycles/s + % HBM roofline; ensemble runs/s, p50/p99 latency",
        "value": value,
        "unit": "cell-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_run": warm,
        "autotuned_variant": variant,
        "variant": g.variant,
        "graph_captures_in_timed_region": captures_timed,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial condition of the configuration)",
        "config": {"workload": f"{args.workload}_lagrange_adi_remap", "mesh": [args.nx, args.ny, args.nz],
                   "cells": nc, "environments": m,
                   "limiter": "minmod" if args.workload == "sedov3d" else "superbee", "order": 2, "cfl": 0.3,
                   "graph_cycles": args.graph},
        "roofline": roof,
        "roofline_cycle": {"alg_bytes_per_cycle": alg_cycle, "achieved_GBs": cyc_gbs,
                           "frac": cyc_gbs / HBM_PEAK_GBS},
        "kernels": kern,
    }
    g.close()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(make, args.cpu_cycles, args.cpu_threads, args.cpu_single_cycles)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def copy_calibration(mib):
    """One device-to-device copy of known size through the HIP runtime (its blit kernel):
    HBM counters of a profiled run are scaled by what they read for it (mib MiB read,
    mib MiB written)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the library is linked with
    n = mib * (1 << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    ok = (hip.hipMemset(a, 1, C.c_size_t(n)) == 0 and hip.hipDeviceSynchronize() == 0
          and hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0 and hip.hipDeviceSynchronize() == 0)  # 3: device to device
    hip.hipFree(a)
    hip.hipFree(b)
    if not ok:
        raise RuntimeError("copy calibration failed")
    print(f"[bench] copy calibration: {mib} MiB read + {mib} MiB written", file=sys.stderr, flush=True)


def _pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v, dtype=float), q))


def bench_batched(args):
    """256 independent 2D bi-material tubes, sides log-uniform in [64, 512], 300 iterations each:
    one ensemble (persistent batched launches, hipGraph, autotuned grid) against the same 256
    runs solved back to back on the single-mesh path, both timed end to end (case creation,
    graph capture, iterations, release)."""
    from mahyco_amd import serve
    batch = serve.batch_cases(args.runs, seed=args.seed)
    cells = sum(c.nb_cells for c in batch)
    print(f"[bench] batched: {len(batch)} runs, {cells} cells", file=sys.stderr, flush=True)
    t_batch, iters, stats, done, bpc = serve.run_batch(batch, graph=True)
    assert all(i == batch[0].max_cycles for i in iters), "a run did not reach its horizon"
    if args.batched_only:
        print(json.dumps({"metric": "ensemble runs/s (batched)", "value": len(batch) / t_batch, "unit": "runs/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "config": {"workload": "batched_2d_bimaterial", "runs": len(batch), "cells": cells,
                                     "iterations": batch[0].max_cycles},
                          "batched_s": t_batch, "batched_phases_s": serve.run_batch.phases}))
        return
    print(f"[bench] batched {t_batch:.2f} s; back to back ...", file=sys.stderr, flush=True)
    t_b2b, done_b2b = serve.run_back_to_back(batch)
    out = {"metric": "ensemble runs/s (batched) vs back-to-back single-mesh solves",
           "value": len(batch) / t_batch, "unit": "runs/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f64", "data": f"synthetic, seed {args.seed}",
           "config": {"workload": "batched_2d_bimaterial", "runs": len(batch), "cells": cells,
                      "iterations": batch[0].max_cycles},
           "basis": "end to end for both: case creation, graph capture, every iteration, release",
           "batched_s": t_batch, "cell_cycles_per_s": cells * batch[0].max_cycles / t_batch,
           "tuned_blocks_per_cu": bpc,
           "back_to_back_s": t_b2b, "back_to_back_runs_per_s": len(batch) / t_b2b,
           "speedup_vs_back_to_back": t_b2b / t_batch,
           "p50_s": _pct(done, 50), "p99_s": _pct(done, 99),
           "back_to_back_p50_s": _pct(done_b2b, 50), "back_to_back_p99_s": _pct(done_b2b, 99),
           "batched_phases_s": serve.run_batch.phases,
           "back_to_back_iterate_s": serve.run_back_to_back.iterate_s,
           "speedup_iterations_only": serve.run_back_to_back.iterate_s / (serve.run_batch.phases["autotune_s"] +
                                                                         serve.run_batch.phases["iterate_s"])}
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline_runs(
        [(c, c.max_cycles) for c in batch], args.cpu_threads, args.cpu_budget)
    print(json.dumps(out))


def bench_serve(args):
    """Poisson stream of 2D (128^2..1024^2) and 3D (32^3..160^3) stiffened-gas runs to t=0.1 at
    an offered rate above what one GPU sustains, served by continuous batching (admission by HBM
    footprint into a device-resident ensemble) and, on the same stream, one run at a time on
    the single-mesh path."""
    from mahyco_amd import serve
    from mahyco_amd import types as T
    reqs = serve.serve_requests(args.runs, seed=args.seed, rate=args.rate)
    print(f"[bench] serve: {len(reqs)} requests at {args.rate}/s", file=sys.stderr, flush=True)
    recs = serve.serve(reqs, budget=96e9, max_slots=64)
    stats = serve.latency_stats(recs)
    if args.batched_only:
        print(json.dumps({"metric": "served runs/s, p50/p99 time-to-solution", "value": stats["runs_per_s"],
                          "unit": "runs/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "config": {"workload": "served_stream_2d3d_stiffened", "runs": len(reqs), "t_end": 0.1,
                                     "hbm_budget_GB": 96, "offered_rate_per_s": args.rate},
                          "p50_s": stats["p50_s"], "p99_s": stats["p99_s"],
                          "overlapped_admissions": sum(r.live_at_admission > 0 for r in recs)}))
        return
    print(f"[bench] served {stats['runs_per_s']:.2f} runs/s; back to back ...", file=sys.stderr, flush=True)
    b2b = serve.latency_stats(serve.serve_back_to_back(reqs))
    out = {"metric": "served runs/s, p50/p99 time-to-solution", "value": stats["runs_per_s"],
           "unit": "runs/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
           "data": f"synthetic Poisson stream, seed {args.seed}, rate {args.rate}/s",
           "config": {"workload": "served_stream_2d3d_stiffened", "runs": len(reqs), "t_end": 0.1,
                      "hbm_budget_GB": 96, "offered_rate_per_s": args.rate},
           "p50_s": stats["p50_s"], "p99_s": stats["p99_s"], "span_s": stats["span_s"],
           "overlapped_admissions": sum(r.live_at_admission > 0 for r in recs),
           "back_to_back_runs_per_s": b2b["runs_per_s"], "back_to_back_p50_s": b2b["p50_s"],
           "back_to_back_p99_s": b2b["p99_s"], "speedup_vs_back_to_back": stats["runs_per_s"] / b2b["runs_per_s"],
           "failed": [{"run": reqs[i][1].name, "status": T.STATUS_NAMES.get(r.status, r.status)}
                      for i, r in enumerate(recs) if r.iterations < 0]}
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline_runs(
        [(reqs[i][1], r.iterations) for i, r in enumerate(recs) if r.iterations > 0], args.cpu_threads,
        args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
