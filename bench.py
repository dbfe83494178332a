"""Benchmark: SSA state transitions per second on one MI355X.

One step = one pass of the simulate/statistics path over one batch: all
trajectories of the workload simulated from their initial state to max_time,
with window (hashed cells), final-state and fixed-point statistics accumulated in
HBM.  The model tables and statistics storage are resident on the device when
the timed region starts; the timed region covers K steps (statistics storage
cleared, persistent simulate kernel, counters read back), bracketed by
torch.cuda.synchronize().

Default workload ``plain_100``: 100-node random network from seed 1 (4 state
words, in-degree 1-4, AND/OR/NOT formulas), 50 internal nodes so the windows
key on 50 observed bits (hashed window cells), rates lognormal(0, 1), 1e6
trajectories, max_time 100, time_tick 1, continuous time.

Prints one JSON line.  Extra fields: ``batched`` (256 variants in one launch
vs launched one by one), ``grad`` (forward vs likelihood-ratio backward
throughput), ``roofline`` (the simulate kernel against
HBM with the reference's record traffic of 8W+16 bytes per transition),
``roofline_valu`` (against the VALU issue rate with the per-transition op
count 2*literals + 4N + 5*n_obs + 40), ``roofline_rate`` (the kernel's
transitions/s against min(HBM/B, VALU/C)), ``cpu_baseline`` (the host oracle with
compiled formulas on this machine's cores, bounded sample), and
``host_inclusive`` (one step that also copies all statistics to the host), and
``sorted_output`` (one step that leaves the window cells as a compact array sorted by
key in HBM instead of the hash table).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz = 39.3 T lane-ops/s
METRIC = "state transitions/sec on 1 MI355X (plain vs batched vs +grad) at matched stats"
WORKLOAD_TEXT = {
    "plain_100": "100-node random network (seed 1, 4 state words, in-degree 1-4, AND/OR/NOT), 50 observed nodes "
                 "(hashed window cells), lognormal(0,1) rates, 1e6 trajectories, max_time 100, tick 1, continuous",
    "plain_100_discrete": "plain_100 network in discrete time",
    "plain_small": "10-node random network (seed 0), all observed, 1e4 trajectories, max_time 5, tick 0.2",
    "scale_1000": "1000-node scale-free network (seed 2, 32 state words, out-degree exponent 2.5, in-degree <= 4, "
                  "threshold/AND-OR formulas), 20 observed nodes, Bernoulli(0.3) initial state, lognormal(0,1) "
                  "rates, 1e6 trajectories, max_time 50, tick 0.5, continuous (one wave per trajectory)",
    "batched_50": "50-node random network (seed 3), base network of the batched form",
    "grad_30": "30-node random network (seed 4), 1e6 trajectories, max_time 20, tick 0.5",
}


def _ensure_built():
    lib = os.path.join(ROOT, "maboss", "gpu_amd", "lib", "libmaboss_gpu.so")
    oracle = os.path.join(ROOT, "oracle", "build", "maboss_oracle")
    if not (os.path.exists(lib) and os.path.exists(oracle)):
        import __graft_entry__
        __graft_entry__.build()


def cpu_baseline(bnd, cfg, target_s, threads):
    """The host oracle (compiled formulas, OpenMP) on a bounded prefix of the
    same trajectories; transitions/s is intensive so the rates compare."""
    oracle = os.path.join(ROOT, "oracle", "build", "maboss_oracle")

    def run(n):
        p = subprocess.run([oracle, "--compiled", "--time", "--threads", str(threads), "--n", str(n), bnd, cfg],
                           capture_output=True, text=True, timeout=600)
        if p.returncode != 0:
            raise RuntimeError(p.stderr)
        return json.loads(p.stderr.strip().splitlines()[-1])

    probe = run(2000)
    n = int(min(1_000_000, max(2000, 2000 * target_s / max(probe["seconds"], 1e-3))))
    res = run(n)
    return {"value": res["transitions_per_s"], "unit": "transitions/s", "cores": threads, "kind": "port",
            "sample": f"first {n} trajectory ids of the same workload ({res['transitions']} transitions, "
                      f"{res['seconds']:.1f} s, formulas compiled with g++ -O3 -march=native)",
            "host": host_cpu()}


def host_cpu():
    """CPU model and logical CPU count of the machine the baseline ran on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def bench_discrete(mg, netgen, wdir, steps):
    """The same 100-node network in discrete time (one record per step at window round(t / tick))."""
    import torch
    bnd, cfg = netgen.write_workload("plain_100_discrete", wdir)
    model = mg.Model.load(bnd, cfg)
    prog = model.compile()
    n = model.info.sample_count
    prog.simulate(n_trajectories=n, device_only=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    trans, sim_ms = 0, 0.0
    for _ in range(steps):
        ri = prog.simulate(n_trajectories=n, device_only=True).info
        trans += ri.transitions
        sim_ms += ri.simulate_ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"workload": "plain_100_discrete: the plain_100 network with discrete_time = 1, 1e6 trajectories",
            "value": trans / wall, "unit": "transitions/s", "ms_per_step": wall / steps * 1e3,
            "kernel_ms": sim_ms / steps, "window_mode": ri.window_mode, "transitions_per_step": trans / steps}


def bench_batched(mg, netgen, wdir, steps):
    """Batched form: the 256 variants of the batched workload (rates x lognormal(0, 0.5), 0-3 mutants,
    1e3-1e5 trajectories, max_time 10-200) in one graph-captured launch, against the same variants
    launched one after another."""
    bnd, cfg = netgen.write_workload("batched_50", wdir)
    model = mg.Model.load(bnd, cfg)
    prog = model.compile()
    variants = netgen.batched_variants(list(model.param_values), model.info.n_nodes)
    total_n = sum(v["n_trajectories"] for v in variants)
    _, info = prog.simulate_batched(variants, device_only=True)  # autotune + graph capture
    dev, wall, trans = 0.0, 0.0, 0
    for _ in range(steps):
        t0 = time.perf_counter()
        _, info = prog.simulate_batched(variants, device_only=True)
        wall += time.perf_counter() - t0
        dev += info.kernel_ms / 1e3
        trans += info.transitions
    one = {"transitions_per_s": trans / dev, "device_ms": dev / steps * 1e3, "wall_ms": wall / steps * 1e3,
           "block": info.block_size, "blocks_per_cu": info.blocks_per_cu, "traj_per_lane": info.traj_per_lane,
           "graph": bool(info.used_graph)}
    sdev, swall, strans = 0.0, 0.0, 0
    t0 = time.perf_counter()
    for v in variants:
        _, vi = prog.simulate_batched([v], device_only=True, use_graph=False, block_size=info.block_size,
                                      blocks_per_cu=info.blocks_per_cu, traj_per_lane=info.traj_per_lane)
        sdev += vi.kernel_ms / 1e3
        strans += vi.transitions
    swall = time.perf_counter() - t0
    seq = {"transitions_per_s": strans / sdev, "device_ms": sdev * 1e3, "wall_ms": swall * 1e3}
    return {"workload": "batched_50: 256 variants, " + f"{total_n} trajectories", "one_launch": one,
            "sequential_256": seq, "speedup_device": seq["device_ms"] / one["device_ms"],
            "speedup_wall": seq["wall_ms"] / one["wall_ms"]}


def bench_grad(mg, netgen, wdir, steps):
    """Gradient form on grad_30 (1e6 trajectories): the forward window probabilities at run-time rates
    (the simulate kernel reading the rates from a table) and the likelihood-ratio backward over the
    same trajectories."""
    import torch
    from maboss.gpu_amd import grad as mgrad
    bnd, cfg = netgen.write_workload("grad_30", wdir)
    model = mg.Model.load(bnd, cfg)
    prog = model.compile()
    info = model.info
    n = info.sample_count
    rates = list(model.param_values)
    G = torch.randn(info.windows, 1 << info.n_observed, dtype=torch.float64,
                    generator=torch.Generator().manual_seed(0)).cuda()
    th = torch.tensor(rates, dtype=torch.float32).cuda()
    prog.simulate(n_trajectories=n, rates=rates, device_only=True)
    mgrad.backward(prog, G, rates=th, n_trajectories=n, info=True)
    fwd, bwd, ft, bt = 0.0, 0.0, 0, 0
    for _ in range(steps):
        fi = prog.simulate(n_trajectories=n, rates=rates, device_only=True).info
        fwd += fi.kernel_ms / 1e3
        ft += fi.transitions
        _, gi = mgrad.backward(prog, G, rates=th, n_trajectories=n, info=True)
        bwd += gi.kernel_ms / 1e3
        bt += gi.transitions
    return {"workload": "grad_30: 30 nodes, 60 rate parameters, 5 observed, 1e6 trajectories, max_time 20",
            "forward_transitions_per_s": ft / fwd, "backward_transitions_per_s": bt / bwd,
            "forward_ms": fwd / steps * 1e3, "backward_ms": bwd / steps * 1e3,
            "backward_over_forward_throughput": (bt / bwd) / (ft / fwd), "accumulator": "f64"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain_100", choices=sorted(WORKLOAD_TEXT))
    ap.add_argument("--trajectories", type=int, default=-1, help="override the workload's trajectory count")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the batched and gradient forms")
    ap.add_argument("--no-sorted-output", action="store_true", help="skip the sorted-pairs step")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run one plain device copy of this many MiB (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.copy_probe_mib:
        import torch
        nbytes = a.copy_probe_mib << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (the engine shards trajectory ids "
              f"across GPUs only by separate runs with first_id)", file=sys.stderr)
        sys.exit(2)

    _ensure_built()
    import torch

    import maboss.gpu_amd as mg
    from maboss.gpu_amd import netgen

    torch.cuda.set_device(0)
    wdir = tempfile.mkdtemp(prefix="mbg_bench_")
    bnd, cfg = netgen.write_workload(a.workload, wdir)
    model = mg.Model.load(bnd, cfg)
    info = model.info
    prog = model.compile()
    n = a.trajectories if a.trajectories > 0 else info.sample_count

    for _ in range(a.warmup):
        prog.simulate(n_trajectories=n, device_only=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    transitions, kernel_ms, sim_ms = 0, 0.0, 0.0
    for _ in range(a.steps):
        r = prog.simulate(n_trajectories=n, device_only=True)
        ri = r.info
        transitions += ri.transitions
        kernel_ms += ri.kernel_ms
        sim_ms += ri.simulate_ms
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per_step = transitions / a.steps
    kernel_s = sim_ms / a.steps / 1e3  # the simulate kernel alone (HIP events on its stream)
    path_s = kernel_ms / a.steps / 1e3  # simulate + tail expansion (+ window sort/reduce in the sorted mode)

    # the same step with every statistic copied back to the host (PCIe, cells in key order)
    t1 = time.perf_counter()
    full = prog.simulate(n_trajectories=n)
    host_s = time.perf_counter() - t1
    fi = full.info

    # the same step with the window cells left as a compact sorted array in HBM (sorted pairs)
    sorted_out = None
    if ri.window_mode == 1 and not a.no_sorted_output:
        prog.simulate(n_trajectories=n, device_only=True, window_mode="sort")
        torch.cuda.synchronize()
        ts = time.perf_counter()
        so = prog.simulate(n_trajectories=n, device_only=True, window_mode="sort").info
        torch.cuda.synchronize()
        sorted_out = {"window_mode": "sort", "value": so.transitions / (time.perf_counter() - ts),
                      "unit": "transitions/s", "device_path_ms": so.kernel_ms, "simulate_ms": so.simulate_ms,
                      "window_pairs": so.window_pairs}

    W, N, nobs = info.n_words, info.n_nodes, info.n_observed
    bytes_per = 8 * W + 16
    ops_per = 2 * info.total_literals + 4 * N + 5 * nobs + 40
    achieved = bytes_per * per_step / kernel_s / 1e9
    if ri.kernel == 1:  # wave per trajectory: rate rows = nodes per lane rounded up to a power of two
        rows = 1
        while rows < (N + 63) // 64:
            rows *= 2
        kname = f"simulate_wave_n{rows}"
    else:
        kname = "simulate"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{kname}_{a.workload}_traffic.md")
    if os.path.exists(tpath):
        with open(tpath) as f:
            for line in f:
                if line.startswith("bytes_per_launch:"):
                    traffic = float(line.split(":", 1)[1])

    out = {
        "metric": METRIC,
        "value": transitions / elapsed,
        "unit": "transitions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded network generator)",
        "config": {"workload": f"{a.workload}: {WORKLOAD_TEXT[a.workload]}", "trajectories": n,
                   "nodes": N, "state_words": W, "observed_nodes": nobs, "windows": info.windows,
                   "form": "plain", "transitions_per_step": per_step},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_ms": kernel_s * 1e3, "device_path_ms": path_s * 1e3,
                     "bytes_per_transition": bytes_per,
                     "note": "effective: the fused kernel writes no records; algorithmic bytes are the "
                             "reference's record write + read, 8W+16 per transition"},
        "roofline_valu": {"achieved": ops_per * per_step / kernel_s / 1e12, "peak": VALU_PEAK_TOPS,
                          "unit": "T lane-ops/s", "ops_per_transition": ops_per,
                          "frac": ops_per * per_step / kernel_s / 1e12 / VALU_PEAK_TOPS},
        # the combined bound R* = min(HBM / B, VALU / C) transitions/s and the kernel's share of it
        "roofline_rate": {"bound_transitions_per_s": min(HBM_PEAK_GBS * 1e9 / bytes_per, VALU_PEAK_TOPS * 1e12 / ops_per),
                          "kernel_transitions_per_s": per_step / kernel_s,
                          "frac": per_step / kernel_s / min(HBM_PEAK_GBS * 1e9 / bytes_per,
                                                            VALU_PEAK_TOPS * 1e12 / ops_per)},
        "sorted_output": sorted_out,
        "host_inclusive": {"value": fi.transitions / host_s, "unit": "transitions/s", "seconds": host_s,
                           "window_cells": fi.n_window_cells},
        "compile_seconds": prog.compile_seconds,
    }
    if not a.no_forms and a.workload == "plain_100":
        out["discrete"] = bench_discrete(mg, netgen, wdir, max(1, min(a.steps, 3)))
    if not a.no_forms:
        out["batched"] = bench_batched(mg, netgen, wdir, max(1, min(a.steps, 3)))
        out["grad"] = bench_grad(mg, netgen, wdir, max(1, min(a.steps, 3)))
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(bnd, cfg, a.cpu_seconds, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
