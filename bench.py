"""Benchmark: SSA state transitions per second on one MI355X.

One step = one pass of the simulate/statistics path over one batch: all
trajectories of the workload simulated from their initial state to max_time,
with window (hashed), final-state and fixed-point statistics accumulated in
HBM.  The model tables and statistics storage are resident on the device when
the timed region starts; the timed region covers K steps (statistics storage
cleared, persistent simulate kernel, counters read back), bracketed by
torch.cuda.synchronize().

Default workload ``plain_100``: 100-node random network from seed 1 (4 state
words, in-degree 1-4, AND/OR/NOT formulas), 50 internal nodes so the windows
key on 50 observed bits (hashed histogram), rates lognormal(0, 1), 1e6
trajectories, max_time 100, time_tick 1, continuous time.

Prints one JSON line.  Extra fields: ``roofline`` (the simulate kernel against
HBM with the reference's record traffic of 8W+16 bytes per transition),
``roofline_valu`` (against the VALU issue rate with the per-transition op
count 2*literals + 4N + 5*n_obs + 40), ``cpu_baseline`` (the host oracle with
compiled formulas on this machine's cores, bounded sample), and
``host_inclusive`` (one step that also copies all statistics to the host).
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
                 "(hashed window histogram), lognormal(0,1) rates, 1e6 trajectories, max_time 100, tick 1, continuous",
    "plain_100_discrete": "plain_100 network in discrete time",
    "plain_small": "10-node random network (seed 0), all observed, 1e4 trajectories, max_time 5, tick 0.2",
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
                      f"{res['seconds']:.1f} s, formulas compiled with g++ -O3 -march=native)"}


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
    transitions, kernel_ms = 0, 0.0
    for _ in range(a.steps):
        r = prog.simulate(n_trajectories=n, device_only=True)
        ri = r.info
        transitions += ri.transitions
        kernel_ms += ri.kernel_ms
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per_step = transitions / a.steps
    kernel_s = kernel_ms / a.steps / 1e3

    # the same step with every statistic copied back to the host (PCIe + host sort)
    t1 = time.perf_counter()
    full = prog.simulate(n_trajectories=n)
    host_s = time.perf_counter() - t1
    fi = full.info

    W, N, nobs = info.n_words, info.n_nodes, info.n_observed
    bytes_per = 8 * W + 16
    ops_per = 2 * info.total_literals + 4 * N + 5 * nobs + 40
    achieved = bytes_per * per_step / kernel_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{a.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")

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
                     "kernel": "simulate", "kernel_ms": kernel_s * 1e3,
                     "bytes_per_transition": bytes_per, "note": "effective: the fused kernel writes no records"},
        "roofline_valu": {"achieved": ops_per * per_step / kernel_s / 1e12, "peak": VALU_PEAK_TOPS,
                          "unit": "T lane-ops/s", "ops_per_transition": ops_per,
                          "frac": ops_per * per_step / kernel_s / 1e12 / VALU_PEAK_TOPS},
        "host_inclusive": {"value": fi.transitions / host_s, "unit": "transitions/s", "seconds": host_s,
                           "window_cells": fi.n_window_cells},
        "compile_seconds": prog.compile_seconds,
    }
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(bnd, cfg, a.cpu_seconds, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
