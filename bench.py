"""Benchmark: SSA state transitions per second on one MI355X.

One step = one pass of the simulate/statistics path over one batch: all
trajectories of the workload simulated from their initial state to max_time,
with window (hashed cells), final-state and fixed-point statistics
accumulated in HBM, compacted and sorted on the device, and copied to pinned
host memory.  The model tables are resident on the device when the timed
region starts; the timed region covers K steps from the first kernel to the
last statistic on the host, bracketed by torch.cuda.synchronize().  Steps are
issued with mbg_simulate_async, so step k's copy to the host overlaps step
k+1's kernels (at most two in flight).

Default workload ``plain_100``: 100-node random network from seed 1 (4 state
words, in-degree 1-4, AND/OR/NOT formulas), 50 internal nodes so the windows
key on 50 observed bits (hashed window cells), rates lognormal(0, 1), 1e6
trajectories, max_time 100, time_tick 1, continuous time.

Prints one JSON line.  Besides ``value`` (the host-inclusive rate above):
``device_only`` (the same steps with the statistics left in HBM),
``latency`` (one synchronous step), ``roofline`` (the simulate kernel against
the VALU peak, MI355X's published FP32 vector rate as 78.6 T lane-ops/s,
with the per-transition op count 2*literals + 4N + 5*n_obs + 40, which
binds; ``issue_probe`` inside it with this GPU's measured single-instruction
rates; ``hbm_effective`` inside it against HBM with the
reference's record traffic of 8W+16 bytes per transition), ``scale_1000``
(the 1000-node workload on the wave kernel, host-inclusive, with its own CPU
baseline), ``discrete``, ``sorted_output``, ``batched`` (256 variants in one
launch vs launched one by one), ``grad`` (forward vs likelihood-ratio
backward throughput), ``cpu_baseline`` (the host oracle with compiled
formulas on this machine's cores: median of 5 runs over bounded id prefixes)
and ``cpu_config1`` (the oracle on the 10-node CPU workload in full).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X's published FP32 vector rate, 157.3 TFLOP/s, as lane-ops (an FMA is 2 flops): 256 CUs x 128
# lane-ops per clock x 2.4 GHz = 78.6 T lane-ops/s.  On gfx950 the 2-operand 32-bit integer ops and
# v_fma_f32 issue at that rate (one wave64 instruction per SIMD every 2 cycles; measured 5.7-6.2e13
# lane-ops/s by lib/mbg_valu_issue at the clock the part holds under that load), v_bfe_u32,
# v_and_or_b32 and v_mul_hi_u32 at half of it.
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
VALU_PROBE = os.path.join(ROOT, "maboss", "gpu_amd", "lib", "mbg_valu_issue")
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


def cpu_threads():
    """Threads the baseline uses: OMP_NUM_THREADS when set (the CPU share of
    this job), else every CPU this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def host_cpu():
    """CPU model, sockets, cores per socket, threads per core (lscpu) and the CPUs this process sees."""
    info = {"logical_cpus": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(bnd, cfg, run_seconds, threads, min_ids=2000, runs=5):
    """The host oracle (formulas compiled with g++, OpenMP over trajectory ids,
    OMP_PROC_BIND=close, OMP_PLACES=cores) on prefixes of the same
    trajectories: one warm-up run, then the median of `runs` runs.
    Transitions/s is intensive, so the rates compare with the GPU's."""
    oracle = os.path.join(ROOT, "oracle", "build", "maboss_oracle")
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_NUM_THREADS=str(threads))

    def run(n):
        p = subprocess.run([oracle, "--compiled", "--time", "--threads", str(threads), "--n", str(n), bnd, cfg],
                           capture_output=True, text=True, timeout=900, env=env)
        if p.returncode != 0:
            raise RuntimeError(p.stderr)
        return json.loads(p.stderr.strip().splitlines()[-1])

    probe = run(min(2000, min_ids))  # warm-up, also sizes the runs
    n = int(min(1_000_000, max(min_ids, min(2000, min_ids) * run_seconds / max(probe["seconds"], 1e-3))))
    res = [run(n) for _ in range(runs)]
    rates = [r["transitions_per_s"] for r in res]
    med = statistics.median(rates)
    return {"value": med, "unit": "transitions/s", "cores": threads, "kind": "port",
            "runs": rates,
            "sample": f"median of {runs} runs over the first {n} trajectory ids of the same workload "
                      f"({res[0]['transitions']} transitions, ~{res[0]['seconds']:.1f} s each) after one warm-up run; "
                      f"formulas compiled with g++ -O3 -march=native; OMP_PROC_BIND=close OMP_PLACES=cores",
            "host": host_cpu()}


def pipelined_steps(prog, n, steps, **kw):
    """K host-inclusive steps, at most two in flight; returns (wall s, transitions, sim ms per step, results)."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prev, trans, sims = None, 0, []
    for _ in range(steps):
        cur = prog.simulate(n_trajectories=n, wait=False, **kw)
        if prev is not None:
            prev.wait()
            pi = prev.info
            trans += pi.transitions
            sims.append(pi.simulate_ms)
        prev = cur
    prev.wait()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    li = prev.info
    trans += li.transitions
    sims.append(li.simulate_ms)
    return wall, trans, sims, li


def device_only_steps(prog, n, steps, **kw):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    trans, sims, paths = 0, [], []
    for _ in range(steps):
        ri = prog.simulate(n_trajectories=n, device_only=True, **kw).info
        trans += ri.transitions
        sims.append(ri.simulate_ms)
        paths.append(ri.kernel_ms)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, trans, sims, paths


def bench_discrete(mg, netgen, wdir, steps):
    """The same 100-node network in discrete time (one record per step at window round(t / tick))."""
    bnd, cfg = netgen.write_workload("plain_100_discrete", wdir)
    model = mg.Model.load(bnd, cfg)
    prog = model.compile()
    n = model.info.sample_count
    pipelined_steps(prog, n, 1)
    wall, trans, sims, li = pipelined_steps(prog, n, steps)
    return {"workload": "plain_100_discrete: the plain_100 network with discrete_time = 1, 1e6 trajectories",
            "value": trans / wall, "unit": "transitions/s", "ms_per_step": wall / steps * 1e3,
            "kernel_ms": sum(sims) / len(sims), "window_mode": li.window_mode,
            "window_cells": li.n_window_cells, "transitions_per_step": trans / steps}


def bench_scale(mg, netgen, wdir, steps, cpu, cpu_seconds):
    """The 1000-node workload (wave per trajectory), host-inclusive, against its own CPU baseline."""
    bnd, cfg = netgen.write_workload("scale_1000", wdir)
    model = mg.Model.load(bnd, cfg)
    prog = model.compile()
    info = model.info
    n = info.sample_count
    pipelined_steps(prog, n, 1)
    wall, trans, sims, li = pipelined_steps(prog, n, steps)
    rate = trans / wall
    kernel_s = sum(sims) / len(sims) / 1e3
    per = trans / steps
    out = {"workload": f"scale_1000: {WORKLOAD_TEXT['scale_1000']}", "value": rate, "unit": "transitions/s",
           "steps": steps, "ms_per_step": wall / steps * 1e3, "kernel": f"simulate_wave_n{_rows(info.n_nodes)}",
           "kernel_ms": kernel_s * 1e3, "kernel_transitions_per_s": per / kernel_s,
           "transitions_per_step": per, "window_cells": li.n_window_cells,
           "hbm_effective_frac": (8 * info.n_words + 16) * per / kernel_s / 1e9 / HBM_PEAK_GBS,
           "ops_per_transition": valu_ops(info), "valu_frac": valu_frac(info, per / kernel_s)}
    if cpu:
        cb = cpu_baseline(bnd, cfg, cpu_seconds, cpu_threads(), min_ids=20000)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = rate / cb["value"]
        out["target_100x_met"] = rate / cb["value"] >= 100.0
    return out


def valu_ops(info):
    """Per-transition VALU lane-ops of the path: 2 per literal (logic), 4 per node (rate select, enable
    bit, running sum), 5 per observed node (index, entropy) and 40 for the step's random numbers, time
    and flip."""
    return 2 * info.total_literals + 4 * info.n_nodes + 5 * info.n_observed + 40


def valu_frac(info, transitions_per_s):
    return valu_ops(info) * transitions_per_s / (VALU_PEAK_TOPS * 1e12)


def valu_issue_probe():
    """Lane-instruction rates of single VALU instructions on this GPU (lib/mbg_valu_issue), or None."""
    if not os.path.exists(VALU_PROBE):
        return None
    p = subprocess.run([VALU_PROBE], capture_output=True, text=True, timeout=120)
    if p.returncode != 0:
        return None
    rows = [json.loads(line) for line in p.stdout.splitlines() if line.startswith("{")]
    return {f"{r['op']}@{r['waves_per_simd']}w": r["lane_instr_per_s"] for r in rows if "op" in r}


def _rows(n_nodes):
    rows = 1
    while rows < (n_nodes + 63) // 64:
        rows *= 2
    return rows


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
           "lds_cells": info.lds_cells, "tuned_from_cache": bool(info.tuned_cached),
           "graph": bool(info.used_graph), "ops_per_transition": valu_ops(model.info),
           "valu_frac": valu_frac(model.info, trans / dev)}
    t0 = time.perf_counter()
    _, hinfo = prog.simulate_batched(variants)
    one["host_inclusive_transitions_per_s"] = hinfo.transitions / (time.perf_counter() - t0)
    # one variant per launch, in chunks of 64 ids (one trajectory per lane): the most waves a single
    # variant of 1e3-1e5 trajectories can keep busy, whatever chunk size the whole batch was tuned to
    sdev, strans = 0.0, 0
    t0 = time.perf_counter()
    for v in variants:
        _, vi = prog.simulate_batched([v], device_only=True, use_graph=False, block_size=info.block_size,
                                      blocks_per_cu=info.blocks_per_cu, traj_per_lane=1,
                                      lds_cells=info.lds_cells or -1)
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
            "backward_over_forward_throughput": (bt / bwd) / (ft / fwd), "accumulator": mgrad.ACCUMULATOR,
            "ops_per_transition": valu_ops(info), "forward_valu_frac": valu_frac(info, ft / fwd),
            "backward_valu_frac_at_forward_ops": valu_frac(info, bt / bwd)}


def _step(what, t0=[None]):
    """One progress line on stderr as each block of the run ends."""
    now = time.perf_counter()
    if t0[0] is None:
        t0[0] = now
    print(f"bench: {what} ({now - t0[0]:.1f} s)", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain_100", choices=sorted(WORKLOAD_TEXT))
    ap.add_argument("--trajectories", type=int, default=-1, help="override the workload's trajectory count")
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="seconds of one CPU baseline run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the batched, gradient, discrete and scale blocks")
    ap.add_argument("--no-sorted-output", action="store_true", help="skip the sorted-pairs step")
    ap.add_argument("--only", choices=["discrete", "scale_1000", "batched", "grad"], default=None,
                    help="run only this form's block and print it (for profiling its kernels)")
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
    if a.only:
        steps = max(1, min(a.steps, 3))
        block = {"discrete": lambda: bench_discrete(mg, netgen, wdir, steps),
                 "scale_1000": lambda: bench_scale(mg, netgen, wdir, steps, not a.no_cpu_baseline, a.cpu_seconds),
                 "batched": lambda: bench_batched(mg, netgen, wdir, steps),
                 "grad": lambda: bench_grad(mg, netgen, wdir, steps)}[a.only]()
        print(json.dumps({"only": a.only, a.only: block}), flush=True)
        return
    bnd, cfg = netgen.write_workload(a.workload, wdir)
    model = mg.Model.load(bnd, cfg)
    info = model.info
    prog = model.compile()
    n = a.trajectories if a.trajectories > 0 else info.sample_count

    _step("model compiled")
    if a.warmup:
        pipelined_steps(prog, n, a.warmup)
    # the headline: K host-inclusive steps
    elapsed, transitions, sims, li = pipelined_steps(prog, n, a.steps)
    per_step = transitions / a.steps
    kernel_s = sum(sims) / len(sims) / 1e3  # the simulate kernel alone (HIP events on its stream)
    # the same steps with the statistics left in HBM
    d_wall, d_trans, d_sims, d_paths = device_only_steps(prog, n, a.steps)
    # one synchronous step: the latency of a single result
    t1 = time.perf_counter()
    full = prog.simulate(n_trajectories=n)
    fi = full.info
    lat_s = time.perf_counter() - t1
    del full
    _step("headline, device-only and latency steps")

    sorted_out = None
    if li.window_mode == 1 and not a.no_sorted_output:
        pipelined_steps(prog, n, 1, window_mode="sort")
        s_wall, s_trans, s_sims, s_li = pipelined_steps(prog, n, max(1, min(a.steps, 10)), window_mode="sort")
        sorted_out = {"window_mode": "sort", "value": s_trans / s_wall, "unit": "transitions/s",
                      "simulate_ms": sum(s_sims) / len(s_sims), "window_pairs": s_li.window_pairs,
                      "note": "window cells appended as (key, value) pairs and sort-reduced instead of hashed"}

    W, N, nobs = info.n_words, info.n_nodes, info.n_observed
    bytes_per = 8 * W + 16
    ops_per = valu_ops(info)
    kname = f"simulate_wave_n{_rows(N)}" if li.kernel == 1 else "simulate"
    rate_k = per_step / kernel_s
    bound_rate = min(HBM_PEAK_GBS * 1e9 / bytes_per, VALU_PEAK_TOPS * 1e12 / ops_per)
    valu_bound = VALU_PEAK_TOPS * 1e12 / ops_per <= HBM_PEAK_GBS * 1e9 / bytes_per
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{kname}_{a.workload}_traffic.md")
    if os.path.exists(tpath):
        with open(tpath) as f:
            for line in f:
                if line.startswith("bytes_per_launch:"):
                    traffic = float(line.split(":", 1)[1])

    valu_ach = ops_per * rate_k / 1e12
    hbm_ach = bytes_per * rate_k / 1e9
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
                   "form": "plain", "transitions_per_step": per_step, "window_cells_per_step": li.n_window_cells,
                   "window_record_bytes": li.window_record_bytes, "window_escapes_per_step": li.window_escapes,
                   "timed": "first kernel to the last statistic in pinned host memory; steps pipelined "
                            "(a step's copy overlaps the next step's kernels)"},
        "roofline": {"bound": "valu" if valu_bound else "hbm",
                     "achieved": valu_ach if valu_bound else hbm_ach,
                     "peak": VALU_PEAK_TOPS if valu_bound else HBM_PEAK_GBS,
                     "unit": "T lane-ops/s" if valu_bound else "GB/s",
                     "frac": rate_k / bound_rate, "traffic": traffic,
                     "kernel": kname, "kernel_ms": kernel_s * 1e3, "transitions_per_launch": per_step,
                     "ops_per_transition": ops_per, "achieved_transitions_per_s": rate_k,
                     "bound_transitions_per_s": bound_rate,
                     "peak_note": "published FP32 vector rate (157.3 TFLOP/s) as lane-ops; issue_probe holds "
                                  "measured single-instruction rates on this GPU",
                     "hbm_effective": {"achieved": hbm_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                       "frac": hbm_ach / HBM_PEAK_GBS, "bytes_per_transition": bytes_per,
                                       "note": "effective: the fused kernel writes no records; algorithmic "
                                               "bytes are the reference's record write + read, 8W+16"}},
        "device_only": {"value": d_trans / d_wall, "unit": "transitions/s", "ms_per_step": d_wall / a.steps * 1e3,
                        "simulate_ms": sum(d_sims) / len(d_sims), "device_path_ms": sum(d_paths) / len(d_paths),
                        "note": "statistics left in HBM (MBG_SIM_DEVICE_ONLY)"},
        "latency": {"seconds": lat_s, "transitions_per_s": fi.transitions / lat_s, "device_path_ms": fi.kernel_ms,
                    "window_cells": fi.n_window_cells,
                    "note": "one synchronous mbg_simulate: kernel, compaction, sort, copy to pinned host memory"},
        "host_over_device": (d_trans / d_wall) / (transitions / elapsed),
        "sorted_output": sorted_out,
        "compile_seconds": prog.compile_seconds,
    }
    if not a.no_forms and a.workload == "plain_100":
        out["discrete"] = bench_discrete(mg, netgen, wdir, max(1, min(a.steps, 10)))
        _step("discrete")
    if not a.no_forms:
        out["scale_1000"] = bench_scale(mg, netgen, wdir, 2, not a.no_cpu_baseline, max(a.cpu_seconds, 4.0))
        _step("scale_1000")
        out["batched"] = bench_batched(mg, netgen, wdir, max(1, min(a.steps, 3)))
        _step("batched")
        out["grad"] = bench_grad(mg, netgen, wdir, max(1, min(a.steps, 3)))
        _step("grad")
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(bnd, cfg, a.cpu_seconds, cpu_threads())
        _step("cpu baseline")
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        b1, c1 = netgen.write_workload("plain_small", wdir)
        cb1 = cpu_baseline(b1, c1, 0.0, cpu_threads(), min_ids=10000)
        cb1["sample"] = "config 1 in full: " + cb1["sample"]
        out["cpu_config1"] = cb1
    out["roofline"]["issue_probe"] = valu_issue_probe()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
