"""Batched form on the GPU: every variant of one launch against the oracle.

A variant (rate vector, mutant nodes, trajectory count, horizon) draws from
the stream tagged with its index, so variant v of a batch must equal, bit for
bit, the oracle run with ``--variant v`` on a .bnd/.cfg pair edited to hold
that variant's rates, horizon and mutations (a node forced on has its
rate_down parameter set to 0 and istate 1; forced off: rate_up 0, istate 0).
"""
import os
import re

import numpy as np
import pytest

import maboss.gpu_amd as mg
from maboss.gpu_amd import netgen
from conftest import model_paths, read_raw, run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def b50(tmp_path_factory):
    d = tmp_path_factory.mktemp("b50")
    return netgen.write_workload("batched_50", str(d))


_programs = {}


def program(bnd, cfg):
    if (bnd, cfg) not in _programs:
        _programs[(bnd, cfg)] = mg.Model.load(bnd, cfg).compile()
    return _programs[(bnd, cfg)]


def f32_text(x):
    return "%.9g" % float(np.float32(x))


def edited_cfg(cfg_text, params=None, max_time=None, istate=None):
    """cfg text with $param values, max_time and node istates replaced."""
    out = cfg_text
    for name, val in (params or {}).items():
        name = name.lstrip("$")
        out, n = re.subn(rf"^\${re.escape(name)}\s*=.*;$", f"${name} = {f32_text(val)};", out, flags=re.M)
        assert n == 1, name
    if max_time is not None:
        out, n = re.subn(r"^max_time\s*=.*;$", f"max_time = {f32_text(max_time)};", out, flags=re.M)
        assert n == 1
    for node, val in (istate or {}).items():
        out, n = re.subn(rf"^{re.escape(node)}\.istate\s*=.*;$", f"{node}.istate = {val};", out, flags=re.M)
        assert n == 1, node
    return out


def oracle_variant(tmp_path, bnd, cfg_text, variant, n, tag):
    cfg = tmp_path / f"{tag}.cfg"
    cfg.write_text(cfg_text)
    raw = tmp_path / f"{tag}.oracle.raw"
    run_oracle(bnd, str(cfg), "--n", n, "--variant", variant, "--raw", raw)
    return read_raw(raw)


def gpu_raw(tmp_path, result, tag):
    path = tmp_path / f"{tag}.gpu.raw"
    result.write_raw(str(path))
    return read_raw(path)


def assert_raw_equal(a, b, what):
    assert a["header"] == b["header"], what
    for k in ("T", "E", "F", "P"):
        assert a.get(k) == b.get(k), (what, k)
    if a["W"] != b["W"]:
        diff = [(k, a["W"].get(k), b["W"].get(k)) for k in set(a["W"]) | set(b["W"])
                if a["W"].get(k) != b["W"].get(k)]
        raise AssertionError(f"{what}: {len(diff)} window cells differ, e.g. {diff[:4]}")


def test_variants_match_oracle(tmp_path, b50):
    """Rates, horizon and mutations per variant, each bit-exact vs the oracle."""
    bnd, cfg = b50
    cfg_text = open(cfg).read()
    prog = program(bnd, cfg)
    names = prog.model.param_names
    base = prog.model.param_values
    node = {nm: i for i, nm in enumerate(prog.model.node_names)}
    n_nodes = len(node)

    scaled = base.copy()
    for k, nm in enumerate(names):
        if nm.startswith("$u_"):
            scaled[k] = np.float32(base[k] * np.float32(1.5))
    mut = np.zeros(n_nodes, dtype=np.int8)
    mut[node["N5"]] = 1
    mut[node["N7"]] = -1
    variants = [
        {"n_trajectories": 1500},
        {"n_trajectories": 1200, "rates": scaled},
        {"n_trajectories": 1000, "max_time": 40.0},
        {"n_trajectories": 1300, "mutations": mut},
        {"n_trajectories": 700, "rates": scaled, "max_time": 25.0, "mutations": mut},
    ]
    print("[gpu] batched_50 5 variants", flush=True)
    results, info = prog.simulate_batched(variants)
    assert len(results) == len(variants)
    i5, i7 = names.index("$d_N5"), names.index("$u_N7")
    expect_cfg = [
        cfg_text,
        edited_cfg(cfg_text, params={nm: scaled[k] for k, nm in enumerate(names) if nm.startswith("$u_")}),
        edited_cfg(cfg_text, max_time=40.0),
        edited_cfg(cfg_text, params={"d_N5": 0.0, "u_N7": 0.0}, istate={"N5": 1, "N7": 0}),
        edited_cfg(cfg_text,
                   params={**{nm: scaled[k] for k, nm in enumerate(names) if nm.startswith("$u_")},
                           "d_N5": 0.0, "u_N7": 0.0},
                   max_time=25.0, istate={"N5": 1, "N7": 0}),
    ]
    assert i5 >= 0 and i7 >= 0
    total_t = 0
    for v, (spec, ctext) in enumerate(zip(variants, expect_cfg)):
        g = gpu_raw(tmp_path, results[v], f"v{v}")
        o = oracle_variant(tmp_path, bnd, ctext, v, spec["n_trajectories"], f"v{v}")
        assert_raw_equal(g, o, f"variant {v}")
        total_t += g["T"][0]
    assert info.transitions == total_t


def test_long_horizon_batch_vs_oracle(tmp_path, b50):
    """A batch whose longest horizon has more windows (300) than the per-wave
    LDS entropy sums hold (256), so its entropy sums go straight to global
    atomics, next to a short one: every variant still equals the oracle."""
    bnd, cfg = b50
    cfg_text = open(cfg).read()
    prog = program(bnd, cfg)
    variants = [{"n_trajectories": 600, "max_time": 300.0}, {"n_trajectories": 700, "max_time": 40.0}]
    print("[gpu] batched_50 300-window horizon", flush=True)
    results, _ = prog.simulate_batched(variants)
    for v, spec in enumerate(variants):
        g = gpu_raw(tmp_path, results[v], f"L{v}")
        o = oracle_variant(tmp_path, bnd, edited_cfg(cfg_text, max_time=spec["max_time"]), v,
                           spec["n_trajectories"], f"L{v}")
        assert_raw_equal(g, o, f"variant {v}")


def test_default_variant_equals_plain_run(tmp_path, b50):
    """Variant v with the model's rates == plain kernel with stream tag v."""
    bnd, cfg = b50
    prog = program(bnd, cfg)
    results, _ = prog.simulate_batched([{"n_trajectories": 900}, {"n_trajectories": 1100}])
    for v, n in ((0, 900), (1, 1100)):
        plain = prog.simulate(n_trajectories=n, variant=v)
        assert_raw_equal(gpu_raw(tmp_path, results[v], f"b{v}"), gpu_raw(tmp_path, plain, f"p{v}"), f"variant {v}")


def test_graph_replays_match_eager(tmp_path, b50):
    """Replays of the captured batch, with other device work between them, equal the eager launch."""
    bnd, cfg = b50
    prog = program(bnd, cfg)
    rng = np.random.default_rng(11)
    base = prog.model.param_values
    variants = [{"n_trajectories": 300 + 37 * v,
                 "rates": (base * rng.uniform(0.5, 2.0, base.size)).astype(np.float32)} for v in range(12)]
    eager, info0 = prog.simulate_batched(variants, use_graph=False, block_size=128, traj_per_lane=2)
    assert info0.used_graph == 0
    want = [gpu_raw(tmp_path, r, f"e{v}") for v, r in enumerate(eager)]
    for rep in range(3):
        got, info = prog.simulate_batched(variants, use_graph=True, block_size=128, traj_per_lane=2)
        assert info.used_graph == 1
        for v, r in enumerate(got):
            assert_raw_equal(gpu_raw(tmp_path, r, f"g{rep}_{v}"), want[v], f"replay {rep} variant {v}")
        prog.simulate(n_trajectories=500)  # eager work between replays


@pytest.mark.parametrize("block,tpl", [(64, 1), (256, 4), (128, 6), (64, 8), (256, 12), (0, 0)])
def test_launch_shape_invariance(tmp_path, b50, block, tpl):
    bnd, cfg = b50
    prog = program(bnd, cfg)
    variants = [{"n_trajectories": 500, "max_time": 30.0 + 10 * v} for v in range(6)]
    ref, _ = prog.simulate_batched(variants, block_size=64, traj_per_lane=1, use_graph=False)
    got, info = prog.simulate_batched(variants, block_size=block, traj_per_lane=tpl)
    if block:
        assert info.block_size == block and info.traj_per_lane == tpl
    for v in range(len(variants)):
        assert_raw_equal(gpu_raw(tmp_path, got[v], f"s{v}"), gpu_raw(tmp_path, ref[v], f"r{v}"), f"variant {v}")


def test_discrete_time_variants(tmp_path):
    """cellcycle runs in discrete time (discrete_time = 1 in its cfg)."""
    bnd, cfg = model_paths("cellcycle")
    prog = program(bnd, cfg)
    assert prog.model.info.discrete_time == 1
    results, _ = prog.simulate_batched([{"n_trajectories": 800}, {"n_trajectories": 600, "max_time": 3.0}])
    text = open(cfg).read()
    for v, n, ctext in ((0, 800, text), (1, 600, edited_cfg(text, max_time=3.0))):
        o = oracle_variant(tmp_path, bnd, ctext, v, n, f"d{v}")
        assert_raw_equal(gpu_raw(tmp_path, results[v], f"d{v}"), o, f"variant {v}")


def test_many_variants_sampled_vs_oracle(tmp_path, b50):
    """256 variants in one launch; a sample of them checked against the oracle."""
    bnd, cfg = b50
    cfg_text = open(cfg).read()
    prog = program(bnd, cfg)
    names = prog.model.param_names
    base = prog.model.param_values
    rng = np.random.default_rng(5)
    rates = [(base * rng.uniform(0.5, 2.0, base.size)).astype(np.float32) for _ in range(256)]
    variants = [{"n_trajectories": 64 + (v % 7) * 16, "rates": rates[v]} for v in range(256)]
    results, info = prog.simulate_batched(variants)
    assert len(results) == 256
    assert info.transitions == sum(r.info.transitions for r in results)
    for v in (0, 77, 255):
        o = oracle_variant(tmp_path, bnd, edited_cfg(cfg_text, params=dict(zip(names, rates[v]))), v,
                           variants[v]["n_trajectories"], f"m{v}")
        assert_raw_equal(gpu_raw(tmp_path, results[v], f"m{v}"), o, f"variant {v}")


def test_bad_variants_rejected(b50):
    bnd, cfg = b50
    prog = program(bnd, cfg)
    with pytest.raises(ValueError):
        prog.simulate_batched([{"n_trajectories": 10, "rates": np.ones(3, np.float32)}])
    with pytest.raises(mg.MbgError):
        prog.simulate_batched([{"n_trajectories": -5}])
    res, _ = prog.simulate_batched([])
    assert res == []


def _mutant_lines(prog, mut):
    names = prog.model.node_names
    return "".join(f"\n{names[i]}.mutant = {'ON' if m > 0 else 'OFF'};" for i, m in enumerate(mut) if m)


def test_benchmark_variant_set_vs_oracle(tmp_path, b50):
    """The batched workload's own 256 variants (netgen.batched_variants: rates x
    lognormal(0, 0.5), 0-3 mutants, 1e3-1e5 trajectories, max_time 10-200) in one
    launch; the variants with mutants and the fewest trajectories are checked against
    the oracle (rates and horizon edited into the .cfg, mutants as X.mutant lines)."""
    bnd, cfg = b50
    cfg_text = open(cfg).read()
    prog = program(bnd, cfg)
    names = prog.model.param_names
    variants = netgen.batched_variants(list(prog.model.param_values), prog.model.info.n_nodes)
    assert len(variants) == 256 and sum(v["n_trajectories"] for v in variants) > 5_000_000
    print("[gpu] batched_50 benchmark variant set, one launch", flush=True)
    results, info = prog.simulate_batched(variants)
    assert info.transitions == sum(r.info.transitions for r in results)
    with_mut = sorted((v for v in range(256) if any(variants[v]["mutations"])),
                      key=lambda v: variants[v]["n_trajectories"])
    assert len(with_mut) > 100
    for v in with_mut[:3]:
        spec = variants[v]
        text = edited_cfg(cfg_text, params=dict(zip(names, spec["rates"])), max_time=spec["max_time"])
        text += _mutant_lines(prog, spec["mutations"]) + "\n"
        print(f"[gpu] oracle variant {v}: {spec['n_trajectories']} trajectories", flush=True)
        o = oracle_variant(tmp_path, bnd, text, v, spec["n_trajectories"], f"b{v}")
        assert_raw_equal(gpu_raw(tmp_path, results[v], f"b{v}"), o, f"variant {v}")


@pytest.mark.parametrize("cells", [-1, 64, 256, 4096])
def test_lds_cell_cache_layouts_agree(tmp_path, b50, cells):
    """The per-wave LDS window-cell cache (none, a small one where most cells collide
    and go to HBM, and larger ones) changes no sum."""
    bnd, cfg = b50
    prog = program(bnd, cfg)
    variants = netgen.batched_variants(list(prog.model.param_values), prog.model.info.n_nodes, count=24, seed=9)
    for v in variants:
        v["n_trajectories"] = min(v["n_trajectories"], 3000)
    ref, _ = prog.simulate_batched(variants, block_size=128, traj_per_lane=2, lds_cells=-1, use_graph=False)
    got, info = prog.simulate_batched(variants, block_size=128, traj_per_lane=2, lds_cells=cells, use_graph=False)
    assert info.lds_cells == max(cells, 0)
    for v in range(len(variants)):
        assert_raw_equal(gpu_raw(tmp_path, got[v], f"c{v}"), gpu_raw(tmp_path, ref[v], f"r{v}"), f"variant {v}")


def test_tuned_shape_is_kept_on_disk(tmp_path, b50):
    """The autotuner's choice (block, trajectories per lane, LDS layout) is written to the
    tuning cache; a new process with the same model and batch reads it instead of timing."""
    import json
    import subprocess
    import sys
    from conftest import ROOT
    bnd, cfg = b50
    code = ("import json, maboss.gpu_amd as mg; from maboss.gpu_amd import netgen; "
            f"p = mg.Model.load({bnd!r}, {cfg!r}).compile(); "
            "vs = netgen.batched_variants(list(p.model.param_values), p.model.info.n_nodes, count=16, seed=4); "
            "_, i = p.simulate_batched(vs, device_only=True); "
            "print(json.dumps([i.block_size, i.traj_per_lane, i.lds_cells, i.tuned_cached]))")
    env = dict(os.environ, MBG_CACHE_DIR=str(tmp_path / "cache"))
    runs = []
    for k in range(2):
        print(f"[gpu] tuning-cache process {k}", flush=True)
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, cwd=ROOT, env=env)
        assert p.returncode == 0, p.stderr[-3000:]
        runs.append(json.loads(p.stdout.strip().splitlines()[-1]))
    assert runs[0][3] == 0 and runs[1][3] == 1
    assert runs[0][:3] == runs[1][:3]
    assert any(f.endswith(".tune") for f in os.listdir(tmp_path / "cache"))
