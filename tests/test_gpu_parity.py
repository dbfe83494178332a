"""GPU engine vs host oracle: bit-exact parity through the C-ABI.

Every comparison is on the exact integer dump (fixed-point window sums,
entropy sums, final-state counts, fixed points and counters), so "equal"
means equal bits.  Sizes are chosen so the oracle finishes in seconds; the
full-size runs check size-independent properties instead (shard additivity,
determinism, conservation)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import maboss.gpu_amd as mg
from maboss.gpu_amd import netgen
from conftest import REF_MODELS, model_paths, read_raw, run_oracle

pytestmark = pytest.mark.gpu

TWO32 = 4294967296.0


@pytest.fixture(scope="module")
def wl(tmp_path_factory):
    d = tmp_path_factory.mktemp("wl")
    return {w: netgen.write_workload(w, str(d)) for w in netgen.WORKLOADS}


_programs = {}


def program(bnd, cfg, entropy=0):
    key = (bnd, cfg, entropy)
    if key not in _programs:
        m = mg.Model.load(bnd, cfg)
        _programs[key] = m.compile(entropy_formula=entropy)
    return _programs[key]


def gpu_raw(tmp_path, bnd, cfg, tag="g", entropy=0, **kw):
    print(f"[gpu] {os.path.basename(bnd)} {kw}", flush=True)
    r = program(bnd, cfg, entropy).simulate(**kw)
    path = tmp_path / f"{tag}.raw"
    r.write_raw(str(path))
    return path, r


def oracle_raw(tmp_path, bnd, cfg, tag="o", n=None, extra=()):
    path = tmp_path / f"{tag}.raw"
    args = ["--raw", path, *extra]
    if n is not None:
        args = ["--n", n, *args]
    run_oracle(bnd, cfg, *args)
    return path


def assert_same(a, b):
    ta, tb = open(a).read(), open(b).read()
    if ta != tb:
        ra, rb = read_raw(a), read_raw(b)
        for k in ("T", "E", "F", "P"):
            assert ra.get(k) == rb.get(k), k
        diff = {k: (ra["W"].get(k), rb["W"].get(k)) for k in set(ra["W"]) | set(rb["W"])
                if ra["W"].get(k) != rb["W"].get(k)}
        raise AssertionError(f"{len(diff)} window cells differ, e.g. {list(diff.items())[:5]}")


@pytest.mark.parametrize("name,n", [("cellcycle", 20000), ("metastasis", 4000), ("sizek", 3000),
                                    ("Montagud2022_Prostate_Cancer", 2000)])
def test_reference_models_bit_exact(tmp_path, name, n):
    bnd, cfg = model_paths(name)
    g, _ = gpu_raw(tmp_path, bnd, cfg, n_trajectories=n)
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=n))


@pytest.mark.parametrize("name,n", [("plain_small", 10000), ("plain_100", 3000), ("plain_100_discrete", 3000),
                                    ("batched_50", 3000), ("grad_30", 5000)])
def test_workloads_bit_exact(tmp_path, wl, name, n):
    bnd, cfg = wl[name]
    g, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=n)
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=n))
    if name.startswith("plain_100"):
        assert r.info.window_mode == 1  # 50 observed nodes: hashed window cells


def test_packed_cells_with_escapes_bit_exact(tmp_path, wl):
    """Window cells copied as 12-byte records: plain_100's keys take 57 bits,
    leaving 39 for the value, so cells above 128 trajectory-ticks go through
    the side list; every cell still equals the oracle's."""
    bnd, cfg = wl["plain_100"]
    n = 30000
    g, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=n)
    assert r.info.window_record_bytes == 12
    assert r.info.window_escapes > 0
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=n))
    # the sorted-pairs mode copies plain 16-byte cells: same result
    s, rs = gpu_raw(tmp_path, bnd, cfg, "s", n_trajectories=n, window_mode="sort")
    assert rs.info.window_record_bytes == 16
    assert_same(g, s)


@pytest.mark.parametrize("name,n", [("plain_small", 4000), ("plain_100", 2000), ("plain_100_discrete", 2000),
                                    ("grad_30", 3000), ("scale_1000", 300)])
def test_wave_kernel_bit_exact(tmp_path, wl, name, n):
    """Wave-per-trajectory kernel against the oracle's wave summation order."""
    bnd, cfg = wl[name]
    g, r = gpu_raw(tmp_path, bnd, cfg, kernel="wave", n_trajectories=n)
    assert r.info.kernel == 1
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=n, extra=("--order", "wave")))


def test_wave_kernel_bit_exact_32_rows(tmp_path):
    """1500 nodes: 24 nodes per lane, a 32-row rate array per wave and the
    per-lane selection loop (the 16-row kernels select with DPP running sums)."""
    net = netgen.scale_free_network(7, 1500, 2.5, 4, 12, 0.0, 1.0, 0.4)
    net.constants = dict(sample_count=100, max_time=4, time_tick=0.5, discrete_time=0, seed_pseudorandom=7)
    bnd, cfg = net.write(str(tmp_path / "sf1500"))
    g, r = gpu_raw(tmp_path, bnd, cfg, kernel="wave", n_trajectories=100)
    assert r.info.kernel == 1 and r.info.transitions > 10000
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=100, extra=("--order", "wave")))


def write_memory_model(tmp_path, n_mem=60, n_dyn=10, n=3000):
    """n_mem internal nodes that keep their initial state (logic = self) and
    n_dyn observed nodes that copy one of them: every trajectory ends in a
    fixed point, nearly all of them distinct (2^n_mem initial states)."""
    nodes = []
    for i in range(n_mem):
        nodes.append(f"Node X{i} {{\n  logic = X{i};\n  rate_up = @logic ? 1.0 : 0.0;\n"
                     f"  rate_down = @logic ? 0.0 : 1.0;\n}}\n")
    for i in range(n_dyn):
        nodes.append(f"Node Y{i} {{\n  logic = X{i} & !X{i + 1};\n  rate_up = @logic ? 2.0 : 0.0;\n"
                     f"  rate_down = @logic ? 0.0 : 0.5;\n}}\n")
    bnd, cfg = tmp_path / "mem.bnd", tmp_path / "mem.cfg"
    bnd.write_text("".join(nodes))
    cfg.write_text(f"sample_count = {n};\nmax_time = 20;\ntime_tick = 1;\ndiscrete_time = 0;\n"
                   f"seed_pseudorandom = 11;\n" + "".join(f"X{i}.is_internal = 1;\n" for i in range(n_mem)))
    return str(bnd), str(cfg)


@pytest.mark.parametrize("kernel,order", [("lane", ()), ("wave", ("--order", "wave"))])
def test_many_distinct_fixed_points(tmp_path, kernel, order):
    """More distinct fixed points than a block's LDS cache holds: the cache's
    fall-through to the global table and its flush at the block's end."""
    bnd, cfg = write_memory_model(tmp_path)
    g, r = gpu_raw(tmp_path, bnd, cfg, kernel=kernel, n_trajectories=3000)
    assert r.info.kernel == (1 if kernel == "wave" else 0)
    assert len(read_raw(g)["P"]) > 2900
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=3000, extra=order))


@pytest.mark.parametrize("n_mem,n_dyn,min_fixed", [(50, 10, 2900), (4, 3, 12)])
def test_fixed_point_counts_kept_per_lane(tmp_path, n_mem, n_dyn, min_fixed):
    """Models of up to 64 nodes: a lane counts repeats of its last fixed point
    in registers and inserts when it ends in another one and when it stops.
    60 nodes, nearly every trajectory its own fixed point (an insert at every
    end); 7 nodes, 16 attractors (lanes switch between few points)."""
    bnd, cfg = write_memory_model(tmp_path, n_mem=n_mem, n_dyn=n_dyn)
    g, r = gpu_raw(tmp_path, bnd, cfg, kernel="lane", n_trajectories=3000)
    assert r.info.kernel == 0
    assert len(read_raw(g)["P"]) >= min_fixed
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=3000))


def write_slow_model(tmp_path):
    """Six slow togglers: a dwell lasts ~5 time units, ~100 windows of 0.05, so
    the windows between a record's first and last ones fill many 64-lane
    passes of the wave."""
    nodes = "".join(f"Node S{i} {{\n  logic = !S{i};\n  rate_up = @logic ? {0.02 + 0.01 * i} : 0.0;\n"
                    f"  rate_down = @logic ? 0.0 : {0.05 - 0.005 * i};\n}}\n" for i in range(6))
    bnd, cfg = tmp_path / "slow.bnd", tmp_path / "slow.cfg"
    bnd.write_text(nodes)
    cfg.write_text("sample_count = 2000;\nmax_time = 100;\ntime_tick = 0.05;\ndiscrete_time = 0;\n"
                   "seed_pseudorandom = 5;\n")
    return str(bnd), str(cfg)


@pytest.mark.parametrize("win", ["dense", "hash", "sort"])
def test_long_dwells_bit_exact(tmp_path, win):
    """Records spanning hundreds of windows (the wave-wide window pass), in
    each HBM window storage, against the oracle; the batched kernel's run of
    the same model equals the plain run."""
    bnd, cfg = write_slow_model(tmp_path)
    g, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=2000, window_mode=win)
    assert r.info.window_mode == {"dense": 0, "hash": 1, "sort": 3}[win]
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=2000))
    if win == "dense":
        results, _ = program(bnd, cfg).simulate_batched([{"n_trajectories": 2000}])
        b = tmp_path / "batched.raw"
        results[0].write_raw(str(b))
        assert_same(b, g)


def test_wave_kernel_reference_model(tmp_path):
    bnd, cfg = model_paths("cellcycle")
    g, _ = gpu_raw(tmp_path, bnd, cfg, kernel="wave", n_trajectories=20000)
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=20000, extra=("--order", "wave")))
    # nodes with more than 6 inputs do not fit the wave kernel's truth tables
    bnd, cfg = model_paths("metastasis")
    with pytest.raises(mg.MbgError) as e:
        program(bnd, cfg).simulate(n_trajectories=10, kernel="wave")
    assert e.value.code == 5


def test_scale_1000_auto_uses_wave(tmp_path, wl):
    bnd, cfg = wl["scale_1000"]
    _, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=64)
    assert r.info.kernel == 1 and r.info.finished == 64


def test_reference_entropy_formula(tmp_path):
    bnd, cfg = model_paths("metastasis")
    g, _ = gpu_raw(tmp_path, bnd, cfg, entropy=1, n_trajectories=3000)
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=3000, extra=("--entropy", "reference")))


@pytest.mark.parametrize("block,per_cu", [(64, 1), (128, 3), (192, 2), (256, 1), (256, 0)])
def test_launch_shape_invariance(tmp_path, wl, block, per_cu):
    for bnd, cfg in (model_paths("cellcycle"), wl["plain_100"]):
        a, _ = gpu_raw(tmp_path, bnd, cfg, "a", n_trajectories=5000, block_size=block, blocks_per_cu=per_cu)
        b, _ = gpu_raw(tmp_path, bnd, cfg, "b", n_trajectories=5000)
        assert_same(a, b)


def test_block_size_validated():
    """Block sizes beyond the kernel's launch bound are refused on the host."""
    bnd, cfg = model_paths("cellcycle")
    for block in (512, 96):
        with pytest.raises(mg.MbgError) as e:
            program(bnd, cfg).simulate(n_trajectories=100, block_size=block)
        assert e.value.code == 1


@pytest.mark.parametrize("win,fin", [("dense", "dense"), ("hash", "hash"), ("lds", "lds"), ("hash", "dense"),
                                     ("sort", "lds"), ("sort", "hash")])
def test_storage_mode_invariance(tmp_path, win, fin):
    bnd, cfg = model_paths("cellcycle")
    a, r = gpu_raw(tmp_path, bnd, cfg, "a", n_trajectories=8000, window_mode=win, final_mode=fin)
    b, _ = gpu_raw(tmp_path, bnd, cfg, "b", n_trajectories=8000)
    assert r.info.window_mode == {"dense": 0, "hash": 1, "lds": 2, "sort": 3}[win]
    assert_same(a, b)


@pytest.mark.parametrize("win", ["dense", "hash", "sort"])
def test_window_modes_with_tails_bit_exact(tmp_path, wl, win):
    """Continuous time with fixed points: tail records expanded into each storage mode."""
    bnd, cfg = wl["plain_small"]
    g, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=6000, window_mode=win)
    assert r.info.window_mode == {"dense": 0, "hash": 1, "sort": 3}[win]
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=6000))


def test_sorted_pairs_capacity(tmp_path, wl):
    """An explicit pair capacity that is too small is an error; the automatic one grows."""
    bnd, cfg = wl["plain_100"]
    prog = program(bnd, cfg)
    with pytest.raises(mg.MbgError) as e:
        prog.simulate(n_trajectories=2000, window_mode="sort", window_capacity=1000)
    assert e.value.code == 8
    g, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=2000, window_mode="sort")
    assert r.info.window_capacity >= r.info.n_window_cells
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=2000))


def test_hash_table_full_spills_instead_of_rerunning(tmp_path, wl):
    """The hashed window table with the automatic capacity spills nothing; a fixed
    capacity far below the run's cells sends the cells past the probe bound to the
    spill list (growing the list, not the caller's table, if it fills) and the
    statistics stay exact."""
    bnd, cfg = wl["plain_100"]
    want = oracle_raw(tmp_path, bnd, cfg, n=2000)
    a, r = gpu_raw(tmp_path, bnd, cfg, "a", n_trajectories=2000, window_mode="hash", window_capacity=0)
    assert r.info.window_spills == 0
    assert_same(a, want)
    for cap in (1 << 14, 1024):
        b, rb = gpu_raw(tmp_path, bnd, cfg, f"b{cap}", n_trajectories=2000, window_mode="hash", window_capacity=cap)
        assert rb.info.window_capacity == cap and rb.info.window_spills > rb.info.n_window_cells // 2
        assert_same(b, want)


def test_edge_cases(tmp_path):
    # empty run
    bnd, cfg = model_paths("cellcycle")
    _, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=0)
    assert r.info.transitions == 0 and r.final_states()[0].size == 0
    # one trajectory, and an odd id offset
    g, _ = gpu_raw(tmp_path, bnd, cfg, "one", n_trajectories=1, first_id=12345)
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, "o1", n=1, extra=("--first-id", 12345)))
    # step cap: truncated trajectories are not final states
    bnd2, cfg2 = model_paths("metastasis")
    g, r = gpu_raw(tmp_path, bnd2, cfg2, "cap", n_trajectories=500, max_steps=7)
    assert r.info.truncated > 0
    assert_same(g, oracle_raw(tmp_path, bnd2, cfg2, "ocap", n=500, extra=("--max-steps", 7)))


def test_fixed_points_discrete_and_continuous(tmp_path):
    for disc in (0, 1):
        b, c = tmp_path / f"d{disc}.bnd", tmp_path / f"d{disc}.cfg"
        b.write_text("node A { rate_up = 0; rate_down = $l; }\nnode B { logic = A; rate_up = @logic ? 2 : 0; "
                     "rate_down = @logic ? 0 : 2; }\nnode C { rate_up = 0; rate_down = 0; }\n")
        c.write_text(f"$l = 0.5; A.istate = 1; C.is_internal = 1; max_time = 6; time_tick = 0.5; "
                     f"discrete_time = {disc}; sample_count = 5000;")
        g, r = gpu_raw(tmp_path, str(b), str(c), f"g{disc}")
        assert_same(g, oracle_raw(tmp_path, str(b), str(c), f"o{disc}"))
        assert r.fixed_points()[0].shape[0] >= 1


def test_determinism_and_shards(tmp_path, wl):
    bnd, cfg = wl["plain_100"]
    n = 20000
    a, ra = gpu_raw(tmp_path, bnd, cfg, "a", n_trajectories=n)
    b, _ = gpu_raw(tmp_path, bnd, cfg, "b", n_trajectories=n)
    assert_same(a, b)
    s1, _ = gpu_raw(tmp_path, bnd, cfg, "s1", n_trajectories=n // 4)
    s2, _ = gpu_raw(tmp_path, bnd, cfg, "s2", n_trajectories=n - n // 4, first_id=n // 4)
    x, y1, y2 = read_raw(a), read_raw(s1), read_raw(s2)
    merged = dict(y1["W"])
    for k, v in y2["W"].items():
        merged[k] = merged.get(k, 0) + v
    assert merged == x["W"]
    assert tuple(p + q for p, q in zip(y1["T"], y2["T"])) == x["T"]


def test_cli_csv_matches_oracle_csv(tmp_path):
    bnd, cfg = model_paths("cellcycle")
    small = tmp_path / "cc.cfg"
    small.write_text(open(cfg).read() + "\nsample_count = 20000;\n")
    p = mg.run_cli("-o", str(tmp_path / "gpu"), bnd, str(small))
    assert p.returncode == 0, p.stderr
    run_oracle(bnd, str(small), "-o", tmp_path / "cpu")
    for suffix in ("_probtraj.csv", "_fp.csv"):
        assert open(tmp_path / ("gpu" + suffix)).read() == open(tmp_path / ("cpu" + suffix)).read(), suffix
    p = mg.run_cli(bnd, str(small))
    assert p.returncode == 0 and p.stdout.startswith("final points:\n")
    run_oracle(bnd, str(small), "--print", tmp_path / "cpu.txt")
    assert p.stdout == open(tmp_path / "cpu.txt").read()


def test_full_size_plain_100_properties(tmp_path, wl):
    """The headline workload at its full size (1e6 trajectories): shard
    additivity, conservation of every window, all trajectories finished."""
    bnd, cfg = wl["plain_100"]
    n = 1000000
    full, r = gpu_raw(tmp_path, bnd, cfg, "full", n_trajectories=n)
    info = r.info
    assert info.finished == n and info.truncated == 0
    fin_states, fin_counts = r.final_states()
    assert int(fin_counts.sum()) == n
    half = n // 2
    h1, _ = gpu_raw(tmp_path, bnd, cfg, "h1", n_trajectories=half)
    h2, _ = gpu_raw(tmp_path, bnd, cfg, "h2", n_trajectories=n - half, first_id=half)
    x, y1, y2 = read_raw(full), read_raw(h1), read_raw(h2)
    merged = dict(y1["W"])
    for k, v in y2["W"].items():
        merged[k] = merged.get(k, 0) + v
    assert merged == x["W"]
    m = mg.Model.load(bnd, cfg)
    tick = np.float32(m.info.time_tick)
    for w in (0, 1, 50, 99):
        tot = sum(v for (ww, _), v in x["W"].items() if ww == w)
        lo = np.float32(w) * tick
        hi = min(np.float32(w + 1) * tick, np.float32(m.info.max_time))
        assert abs(tot / (n * TWO32) - float(hi - lo) / float(tick)) < 2e-6


def test_framework_hip_loaded_first(tmp_path):
    """With torch (and its bundled HIP runtime / code-object manager) loaded
    before the engine, programs still compile and stay bit-exact."""
    import sys
    bnd, cfg = model_paths("cellcycle")
    raw = tmp_path / "g.raw"
    code = ("import torch; torch.cuda.init(); import maboss.gpu_amd as mg; "
            f"r = mg.Model.load({bnd!r}, {cfg!r}).compile(no_cache=True).simulate(n_trajectories=3000); "
            f"r.write_raw({str(raw)!r}); print('ok')")
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert p.returncode == 0 and "ok" in p.stdout, p.stderr[-3000:]
    assert_same(raw, oracle_raw(tmp_path, bnd, cfg, n=3000))


def test_runtime_rates_bit_exact(tmp_path, wl):
    """Rates given at run time (plain kernel, third flavour) equal rates edited into the model."""
    import re
    bnd, cfg = wl["grad_30"]
    prog = program(bnd, cfg)
    names = prog.model.param_names
    base = prog.model.param_values
    same = gpu_raw(tmp_path, bnd, cfg, "s", n_trajectories=4000, rates=base)[0]
    assert_same(same, gpu_raw(tmp_path, bnd, cfg, "p", n_trajectories=4000)[0])
    rates = (base * np.float32(0.7)).astype(np.float32)
    text = open(cfg).read()
    for k, nm in enumerate(names):
        text, c = re.subn(rf"^\{nm}\s*=.*;$", f"{nm} = {float(rates[k]):.9g};", text, flags=re.M)
        assert c == 1
    cfg2 = tmp_path / "scaled.cfg"
    cfg2.write_text(text)
    g, _ = gpu_raw(tmp_path, bnd, cfg, "r", n_trajectories=4000, rates=rates)
    assert_same(g, oracle_raw(tmp_path, bnd, str(cfg2), n=4000))
    with pytest.raises(ValueError):
        prog.simulate(n_trajectories=10, rates=rates[:3])


def test_scale_1000_bit_exact_2000(tmp_path, wl):
    """The 1000-node workload on the wave kernel, 2000 trajectories, against
    the oracle in the wave kernel's summation order (compiled formulas)."""
    bnd, cfg = wl["scale_1000"]
    g, r = gpu_raw(tmp_path, bnd, cfg, n_trajectories=2000)
    assert r.info.kernel == 1 and r.info.finished == 2000
    assert_same(g, oracle_raw(tmp_path, bnd, cfg, n=2000, extra=("--order", "wave", "--compiled")))


def _cells(res):
    """All window cells of a result as (window << n_obs | state) keys and raw sums."""
    n_obs = res.info.n_observed
    keys, vals = [], []
    for w in range(res.info.windows):
        win = res.window(w)
        keys.append((np.uint64(w) << np.uint64(n_obs)) | win.states)
        vals.append(win.raw)
    return np.concatenate(keys), np.concatenate(vals)


def test_full_size_scale_1000_properties(tmp_path, wl):
    """The 1000-node workload at its full size (1e6 trajectories): every
    trajectory finishes, the final states count them all, each window's cells
    cover its float32 extent for every trajectory, and two id-range shards add
    up to the full run cell for cell."""
    bnd, cfg = wl["scale_1000"]
    prog = program(bnd, cfg)
    n = 1000000
    print("[gpu] scale_1000 full size", flush=True)
    full = prog.simulate(n_trajectories=n)
    info = full.info
    assert info.kernel == 1 and info.finished == n and info.truncated == 0
    assert int(full.final_states()[1].sum()) == n
    m = prog.model.info
    tick, mt = np.float32(m.time_tick), np.float32(m.max_time)
    fk, fv = _cells(full)
    for w in (0, 1, 37, m.windows - 1):
        tot = sum(int(v) for v in full.window(w).raw)
        lo = np.float32(w) * tick
        hi = min(np.float32(w + 1) * tick, mt)
        assert abs(tot / (n * TWO32) - float(hi - lo) / float(tick)) < 2e-6
    print("[gpu] scale_1000 two shards", flush=True)
    half = n // 2
    a = prog.simulate(n_trajectories=half)
    b = prog.simulate(n_trajectories=n - half, first_id=half)
    ak, av = _cells(a)
    bk, bv = _cells(b)
    keys, inv = np.unique(np.concatenate([ak, bk]), return_inverse=True)
    sums = np.zeros(len(keys), dtype=np.uint64)
    np.add.at(sums, inv, np.concatenate([av, bv]))
    assert np.array_equal(keys, fk) and np.array_equal(sums, fv)
    assert a.info.transitions + b.info.transitions == info.transitions


def test_lds_budget_edge(tmp_path):
    """2048 windows of one observed node: the LDS entropy sums (16 KB) and a
    window histogram in LDS (32 KB) would fill the 48 KB budget, leaving no
    room for the per-wave tables; the plan puts the cells in HBM instead, and
    forcing them into LDS is a clear error."""
    b, c = tmp_path / "edge.bnd", tmp_path / "edge.cfg"
    b.write_text("node A { logic = !A; rate_up = @logic ? 3 : 0; rate_down = @logic ? 0 : 2; }\n")
    c.write_text("max_time = 204.8; time_tick = 0.1; sample_count = 500; seed_pseudorandom = 3;\n")
    g, r = gpu_raw(tmp_path, str(b), str(c), n_trajectories=500)
    assert r.info.windows == 2048 and r.info.window_mode == 0
    assert_same(g, oracle_raw(tmp_path, str(b), str(c), n=500))
    with pytest.raises(mg.MbgError) as e:
        program(str(b), str(c)).simulate(n_trajectories=10, window_mode="lds")
    assert e.value.code == 5
