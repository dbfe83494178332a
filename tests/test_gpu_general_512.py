"""General models above 256 nodes (the reference serves up to 512 nodes with
any formula, CMakeLists.txt:12-13, src/generator.cpp:108-119).

A 400-node model (13 state words) with 8-13-input nodes (threshold and
AND/OR formulas) and rates that are not of the fast form (rates scaled by
other nodes' states, rates not gated by @logic, rates from $parameters) is
outside the wave kernel's table-driven evaluator, so it runs on the lane
kernel with the model's generated code: bit-exact against the oracle in its
sequential order, with and without the error columns.  A 300-node fast-form
model runs on the wave kernel by default and on the lane kernel when asked,
each bit-exact against the oracle's matching order."""
import pytest

import maboss.gpu_amd as mg
from conftest import read_raw, run_oracle
from maboss.gpu_amd.netgen import SplitMix64, _and_or, _threshold, fmt, random_network

pytestmark = pytest.mark.gpu


def general_model(tmp_path, seed=5, n=400, n_obs=10, n_traj=2000):
    rng = SplitMix64(seed)
    names = [f"G{i}" for i in range(n)]
    bnd, wide = [], 0
    for i in range(n):
        big = rng.uniform() < 0.2
        k = rng.randint(8, 13) if big else rng.randint(1, 4)
        wide += big
        par = rng.sample([j for j in range(n) if j != i], k)
        pn = [names[j] for j in par]
        logic = _threshold(rng, pn) if big and rng.uniform() < 0.5 else _and_or(rng, pn)
        u, d = fmt(rng.lognormal(0, 0.7)), fmt(rng.lognormal(0, 0.7))
        a, b = names[par[0]], names[par[-1]]
        form = rng.randint(0, 3)
        if form == 0:
            up, dn = f"@logic ? {u} : 0", f"@logic ? 0 : {d}"
        elif form == 1:
            up, dn = f"@logic ? {u} * (1 + {a}) : 0.0", f"@logic ? 0.0 : {d} / (1 + {b})"
        elif form == 2:
            up, dn = f"({a} & !{b}) ? {u} : 0.25 * {u}", f"@logic ? 0 : {d}"
        else:
            up, dn = f"@logic ? $k{i} : 0", f"({a} | {b}) ? $k{i} * 0.5 : {d}"
        bnd.append(f"Node {names[i]} {{ logic = {logic}; rate_up = {up}; rate_down = {dn}; }}")
    assert wide > 40
    cfg = [f"$k{i} = {fmt(rng.lognormal(0, 0.5))};" for i in range(n)]
    obs = set(rng.sample(list(range(n)), n_obs))
    cfg += [f"{names[i]}.is_internal = {0 if i in obs else 1};" for i in range(n)]
    cfg += [f"{names[i]}.istate = {fmt(rng.uniform())};" for i in range(0, n, 7)]
    cfg.append(f"sample_count = {n_traj}; max_time = 4; time_tick = 0.5; discrete_time = 0; seed_pseudorandom = 11;")
    b, c = tmp_path / "g400.bnd", tmp_path / "g400.cfg"
    b.write_text("\n".join(bnd) + "\n")
    c.write_text("\n".join(cfg) + "\n")
    return str(b), str(c)


def test_400_node_general_model_on_the_lane_kernel(tmp_path):
    bnd, cfg = general_model(tmp_path)
    m = mg.Model.load(bnd, cfg)
    assert m.info.n_nodes == 400 and m.info.n_words == 13
    print("[gpu] 400-node general model: compile", flush=True)
    prog = m.compile()
    print("[gpu] 400-node general model: simulate", flush=True)
    r = prog.simulate()
    assert r.info.kernel == 0  # the wave kernel cannot run it
    r.write_raw(str(tmp_path / "g.raw"))
    run_oracle(bnd, cfg, "--compiled", "--raw", tmp_path / "o.raw")
    g, o = read_raw(tmp_path / "g.raw"), read_raw(tmp_path / "o.raw")
    assert g["T"][0] > 2000 * 20
    assert g == o
    print("[gpu] 400-node general model: error columns", flush=True)
    prog.simulate(n_trajectories=1000, errors=True).write_raw(str(tmp_path / "ge.raw"))
    run_oracle(bnd, cfg, "--compiled", "--n", 1000, "--errors", "--raw", tmp_path / "oe.raw")
    ge = read_raw(tmp_path / "ge.raw")
    assert ge["Q"] and ge == read_raw(tmp_path / "oe.raw")


def test_300_node_fast_form_model_wave_by_default_lane_on_request(tmp_path):
    net = random_network(21, 300, (1, 4), 290, 0.0, 1.0)
    net.constants = dict(sample_count=2000, max_time=5, time_tick=0.5, discrete_time=0, seed_pseudorandom=21)
    bnd, cfg = net.write(str(tmp_path / "f300"))
    prog = mg.Model.load(bnd, cfg).compile()
    print("[gpu] 300-node fast-form model: wave kernel", flush=True)
    rw = prog.simulate()
    assert rw.info.kernel == 1
    rw.write_raw(str(tmp_path / "w.raw"))
    run_oracle(bnd, cfg, "--compiled", "--order", "wave", "--raw", tmp_path / "ow.raw")
    assert read_raw(tmp_path / "w.raw") == read_raw(tmp_path / "ow.raw")
    print("[gpu] 300-node fast-form model: lane kernel", flush=True)
    rl = prog.simulate(kernel="lane")
    assert rl.info.kernel == 0
    rl.write_raw(str(tmp_path / "l.raw"))
    run_oracle(bnd, cfg, "--compiled", "--raw", tmp_path / "ol.raw")
    assert read_raw(tmp_path / "l.raw") == read_raw(tmp_path / "ol.raw")
