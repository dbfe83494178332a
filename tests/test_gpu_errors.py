"""Error columns on the GPU: the second-moment replay against the oracle.

The replay (mbg_moments_kernel.h) must give exactly the oracle's 128-bit sums
of per-trajectory squares (R and Q lines of the raw dump) on reference models,
in continuous and discrete time, with run-time rates and mutants; a lane
buffer of 1 state (every window with two states goes through the host-side
spill list) must give the same sums as the default 8; and the command line's
--errors CSV must equal the oracle's byte for byte."""
import numpy as np
import pytest

import maboss.gpu_amd as mg
from conftest import model_paths, read_raw, run_oracle
from test_cpu_mutants import CFG, _bnd, _write

pytestmark = pytest.mark.gpu


def _gpu(tmp_path, bnd, cfg, tag, **kw):
    r = mg.Model.load(bnd, cfg).compile().simulate(errors=True, **kw)
    r.write_raw(str(tmp_path / f"{tag}.raw"))
    return read_raw(tmp_path / f"{tag}.raw"), r


@pytest.mark.parametrize("name,n", [("cellcycle", 20000), ("sizek", 5000), ("metastasis", 5000)])
def test_moments_match_oracle_on_reference_models(tmp_path, name, n):
    bnd, cfg = model_paths(name)
    print(f"[gpu] moments {name} n={n}", flush=True)
    g, r = _gpu(tmp_path, bnd, cfg, "g", n_trajectories=n)
    run_oracle(bnd, cfg, "--n", n, "--errors", "--raw", tmp_path / "o.raw")
    o = read_raw(tmp_path / "o.raw")
    assert g["R"] and g["Q"]
    assert g == o
    assert r.info.has_errors == 1


def test_moments_discrete_time_and_mutants(tmp_path):
    bnd, cfg = _write(tmp_path, "m", _bnd(), CFG + "discrete_time = 1; B.mutant = ON;\nsample_count = 10000;\n")
    print("[gpu] moments discrete + mutant", flush=True)
    g, _ = _gpu(tmp_path, bnd, cfg, "g")
    run_oracle(bnd, cfg, "--errors", "--raw", tmp_path / "o.raw")
    assert g == read_raw(tmp_path / "o.raw")
    # one step per window and trajectory: every P is 0 or 1
    assert all(g["Q"][k] == v for k, v in g["W"].items())


def _continuous_cellcycle(tmp_path):
    bnd, cfg = model_paths("cellcycle")
    c = tmp_path / "cc_cont.cfg"
    c.write_text(open(cfg).read() + "\ndiscrete_time = 0;\n")
    return bnd, str(c)


def test_moments_match_oracle_continuous_cellcycle(tmp_path):
    bnd, cfg = _continuous_cellcycle(tmp_path)
    print("[gpu] moments cellcycle continuous", flush=True)
    g, _ = _gpu(tmp_path, bnd, cfg, "g", n_trajectories=20000)
    run_oracle(bnd, cfg, "--n", 20000, "--errors", "--raw", tmp_path / "o.raw")
    assert g == read_raw(tmp_path / "o.raw")


def test_lane_buffer_of_one_state_spills_and_agrees(tmp_path):
    bnd, cfg = _continuous_cellcycle(tmp_path)
    print("[gpu] moments K=1 vs K=8", flush=True)
    g8, r8 = _gpu(tmp_path, bnd, cfg, "g8", n_trajectories=20000)
    g1, r1 = _gpu(tmp_path, bnd, cfg, "g1", n_trajectories=20000, moment_states=1)
    assert r1.info.moment_spills > 1000
    assert g1 == g8


def test_moments_with_run_time_rates(tmp_path):
    bnd, cfg = model_paths("cellcycle")
    m = mg.Model.load(bnd, cfg)
    rates = np.array(m.param_values, dtype=np.float32) * np.float32(1.5)
    print("[gpu] moments with run-time rates", flush=True)
    r = m.compile().simulate(n_trajectories=8000, rates=rates, errors=True)
    r.write_raw(str(tmp_path / "g.raw"))
    # the oracle reads the same values from an edited .cfg
    text = open(cfg).read()
    for name, v in zip(m.param_names, rates):
        text += f"\n{name} = {float(v)!r};"
    ecfg = tmp_path / "e.cfg"
    ecfg.write_text(text + "\n")
    run_oracle(bnd, str(ecfg), "--n", 8000, "--errors", "--raw", tmp_path / "o.raw")
    assert read_raw(tmp_path / "g.raw") == read_raw(tmp_path / "o.raw")


def test_cli_errors_csv_matches_oracle(tmp_path):
    bnd, cfg = model_paths("cellcycle")
    small = tmp_path / "cc.cfg"
    small.write_text(open(cfg).read() + "\nsample_count = 20000;\n")
    print("[gpu] cli --errors", flush=True)
    p = mg.run_cli("--errors", "-o", str(tmp_path / "gpu"), bnd, str(small))
    assert p.returncode == 0, p.stderr
    run_oracle(bnd, str(small), "--errors", "-o", tmp_path / "cpu")
    a, b = open(tmp_path / "gpu_probtraj.csv").read(), open(tmp_path / "cpu_probtraj.csv").read()
    assert a == b
    # the columns carry values: some ErrorProba is positive
    row = a.splitlines()[5].split("\t")
    assert any(float(x) > 0 for x in row[7::3])


def test_errors_rejected_where_unsupported(tmp_path):
    bnd, cfg = model_paths("cellcycle")
    prog = mg.Model.load(bnd, cfg).compile()
    with pytest.raises(mg.MbgError):
        prog.simulate(n_trajectories=100, errors=True, device_only=True)
    with pytest.raises(mg.MbgError):
        prog.simulate(n_trajectories=100, errors=True, moment_states=3)
