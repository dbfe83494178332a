"""Mutant .cfg statements on the GPU: the command line and the Python API
against the oracle on the same mutant .cfg, and the plain path's mutants
against the batched form's mutant masks (two implementations of the same
semantics: forced on = rate_down 0 and istate 1, forced off = rate_up 0 and
istate 0)."""
import pytest

import maboss.gpu_amd as mg
from conftest import model_paths, read_raw, run_oracle
from test_cpu_mutants import CFG, _bnd, _write

pytestmark = pytest.mark.gpu


def test_cli_mutant_cfg_matches_oracle(tmp_path):
    bnd, cfg = model_paths("cellcycle")
    mcfg = tmp_path / "cc_mut.cfg"
    mcfg.write_text(open(cfg).read() + "\nsample_count = 20000;\nCycE.mutant = ON;\nRb.mutant = OFF;\n")
    print("[gpu] cli cellcycle CycE ON, Rb OFF", flush=True)
    p = mg.run_cli("-o", str(tmp_path / "gpu"), bnd, str(mcfg))
    assert p.returncode == 0, p.stderr
    run_oracle(bnd, str(mcfg), "-o", tmp_path / "cpu")
    for suffix in ("_probtraj.csv", "_fp.csv"):
        assert open(tmp_path / ("gpu" + suffix)).read() == open(tmp_path / ("cpu" + suffix)).read(), suffix
    # the mutant run differs from the plain one
    plain = tmp_path / "cc.cfg"
    plain.write_text(open(cfg).read() + "\nsample_count = 20000;\n")
    p = mg.run_cli("-o", str(tmp_path / "plain"), bnd, str(plain))
    assert p.returncode == 0, p.stderr
    assert open(tmp_path / "plain_probtraj.csv").read() != open(tmp_path / "gpu_probtraj.csv").read()


@pytest.mark.parametrize("mut", ["B.mutant = ON;", "C.mutant = OFF; E.mutant = ON;"])
def test_api_mutant_raw_matches_oracle(tmp_path, mut):
    bnd, cfg = _write(tmp_path, "m", _bnd(), CFG + mut + "\nsample_count = 20000;\n")
    print(f"[gpu] api {mut}", flush=True)
    mg.Model.load(bnd, cfg).compile().simulate().write_raw(str(tmp_path / "g.raw"))
    run_oracle(bnd, cfg, "--raw", tmp_path / "o.raw")
    assert read_raw(tmp_path / "g.raw") == read_raw(tmp_path / "o.raw")


def test_plain_mutants_equal_batched_mutant_masks(tmp_path):
    n = 20000
    bnd, cfg = _write(tmp_path, "m", _bnd(), CFG + "B.mutant = ON; D.mutant = OFF;\n")
    print("[gpu] plain mutant cfg vs batched masks", flush=True)
    plain = mg.Model.load(bnd, cfg).compile().simulate(n_trajectories=n, variant=1)
    plain.write_raw(str(tmp_path / "plain.raw"))
    base = mg.Model.load(*_write(tmp_path, "b", _bnd(), CFG))
    muts = [0, 1, 0, -1, 0]
    res, _ = base.compile().simulate_batched([{"n_trajectories": n}, {"n_trajectories": n, "mutations": muts}])
    res[1].write_raw(str(tmp_path / "batched.raw"))
    a, b = read_raw(tmp_path / "plain.raw"), read_raw(tmp_path / "batched.raw")
    assert a["W"] == b["W"] and a["E"] == b["E"] and a["F"] == b["F"] and a["P"] == b["P"]
