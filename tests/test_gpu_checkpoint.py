"""Checkpoint and resume at trajectory-id granularity (the one-GPU part of
sharding).  The reference's runner keeps state only between its epochs
(src/simulation_runner.cu:123-161).  Here a saved result, reloaded and merged
with the run of the following ids, and a run in chunks that is stopped and
started again on the same checkpoint file, must both equal the single run of
the whole id range bit for bit (window cells, entropy sums, final states,
fixed points, counters, second moments)."""
import os

import pytest

import maboss.gpu_amd as mg
from conftest import model_paths, read_raw

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    bnd, cfg = model_paths("cellcycle")
    d = tmp_path_factory.mktemp("ck")
    c = d / "cc.cfg"
    c.write_text(open(cfg).read() + "\ndiscrete_time = 0;\n")
    return mg.Model.load(bnd, str(c)).compile()


def _raw(tmp_path, r, tag):
    r.write_raw(str(tmp_path / f"{tag}.raw"))
    return read_raw(tmp_path / f"{tag}.raw")


@pytest.mark.parametrize("errors", [False, True])
def test_save_load_merge_equals_single_run(tmp_path, prog, errors):
    n = 30000
    print(f"[gpu] checkpoint merge errors={errors}", flush=True)
    full = _raw(tmp_path, prog.simulate(n_trajectories=n, errors=errors), "full")
    a = prog.simulate(n_trajectories=11000, errors=errors)
    a.save(str(tmp_path / "a.ckpt"))
    back = prog.load_result(str(tmp_path / "a.ckpt"))
    assert _raw(tmp_path, back, "back") == _raw(tmp_path, a, "a")
    b = prog.simulate(n_trajectories=n - 11000, first_id=11000, errors=errors)
    back.merge(b)
    assert back.info.n_trajectories == n
    assert _raw(tmp_path, back, "merged") == full


def test_chunked_run_stopped_and_resumed(tmp_path, prog):
    n, path = 40000, str(tmp_path / "run.ckpt")
    print("[gpu] checkpointed run", flush=True)
    full = _raw(tmp_path, prog.simulate(n_trajectories=n, errors=True), "full")
    # a first process covers 17000 ids (in chunks of 6000) and stops
    part = prog.simulate_checkpointed(path, chunk=6000, n_trajectories=17000, errors=True)
    assert part.info.n_trajectories == 17000 and os.path.exists(path)
    # started again with the whole range: resumes after id 17000
    done = prog.simulate_checkpointed(path, chunk=6000, n_trajectories=n, errors=True)
    assert done.info.n_trajectories == n
    assert _raw(tmp_path, done, "resumed") == full
    # the file now covers the whole range: a third call does no work and returns it
    again = prog.simulate_checkpointed(path, chunk=6000, n_trajectories=n, errors=True)
    assert _raw(tmp_path, again, "again") == full


def test_checkpoint_mismatches_are_refused(tmp_path, prog):
    path = str(tmp_path / "s.ckpt")
    prog.simulate_checkpointed(path, n_trajectories=2000)
    with pytest.raises(mg.MbgError):  # another seed
        prog.simulate_checkpointed(path, n_trajectories=4000, seed=7)
    with pytest.raises(mg.MbgError):  # other ids
        prog.simulate_checkpointed(path, n_trajectories=4000, first_id=100)
    a = prog.simulate(n_trajectories=1000)
    with pytest.raises(mg.MbgError):  # not adjacent
        a.merge(prog.simulate(n_trajectories=1000, first_id=5000))
    data = open(path, "rb").read()
    with open(tmp_path / "cut.ckpt", "wb") as f:
        f.write(data[: len(data) // 2])
    with pytest.raises(mg.MbgError):  # truncated
        prog.load_result(str(tmp_path / "cut.ckpt"))
