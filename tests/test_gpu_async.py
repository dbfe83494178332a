"""Runs whose statistics are copied out while the next run computes.

mbg_simulate_async returns once a run is enqueued; its compacted, sorted
cells reach pinned host memory on the program's copy stream while the
caller's stream moves on.  These tests check that overlapping runs, runs on
non-blocking streams and results freed before they complete give exactly the
statistics of one synchronous run at a time."""
import pytest

import maboss.gpu_amd as mg
from maboss.gpu_amd import netgen
from conftest import model_paths, run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl(tmp_path_factory):
    d = tmp_path_factory.mktemp("wla")
    return {w: netgen.write_workload(w, str(d)) for w in ("plain_100", "plain_small")}


def raw_text(r, path):
    r.write_raw(str(path))
    return open(path).read()


@pytest.mark.parametrize("case,win", [("plain_100", "auto"), ("plain_100", "sort"), ("cellcycle", "auto"),
                                      ("plain_small", "hash")])
def test_overlapped_runs_equal_sync_runs(tmp_path, wl, case, win):
    bnd, cfg = model_paths(case) if case == "cellcycle" else wl[case]
    print(f"[gpu] async {case} window_mode={win}", flush=True)
    prog = mg.Model.load(bnd, cfg).compile()
    shards = [(0, 3000), (3000, 2000), (5000, 4000), (9000, 1000)]
    pending = [prog.simulate(n_trajectories=n, first_id=f, window_mode=win, wait=False) for f, n in shards]
    got = [raw_text(r, tmp_path / f"a{i}.raw") for i, r in enumerate(pending)]
    for i, (f, n) in enumerate(shards):
        want = raw_text(prog.simulate(n_trajectories=n, first_id=f, window_mode=win), tmp_path / f"s{i}.raw")
        assert got[i] == want, (case, win, i)
    # and the first shard against the oracle
    run_oracle(bnd, cfg, "--n", shards[0][1], "--raw", tmp_path / "o.raw")
    assert got[0] == open(tmp_path / "o.raw").read()


def test_non_blocking_stream(tmp_path, wl):
    """Hashed window cells on a stream that does not synchronise with the null stream."""
    import torch
    bnd, cfg = wl["plain_100"]
    print("[gpu] side stream plain_100", flush=True)
    prog = mg.Model.load(bnd, cfg).compile()
    base = raw_text(prog.simulate(n_trajectories=4000), tmp_path / "base.raw")
    s = torch.cuda.Stream()
    for wait in (True, False):
        r = prog.simulate(n_trajectories=4000, stream=s.cuda_stream, wait=wait)
        assert r.info.window_mode == 1
        assert raw_text(r, tmp_path / f"side{int(wait)}.raw") == base


def test_results_freed_before_completion(tmp_path, wl):
    """Dropping results whose copies are in flight; the pinned blocks are reused."""
    bnd, cfg = wl["plain_100"]
    print("[gpu] dropped async results plain_100", flush=True)
    prog = mg.Model.load(bnd, cfg).compile()
    for _ in range(4):
        r = prog.simulate(n_trajectories=3000, wait=False)
        del r
    last = prog.simulate(n_trajectories=3000, wait=False)
    want = raw_text(prog.simulate(n_trajectories=3000), tmp_path / "w.raw")
    assert raw_text(last, tmp_path / "l.raw") == want
    # counters are final when the async call returns
    r = prog.simulate(n_trajectories=3000, wait=False)
    assert r.info.transitions == last.info.transitions


def test_program_freed_with_copy_in_flight(tmp_path, wl):
    bnd, cfg = wl["plain_100"]
    print("[gpu] program freed with a copy in flight", flush=True)
    prog = mg.Model.load(bnd, cfg).compile()
    r = prog.simulate(n_trajectories=3000, wait=False)
    del prog
    r.wait()
    assert r.info.finished == 3000
