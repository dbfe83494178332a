"""The on-disk code-object cache of the JIT (hipRTC) compiler.

Every program compile goes through it: a miss compiles and stores a code
object keyed by a hash of the source, flags, embedded headers and compiler
and runtime versions; a hit loads it; a file that does not load (damaged)
is removed and compiled again.  Each step runs in a fresh process with its
own MBG_CACHE_DIR, so the hits are real disk hits."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT, model_paths

pytestmark = pytest.mark.gpu


def _run(cache, raw, entropy=0):
    bnd, cfg = model_paths("cellcycle")
    code = ("import maboss.gpu_amd as mg; "
            f"p = mg.Model.load({bnd!r}, {cfg!r}).compile(entropy_formula={entropy}); "
            f"p.simulate(n_trajectories=2000).write_raw({str(raw)!r}); print('ok')")
    env = dict(os.environ, MBG_CACHE_DIR=str(cache))
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, cwd=ROOT, env=env)
    assert p.returncode == 0 and "ok" in p.stdout, p.stderr[-3000:]
    return open(raw).read()


def _objects(cache):
    return {f: os.path.getsize(os.path.join(cache, f)) for f in os.listdir(cache) if f.endswith(".co")}


def test_code_object_cache(tmp_path):
    cache = tmp_path / "cache"
    print("[gpu] cache miss", flush=True)
    first = _run(cache, tmp_path / "a.raw")
    files = _objects(cache)
    assert files, "a compile stores its code object"

    print("[gpu] cache hit", flush=True)
    assert _run(cache, tmp_path / "b.raw") == first
    assert _objects(cache) == files  # nothing compiled anew

    print("[gpu] flag change", flush=True)
    _run(cache, tmp_path / "c.raw", entropy=1)
    changed = _objects(cache)
    assert set(files) < set(changed)  # another key for the other entropy form

    print("[gpu] damaged cache files", flush=True)
    for f in files:
        # no ELF magic: the loader rejects the image before reading anything else
        with open(os.path.join(cache, f), "wb") as fh:
            fh.write(b"\0" * 32)
    assert _run(cache, tmp_path / "d.raw") == first
    again = _objects(cache)
    assert all(again[f] == files[f] for f in files)  # recompiled and stored again
