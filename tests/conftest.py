import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

MODELS = os.path.join(ROOT, "tests", "golden", "models")
RES = os.path.join(ROOT, "tests", "golden", "res")
ORACLE = os.path.join(ROOT, "oracle", "build", "maboss_oracle")
LIB = os.path.join(ROOT, "maboss", "gpu_amd", "lib", "libmaboss_gpu.so")
REF_MODELS = ["cellcycle", "metastasis", "sizek", "Montagud2022_Prostate_Cancer"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def model_paths(name):
    return os.path.join(MODELS, name + ".bnd"), os.path.join(MODELS, name + ".cfg")


def run_oracle(bnd, cfg, *args, timeout=600):
    """Run the host oracle (the checker); returns the completed process."""
    if not os.path.exists(ORACLE):
        pytest.fail(f"{ORACLE} missing: run __graft_entry__.build()")
    p = subprocess.run([ORACLE, *map(str, args), bnd, cfg], capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f"oracle failed: {p.stderr}")
    return p


def read_raw(path):
    """Parse an exact integer dump (engine or oracle) into a dict."""
    out = {"W": {}, "E": {}, "F": {}, "P": {}, "Q": {}, "R": {}}
    with open(path) as f:
        assert f.readline().strip() == "MBG_RAW 1"
        head = f.readline().split()
        out["header"] = dict(zip(head[::2], map(int, head[1::2])))
        for line in f:
            t = line.split()
            if t[0] == "T":
                out["T"] = tuple(map(int, t[1:]))
            elif t[0] == "E":
                out["E"][int(t[1])] = int(t[2])
            elif t[0] == "W":
                out["W"][(int(t[1]), int(t[2]))] = int(t[3])
            elif t[0] == "F":
                out["F"][int(t[1])] = int(t[2])
            elif t[0] == "P":
                out["P"][tuple(map(int, t[1:-1]))] = int(t[-1])
            elif t[0] == "Q":
                out["Q"][(int(t[1]), int(t[2]))] = (int(t[3]) << 64) | int(t[4])
            elif t[0] == "R":
                out["R"][int(t[1])] = (int(t[2]) << 64) | int(t[3])
    return out


@pytest.fixture(scope="session")
def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def workdir(tmp_path_factory):
    return tmp_path_factory.mktemp("mbg")


NATIVE = os.path.join(ROOT, "tests", "native")
DEVICE_INC = os.path.join(ROOT, "maboss", "gpu_amd", "csrc", "device")


@pytest.fixture(scope="session")
def native_bin(tmp_path_factory):
    """Compile a test probe from tests/native; returns a function name -> path."""
    out = tmp_path_factory.mktemp("native")
    cache = {}

    def build(name, extra=(), compiler="g++", tag=None):
        key = (name, tuple(extra), tag)
        if key in cache:
            return cache[key]
        exe = str(out / f"{name}_{len(cache)}")
        cmd = [compiler, "-O2", "-std=c++17", "-ffp-contract=off", f"-I{DEVICE_INC}",
               os.path.join(NATIVE, name + ".cpp"), "-o", exe, *extra]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            pytest.fail(f"building {name} failed: {p.stderr[-2000:]}")
        cache[key] = exe
        return exe

    return build


def pipe(exe, text, *args, timeout=300):
    p = subprocess.run([exe, *args], input=text, capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(p.stderr)
    return p.stdout.split("\n")[:-1]
