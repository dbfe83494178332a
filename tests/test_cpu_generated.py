"""Generated per-model code, run on the host, against the oracle.

The engine's generator emits HIP source; tests/native/gen_harness.cpp
compiles it for the host with the same float rules (-ffp-contract=off) and
evaluates total rate, transition entropy, flip node and observed index for
many states.  The oracle evaluates the same states with its own parser and
interpreter.  Both must agree bit for bit (both entropy formulas)."""
import os
import subprocess

import numpy as np
import pytest

import maboss.gpu_amd as mg
from maboss.gpu_amd import netgen
from conftest import DEVICE_INC, NATIVE, ORACLE, REF_MODELS, model_paths, pipe

HIPRTC_FLAGS = ["--offload-arch=gfx950", "-fgpu-rdc", "-O3", "-std=c++17", "-ffp-contract=off",
                "-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _models(tmp):
    out = [(n, *model_paths(n)) for n in REF_MODELS]
    for w in ("plain_small", "plain_100", "batched_50", "grad_30"):
        out.append((w, *netgen.write_workload(w, str(tmp))))
    return out


@pytest.fixture(scope="module")
def models(tmp_path_factory):
    return _models(tmp_path_factory.mktemp("wl"))


def _states(n_nodes, count, seed):
    rng = np.random.default_rng(seed)
    W = (n_nodes + 31) // 32
    if n_nodes <= 10:
        st = np.arange(1 << n_nodes, dtype=np.uint64)[:, None].astype(np.uint32)
    else:
        st = rng.integers(0, 2**32, size=(count, W), dtype=np.uint64).astype(np.uint32)
        if n_nodes % 32:
            st[:, -1] &= np.uint32((1 << (n_nodes % 32)) - 1)
    u2 = rng.uniform(0, 1, size=len(st)).astype(np.float32)
    u2[:3] = [np.float32(1.0), np.float32(2**-24), np.float32(0.5)]
    return st, u2


@pytest.mark.parametrize("entropy", [0, 1])
def test_generated_matches_oracle(models, native_bin, tmp_path, entropy):
    for name, bnd, cfg in models:
        m = mg.Model.load(bnd, cfg)
        gen = tmp_path / f"{name}_{entropy}.hip"
        gen.write_text(m.generate_code(entropy_formula=entropy))
        exe = native_bin("gen_harness", extra=[f'-DGEN_FILE="{gen}"'], tag=str(gen))
        st, u2 = _states(m.info.n_nodes, 3000, 5)
        text = "".join(" ".join(f"{w:x}" for w in row) + f" {u.view(np.uint32):08x}\n" for row, u in zip(st, u2))
        a = pipe(exe, text)
        b = pipe(ORACLE, text, "--eval-states", "--entropy", "reference" if entropy else "log_sum", bnd, cfg)
        assert len(a) == len(st)
        assert a == b, name


def test_generated_compiles_for_gfx950(models, tmp_path):
    """The generated source compiles for gfx950 with the JIT's flags (hipcc
    runs the same clang front-end as hipRTC)."""
    for name, bnd, cfg in models[:5]:
        src = tmp_path / f"{name}.hip"
        src.write_text(mg.Model.load(bnd, cfg).generate_code())
        p = subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-device-only", "-emit-llvm", "-c",
                            *HIPRTC_FLAGS, "-include", "hip/hip_runtime.h", f"-I{DEVICE_INC}", str(src), "-o", str(tmp_path / f"{name}.bc")],
                           capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-3000:]


def test_entropy_special_cases(native_bin, tmp_path):
    """Zero observed rate -> 0; one nonzero -> 0; k equal rates -> log2 k."""
    bnd = "".join(f"node X{i} {{ logic = !X{i}; rate_up = @logic ? $r{i} : 0; rate_down = @logic ? 0 : $r{i}; }}\n"
                  for i in range(6))
    for rates, expect in [([0, 0, 0, 0, 0, 0], 0.0), ([0, 3, 0, 0, 0, 0], 0.0), ([2] * 4 + [0, 0], 2.0),
                          ([1.5] * 3 + [0] * 3, np.log2(3)), ([7] * 6, np.log2(6))]:
        cfg = "".join(f"$r{i} = {r};\n" for i, r in enumerate(rates))
        m = mg.Model.from_strings(bnd, cfg)
        for ent in (0, 1):
            gen = tmp_path / f"e{ent}_{len(str(rates))}_{sum(rates)}.hip"
            gen.write_text(m.generate_code(entropy_formula=ent))
            exe = native_bin("gen_harness", extra=[f'-DGEN_FILE="{gen}"'], tag=str(gen))
            out = pipe(exe, "0 3f000000\n")[0].split()
            H = np.uint32(int(out[1], 16)).view(np.float32)
            assert abs(H - expect) <= 2 * np.spacing(np.float32(max(expect, 1.0))), (rates, ent, H)
            total = np.uint32(int(out[0], 16)).view(np.float32)
            assert total == np.float32(sum(np.float32(r) for r in rates))


def test_flip_selection_brute_force(native_bin, tmp_path):
    """flip = first node whose prefix sum reaches u2*total; zero-rate nodes
    are never chosen; u2 = 1 picks the last nonzero node."""
    rates = [0.0, 1.25, 0.0, 3.0, 0.5, 0.0, 2.0, 0.0]
    bnd = "".join(f"node X{i} {{ logic = !X{i}; rate_up = @logic ? $r{i} : 0; rate_down = @logic ? 0 : $r{i}; }}\n"
                  for i in range(8))
    cfg = "".join(f"$r{i} = {r};\n" for i, r in enumerate(rates))
    m = mg.Model.from_strings(bnd, cfg)
    gen = tmp_path / "flip.hip"
    gen.write_text(m.generate_code())
    exe = native_bin("gen_harness", extra=[f'-DGEN_FILE="{gen}"'], tag=str(gen))
    us = np.concatenate([np.linspace(1e-6, 1, 997, dtype=np.float32), [np.float32(1.0)]])
    out = pipe(exe, "".join(f"0 {u.view(np.uint32):08x}\n" for u in us))
    r32 = np.array(rates, dtype=np.float32)
    pref = np.cumsum(r32, dtype=np.float32)
    for u, line in zip(us, out):
        flip = int(line.split()[1 + 1])
        target = np.float32(u * pref[-1])
        expect = int(np.sum(pref < target))
        assert flip == expect
        assert r32[flip] > 0
    assert int(out[-1].split()[2]) == 6


def test_grouped_flip_selection_brute_force(native_bin, tmp_path):
    """Models above 40 nodes select the flip from groups of 8 nodes: the same
    node as a full sequential float32 prefix scan, for states and u2 values
    that put the target on group boundaries, inside runs of zero-rate nodes
    (+0 and -0 rates) and in the last, partial group.  The observed nodes'
    nonzero-rate count (entropy is 0 unless two or more are nonzero) is
    checked against the same brute force."""
    n = 45
    rng = np.random.default_rng(11)
    up = rng.lognormal(0, 1, n).astype(np.float32)
    dn = rng.lognormal(0, 1, n).astype(np.float32)
    up[[3, 8, 9, 10, 15, 16, 23, 40, 44]] = 0.0
    dn[[0, 7, 8, 24, 31, 32, 41]] = 0.0
    neg0 = {(5, "u"), (33, "d")}
    up[5] = dn[33] = np.float32(-0.0)
    lit = lambda v, neg: "-0" if neg else repr(float(v))
    bnd = "".join(f"node X{i} {{ logic = !X{i}; rate_up = @logic ? $u{i} : 0; rate_down = @logic ? 0 : $d{i}; }}\n"
                  for i in range(n))
    cfg = "".join(f"$u{i} = {lit(up[i], (i, 'u') in neg0)};\n$d{i} = {lit(dn[i], (i, 'd') in neg0)};\n"
                  for i in range(n))
    cfg += "".join(f"X{i}.is_internal = 1;\n" for i in range(0, n, 3))
    m = mg.Model.from_strings(bnd, cfg)
    gen = tmp_path / "grouped.hip"
    src = m.generate_code()
    assert "pass 2 (grouped)" in src
    gen.write_text(src)
    exe = native_bin("gen_harness", extra=[f'-DGEN_FILE="{gen}"'], tag=str(gen))
    states = rng.integers(0, 2, size=(400, n)).astype(np.uint8)
    states[0] = 0
    states[1] = 1
    lines, expect = [], []
    obs = np.array([i % 3 != 0 for i in range(n)])
    for st in states:
        r = np.where(st == 1, dn, up).astype(np.float32)  # every node is enabled (logic = !x)
        pref = np.cumsum(r, dtype=np.float32)
        total = pref[-1]
        # u2 values landing exactly on group-end prefix sums, just above and below them, and random
        cand = [np.float32(1.0), np.float32(2 ** -24)] + list(rng.uniform(0, 1, 6).astype(np.float32))
        for g in range(7, n, 8):
            if total > 0:
                q = np.float32(pref[g] / total)
                cand += [q, np.nextafter(q, np.float32(2)), np.nextafter(q, np.float32(0))]
        words = [int(sum(int(b) << j for j, b in enumerate(st[w * 32:(w + 1) * 32]))) for w in range((n + 31) // 32)]
        nnz = int(np.sum((r != 0) & obs))
        for u in cand:
            u = np.float32(min(max(u, np.float32(2 ** -24)), np.float32(1.0)))
            lines.append(" ".join(f"{x:x}" for x in words) + f" {u.view(np.uint32):08x}\n")
            expect.append((int(np.sum(pref < np.float32(u * total))), total, nnz))
    out = pipe(exe, "".join(lines))
    assert len(out) == len(expect)
    for line, (flip, total, nnz) in zip(out, expect):
        f = line.split()
        assert np.uint32(int(f[0], 16)).view(np.float32) == total
        assert int(f[2]) == flip, (line, flip)
        if nnz < 2:
            assert int(f[1], 16) == 0


def test_gradient_kernel_compiles_for_gfx950(tmp_path):
    """The gradient kernel's translation unit (generated model code + the
    kernel header) compiles for gfx950; models without a gradient form are
    refused with MBG_ERR_LIMIT."""
    d = tmp_path / "wl"
    d.mkdir()
    bnd, cfg = netgen.write_workload("grad_30", str(d))
    src = tmp_path / "grad.hip"
    src.write_text(mg.Model.load(bnd, cfg).generate_code(with_gradient=True))
    flags = [f for f in HIPRTC_FLAGS if f != "-fgpu-rdc"]
    p = subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-device-only", "-c", *flags,
                        "-include", "hip/hip_runtime.h", f"-I{DEVICE_INC}", str(src), "-o", str(tmp_path / "grad.o")],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    # cellcycle runs in discrete time: no gradient form
    with pytest.raises(mg.MbgError) as e:
        mg.Model.load(*model_paths("cellcycle")).generate_code(with_gradient=True)
    assert e.value.code == 5


def test_single_unit_kernels_compile_with_hiprtc(native_bin, tmp_path):
    """The engine's hipRTC-only compile paths: lane kernels as one translation
    unit (used when the process's code-object manager cannot read the
    embedded bitcode) and the gradient kernel, compiled by hipRTC itself."""
    exe = native_bin("rtc_compile", extra=["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-L/opt/rocm/lib",
                                           "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"])
    d = tmp_path / "wl"
    d.mkdir()
    bnd, cfg = netgen.write_workload("grad_30", str(d))
    m = mg.Model.load(bnd, cfg)
    lane = open(os.path.join(DEVICE_INC, "simulate_lane.hip")).read()
    headers = [os.path.join(DEVICE_INC, h) for h in ("mbg_numeric.h", "mbg_abi.h", "mbg_kernel_args.h",
                                                     "mbg_stats_device.h", "mbg_lane_core.h", "mbg_grad_kernel.h")]
    # the plain flavour with hashed window cells and dense final states (MBG_HIST_HASH, MBG_HIST_DENSE)
    modes = {0: "#define MBG_WIN_MODE 1\n#define MBG_FIN_MODE 0\n", 1: ""}
    units = {f"lane_p{p}": f"#define MBG_KW {m.info.n_words}\n#define MBG_PARAM {p}\n" + modes[p] + m.generate_code()
             + "\n" + lane for p in (0, 1)}
    units["grad"] = m.generate_code(with_gradient=True)
    for name, text in units.items():
        src = tmp_path / f"{name}.hip"
        src.write_text(text)
        p = subprocess.run([exe, str(src), *headers], capture_output=True, text=True, timeout=600)
        assert p.returncode == 0 and p.stdout.startswith("OK"), (name, p.stderr[-3000:])


def _mutant_masks(m, rng, k):
    """k random mutant nodes: (cfg text, up_en words, dn_en words)."""
    N, W = m.info.n_nodes, m.info.n_words
    up = [0xFFFFFFFF] * W
    dn = [0xFFFFFFFF] * W
    lines = []
    names = m.node_names
    for i in rng.choice(N, size=k, replace=False):
        on = bool(rng.integers(0, 2))
        if on:
            dn[i // 32] &= ~(1 << (i % 32))
        else:
            up[i // 32] &= ~(1 << (i % 32))
        lines.append(f"{names[i]}.mutant = {'ON' if on else 'OFF'};\n")
    return "".join(lines), up, dn


@pytest.mark.parametrize("workload", ["plain_100", "batched_50", "grad_30", "cellcycle"])
def test_generated_runtime_rates_with_mutant_masks(native_bin, tmp_path, workload):
    """mbg_eval_p (rates from a table, mutant masks as up_en / dn_en words),
    run on the host, equals the oracle's evaluation of the same model with
    those nodes written as mutants in the .cfg (rate_down or rate_up 0) bit
    for bit; and its enabled words equal, for every node, whether the
    engine's interpreter gives that mutant model a nonzero rate."""
    if workload in REF_MODELS:
        bnd, cfg = model_paths(workload)
    else:
        bnd, cfg = netgen.write_workload(workload, str(tmp_path))
    m = mg.Model.load(bnd, cfg)
    rng = np.random.default_rng(11)
    mut, up, dn = _mutant_masks(m, rng, min(6, m.info.n_nodes // 3))
    mcfg = tmp_path / "mutant.cfg"
    mcfg.write_text(open(cfg).read() + "\n" + mut)
    gen = tmp_path / f"{workload}_p.hip"
    gen.write_text(m.generate_code())
    exe = native_bin("gen_harness", extra=[f'-DGEN_FILE="{gen}"', "-DGEN_PARAM"], tag=str(gen))
    W, N = m.info.n_words, m.info.n_nodes
    st, u2 = _states(N, 2000, 7)
    pv = m.param_values
    head = f"{len(pv)} " + " ".join(f"{v.view(np.uint32):x}" for v in pv) + " " + \
        " ".join(f"{w:x}" for w in up) + " " + " ".join(f"{w:x}" for w in dn) + "\n"
    text = "".join(" ".join(f"{w:x}" for w in row) + f" {u.view(np.uint32):08x}\n" for row, u in zip(st, u2))
    a = pipe(exe, head + text)
    b = pipe(ORACLE, text, "--eval-states", bnd, str(mcfg))
    assert len(a) == len(b) == len(st)
    assert [" ".join(x.split()[:4]) for x in a] == b
    mm = mg.Model.load(bnd, str(mcfg))
    for row, line in list(zip(st, a))[:200]:
        en = [int(x, 16) for x in line.split()[4:]]
        words = [int(x) for x in row]
        for i in range(N):
            bit = (en[i // 32] >> (i % 32)) & 1
            assert bit == (mm.rate(i, words) != 0.0), (workload, i)
