"""Pins of the host oracle (the checker the GPU engine is compared with).

Each check is fixed by mathematics or by the reference's own captured
results, not by the oracle's code: closed-form Markov chains, exact counts in
discrete time, conservation invariants, fixed points verified by the engine's
separate expression interpreter, and 3-sigma binomial agreement with the
reference's .res captures (test/*.res of MaBoSS.GPU).  The closed forms and
the .res comparisons run in both summation orders the oracle has: the lane
kernel's sequential order and the wave kernel's lane-slice + scan order
(--order wave), so the wave branch is pinned by the same mathematics."""
import math
import os

import numpy as np
import pytest

import maboss.gpu_amd as mg
from conftest import RES, model_paths, read_raw, run_oracle
from resfile import parse_report, zscores

TWO32 = 4294967296.0
ORDERS = pytest.mark.parametrize("order", ["seq", "wave"])


def _write(tmp_path, name, bnd, cfg):
    b, c = tmp_path / f"{name}.bnd", tmp_path / f"{name}.cfg"
    b.write_text(bnd)
    c.write_text(cfg)
    return str(b), str(c)


def _window_probs(raw, n, w, discrete=False):
    scale = n if discrete else n * TWO32
    return {s: v / scale for (ww, s), v in raw["W"].items() if ww == w}


@ORDERS
def test_toggle_closed_form(tmp_path, order):
    lam, p0, T, tick, n = 2.0, 0.2, 2.0, 0.25, 200000
    b, c = _write(tmp_path, "toggle", "node A { logic = !A; rate_up = @logic ? $l : 0; rate_down = @logic ? 0 : $l; }",
                  f"$l = {lam}; A.istate = {p0}; max_time = {T}; time_tick = {tick}; sample_count = {n};")
    run_oracle(b, c, "--raw", tmp_path / "t.raw", "--threads", 4, "--order", order)
    raw = read_raw(tmp_path / "t.raw")
    for w in range(8):
        a_, b_ = w * tick, (w + 1) * tick
        exact = 0.5 + (p0 - 0.5) * (math.exp(-2 * lam * a_) - math.exp(-2 * lam * b_)) / (2 * lam * (b_ - a_))
        got = _window_probs(raw, n, w).get(1, 0.0)
        assert abs(got - exact) < 5 * math.sqrt(exact * (1 - exact) / n), (w, got, exact)
        assert abs(sum(_window_probs(raw, n, w).values()) - 1) < 1e-6
    assert not raw["E"]  # one observed node: transition entropy 0
    assert raw["T"][1] == n and sum(raw["F"].values()) == n


@ORDERS
def test_decay_closed_form_and_fixed_points(tmp_path, order):
    lam, T, tick, n = 0.7, 3.0, 0.5, 200000
    b, c = _write(tmp_path, "decay", "node A { rate_up = 0; rate_down = $l; }\nnode B { rate_up = 0; rate_down = 0; }",
                  f"$l = {lam}; A.istate = 1; B.istate = 0; B.is_internal = 1; max_time = {T}; time_tick = {tick}; "
                  f"sample_count = {n};")
    run_oracle(b, c, "--raw", tmp_path / "d.raw", "--threads", 4, "--order", order)
    raw = read_raw(tmp_path / "d.raw")
    for w in range(6):
        a_, b_ = w * tick, (w + 1) * tick
        exact = (math.exp(-lam * a_) - math.exp(-lam * b_)) / (lam * (b_ - a_))
        got = _window_probs(raw, n, w).get(1, 0.0)
        assert abs(got - exact) < 5 * math.sqrt(exact * (1 - exact) / n)
    off = raw["F"].get(0, 0)
    assert abs(off / n - (1 - math.exp(-lam * T))) < 5 * math.sqrt(0.25 / n)
    # every trajectory that ended off is a fixed point, and nothing else is
    assert raw["P"] == {(0,): off}
    assert raw["F"].get(0, 0) + raw["F"].get(1, 0) == n


@ORDERS
def test_discrete_exact_counts(tmp_path, order):
    n = 1000
    b, c = _write(tmp_path, "disc", "node A { logic = !A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }",
                  f"A.istate = 0; max_time = 5; time_tick = 1; discrete_time = 1; sample_count = {n};")
    run_oracle(b, c, "--raw", tmp_path / "x.raw", "--order", order)
    raw = read_raw(tmp_path / "x.raw")
    expect = {(w, w % 2): n for w in range(5)}
    assert raw["W"] == expect
    assert raw["T"] == (5 * n, n, 0)
    assert raw["F"] == {0: n}  # state of the last window: no flip after the last record


@ORDERS
def test_two_equal_togglers_entropy_one_bit(tmp_path, order):
    n, T = 20000, 4.0
    bnd = "".join(f"node {x} {{ logic = !{x}; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }}\n"
                  for x in "AB")
    b, c = _write(tmp_path, "two", bnd, f"max_time = {T}; time_tick = 1; sample_count = {n};")
    run_oracle(b, c, "--raw", tmp_path / "e.raw", "--order", order)
    raw = read_raw(tmp_path / "e.raw")
    for w in range(4):
        assert abs(raw["E"][w] / (n * TWO32) - 1.0) < 1e-7
        # uniform stationary distribution reached from istate 0.5
        for s in range(4):
            assert abs(_window_probs(raw, n, w)[s] - 0.25) < 5 * math.sqrt(0.1875 / n)


@pytest.mark.parametrize("name,n", [("cellcycle", 20000), ("sizek", 5000), ("metastasis", 5000),
                                    ("Montagud2022_Prostate_Cancer", 3000)])
def test_invariants(tmp_path, name, n):
    bnd, cfg = model_paths(name)
    run_oracle(bnd, cfg, "--n", n, "--raw", tmp_path / "r.raw")
    raw = read_raw(tmp_path / "r.raw")
    nw, disc = raw["header"]["windows"], raw["header"]["discrete"]
    assert raw["T"][1] == n and raw["T"][2] == 0 and sum(raw["F"].values()) == n
    assert raw["T"][0] >= n
    m = mg.Model.load(bnd, cfg)
    tick = np.float32(m.info.time_tick)
    mt = np.float32(m.info.max_time)
    for w in range(nw):
        tot = sum(v for (ww, _), v in raw["W"].items() if ww == w)
        if disc:
            assert tot == n  # cellcycle has no fixed points: one step per window
        else:
            # each trajectory covers the window's float32 extent, as the
            # window loop computes its boundaries: w*tick .. min((w+1)*tick, max_time)
            lo = np.float32(w) * tick if w else np.float32(0)
            hi = min(np.float32(w + 1) * tick, mt)
            cover = float(np.float32(hi - lo)) / float(tick)
            assert abs(tot / (n * TWO32) - cover) < 2e-6 * max(1.0, w * float(tick))
    # fixed points: every one has total rate 0 (engine's separate interpreter)
    for words in list(raw["P"])[:50]:
        assert all(m.rate(i, list(words)) == 0.0 for i in range(m.info.n_nodes))
    if name == "cellcycle":
        assert not raw["P"]


RES_CASES = [
    # name, oracle trajectories, compare final points
    ("cellcycle", 100000, False),   # capture's final points disagree with its own last window
    ("sizek", 40000, True),
    ("metastasis", 20000, True),
    ("Montagud2022_Prostate_Cancer", 20000, True),
]


@ORDERS
@pytest.mark.parametrize("name,n,finals", RES_CASES)
def test_reference_res_3sigma(tmp_path, name, n, finals, order):
    bnd, cfg = model_paths(name)
    run_oracle(bnd, cfg, "--n", n, "--compiled", "--order", order, "--print", tmp_path / "o.txt", timeout=900)
    ref = parse_report(os.path.join(RES, name + ".res.txt"))
    new = parse_report(str(tmp_path / "o.txt"))
    n_ref = mg.Model.load(bnd, cfg).info.sample_count
    assert len(ref["windows"]) == len(new["windows"])
    zs = []
    for wr, wn in zip(ref["windows"], new["windows"]):
        assert wr["a"] == pytest.approx(wn["a"]) and wr["b"] == pytest.approx(wn["b"])
        zs += zscores(wr["probs"], wn["probs"], n_ref, n)
        nobs = max(1, max((len(k) for k in wr["probs"]), default=1))
        tol = 5 * max(1.0, math.log2(1 + nobs)) * math.sqrt(1.0 / n_ref + 1.0 / n) + 1e-3
        assert abs(wr["TH"] - wn["TH"]) < tol
    if finals:
        zs += zscores(ref["final"], new["final"], n_ref, n)
        zs += zscores(ref["fixed"], new["fixed"], n_ref, n)
    z = np.array([x[0] for x in zs])
    worst = max(zs)
    assert (z > 3).mean() <= 0.01, (name, (z > 3).mean(), worst)
    assert z.max() < 5, (name, worst)


# ---- the oracle's likelihood-ratio gradient (--grad-in) against closed forms

TOGGLE_UD = """Node A {
	logic = !A;
	rate_up = @logic ? $u : 0;
	rate_down = @logic ? 0 : $d;
}
"""
DECAY_D = """Node A {
	logic = !A;
	rate_up = @logic ? 0 : 0;
	rate_down = @logic ? 0 : $d;
}
"""


def _grad_batches(tmp_path, b, c, G, n_batch, k_batches):
    gpath = tmp_path / "G.bin"
    np.ascontiguousarray(G, dtype=np.float64).tofile(gpath)
    vals = []
    for i in range(k_batches):
        out = tmp_path / f"g{i}.txt"
        run_oracle(b, c, "--n", n_batch, "--first-id", i * n_batch, "--threads", 8, "--grad-in", gpath,
                   "--grad-out", out)
        d = dict(line.split() for line in open(out))
        vals.append({k: float(v) for k, v in d.items() if k != "transitions"})
    names = sorted(vals[0])
    arr = np.array([[v[k] for k in names] for v in vals])
    return names, arr.mean(0), arr.std(0, ddof=1) / math.sqrt(k_batches)


def test_oracle_gradient_toggle_closed_form(tmp_path):
    """Two-state chain, u (off->on), d (on->off): P(on, t) = pi + (p0 - pi) e^{-(u+d) t}, pi = u/(u+d);
    L = sum_w P_w(on), differentiated analytically (central differences of the closed form)."""
    u, d, p0, T, tick = 1.25, 0.5, 0.3, 4.0, 0.5
    b, c = _write(tmp_path, "tog", TOGGLE_UD, f"$u = {u}; $d = {d}; A.istate = {p0}; sample_count = 10; "
                  f"max_time = {T}; time_tick = {tick}; discrete_time = 0; seed_pseudorandom = 7;")
    W = int(round(T / tick))
    G = np.zeros((W, 2))
    G[:, 1] = 1.0

    def L(u, d):
        lam = u + d
        pi = u / lam
        return sum(pi + (p0 - pi) * (math.exp(-lam * w * tick) - math.exp(-lam * (w + 1) * tick)) / (lam * tick)
                   for w in range(W))

    h = 1e-6
    exact = {"$u": (L(u + h, d) - L(u - h, d)) / (2 * h), "$d": (L(u, d + h) - L(u, d - h)) / (2 * h)}
    names, mean, se = _grad_batches(tmp_path, b, c, G, 250000, 8)
    for k in ("$u", "$d"):
        j = names.index(k)
        assert abs(mean[j] - exact[k]) < 4.5 * se[j], (k, mean[j], exact[k], se[j])
        assert se[j] < 0.02 * abs(exact[k])


def test_oracle_gradient_decay_closed_form(tmp_path):
    """Pure decay from on: P(on, t) = e^{-d t}, with window weights 1..2 on P_w(on)."""
    d, T, tick = 0.8, 3.0, 0.25
    b, c = _write(tmp_path, "dec", DECAY_D, f"$d = {d}; A.istate = 1; sample_count = 10; max_time = {T}; "
                  f"time_tick = {tick}; discrete_time = 0; seed_pseudorandom = 7;")
    W = int(round(T / tick))
    G = np.zeros((W, 2))
    G[:, 1] = np.linspace(1.0, 2.0, W)

    def L(d):
        return sum(G[w, 1] * (math.exp(-d * w * tick) - math.exp(-d * (w + 1) * tick)) / (d * tick) for w in range(W))

    exact = (L(d + 1e-6) - L(d - 1e-6)) / 2e-6
    names, mean, se = _grad_batches(tmp_path, b, c, G, 250000, 8)
    assert names == ["$d"]
    assert abs(mean[0] - exact) < 4.5 * se[0], (mean[0], exact, se[0])
    assert se[0] < 0.02 * abs(exact)


@pytest.mark.parametrize("order", ["seq", "wave"])
def test_compiled_formulas_equal_interpreter(tmp_path, order):
    """--compiled (formulas emitted as C++ and built with g++, the CPU
    baseline's mode) gives the interpreter's exact dump."""
    from maboss.gpu_amd import netgen
    for w in ("plain_100", "plain_small"):
        bnd, cfg = netgen.write_workload(w, str(tmp_path))
        run_oracle(bnd, cfg, "--n", 400, "--order", order, "--raw", tmp_path / "i.raw")
        run_oracle(bnd, cfg, "--n", 400, "--order", order, "--compiled", "--raw", tmp_path / "c.raw")
        assert open(tmp_path / "i.raw").read() == open(tmp_path / "c.raw").read(), w
