"""ErrorTH / ErrorProba from exact per-trajectory second moments (oracle).

The reference prints both columns as 0 (src/statistics/window_average_small.cpp:
235, :244-245).  With --errors the oracle sums, per window w and observed state
x, P^2 over trajectories (P = the trajectory's fixed-point time in x during w,
all its visits to x in w together) and T^2 (T = its entropy sum in w), as
exact 128-bit integers, and prints the standard error of the mean
sqrt((E[p^2] - E[p]^2) / (n - 1)).  Pinned here by what the mathematics fixes:
- no transitions and a Bernoulli initial state: every p is 0 or 1, so
  sum P^2 = 2^32 sum P exactly and the error is sqrt(p(1 - p) / (n - 1));
- a decaying node (time to decay ~ Exp(lambda)): p = clamp(tau - w, 0, 1),
  E[p] = e^{-lambda w} (1 - e^{-lambda}) / lambda,
  E[p^2] = 2 e^{-lambda w} (1 - e^{-lambda} (1 + lambda)) / lambda^2;
- a stationary two-state toggle (rates a up, b down, k = a + b, pi = a / k):
  Var(occupation fraction of a unit window) = 2 pi (1 - pi) (k - 1 + e^{-k}) / k^2
  (the covariance pi (1 - pi) e^{-k s} integrated over the window).  Squaring
  each dwell slice instead of the trajectory's window total drops the cross
  terms of revisits and fails this one;
- discrete time, one step per window: every P is 0 or 1, sum P^2 = sum P."""
import math

import numpy as np
import pytest

from conftest import read_raw, run_oracle

TWO32 = 2.0 ** 32


def _model(tmp_path, tag, nodes, cfg):
    b, c = tmp_path / f"{tag}.bnd", tmp_path / f"{tag}.cfg"
    b.write_text("\n".join(f"Node {n} {{ logic = {lg}; rate_up = {u}; rate_down = {d}; }}" for n, lg, u, d in nodes)
                 + "\n")
    c.write_text(cfg)
    return str(b), str(c)


def _read_q(path):
    q, r = {}, {}
    for line in open(path):
        t = line.split()
        if t[0] == "Q":
            q[(int(t[1]), int(t[2]))] = (int(t[3]) << 64) | int(t[4])
        elif t[0] == "R":
            r[int(t[1])] = (int(t[2]) << 64) | int(t[3])
    return q, r


def _csv(path):
    rows = []
    for line in open(path).read().splitlines()[1:]:
        f = line.split("\t")
        cells = {f[i]: (float(f[i + 1]), float(f[i + 2])) for i in range(5, len(f), 3)}
        rows.append({"TH": float(f[1]), "ErrTH": float(f[2]), "cells": cells})
    return rows


def _run(tmp_path, bnd, cfg, n, errors=True):
    args = ["--n", n, "--raw", tmp_path / "o.raw", "-o", tmp_path / "o"]
    run_oracle(bnd, cfg, *(args + (["--errors"] if errors else [])))
    return read_raw(tmp_path / "o.raw"), _read_q(tmp_path / "o.raw"), _csv(tmp_path / "o_probtraj.csv")


def test_bernoulli_initial_state_without_transitions(tmp_path):
    n = 50000
    bnd, cfg = _model(tmp_path, "b", [("A", "A", "0", "0"), ("B", "B", "0", "0")],
                      "A.istate = 0.3; B.istate = 0.75; max_time = 3; time_tick = 1;\n")
    raw, (q, r), rows = _run(tmp_path, bnd, cfg, n)
    assert raw["W"] and set(q) == set(raw["W"])
    for key, s1 in raw["W"].items():
        assert q[key] == s1 << 32  # P in {0, 2^32}: P^2 = 2^32 P
    for w, row in enumerate(rows):
        assert row["ErrTH"] == 0.0 and row["TH"] == 0.0
        for name, (p, err) in row["cells"].items():
            exact = math.sqrt(p * (1 - p) / (n - 1))
            assert err == pytest.approx(exact, rel=1e-5), (w, name)


def test_errors_off_by_default(tmp_path):
    bnd, cfg = _model(tmp_path, "b", [("A", "!A", "1", "2")], "A.istate = 0.5; max_time = 2; time_tick = 1;\n")
    raw, (q, r), rows = _run(tmp_path, bnd, cfg, 2000, errors=False)
    assert not q and not r
    assert all(row["ErrTH"] == 0.0 and all(e == 0.0 for _, e in row["cells"].values()) for row in rows)


def _var_hat(err, n):
    # ErrorProba^2 (n - 1) = E[p^2] - E[p]^2 over the sample
    return err * err * (n - 1)


def test_decay_closed_form(tmp_path):
    n, lam = 200000, 1.0
    bnd, cfg = _model(tmp_path, "d", [("A", "A", "0", f"{lam}")], "A.istate = 1; max_time = 3; time_tick = 1;\n")
    raw, (q, r), rows = _run(tmp_path, bnd, cfg, n)
    for w in (0, 1):
        p, err = rows[w]["cells"]["A"]
        m1 = math.exp(-lam * w) * (1 - math.exp(-lam)) / lam
        m2 = 2 * math.exp(-lam * w) * (1 - math.exp(-lam) * (1 + lam)) / lam ** 2
        assert p == pytest.approx(m1, rel=0.01)
        assert _var_hat(err, n) == pytest.approx(m2 - m1 * m1, rel=0.04), w


def test_stationary_toggle_closed_form(tmp_path):
    n, a, b = 200000, 2.0, 3.0
    k, pi = a + b, a / (a + b)
    bnd, cfg = _model(tmp_path, "t", [("A", "!A", f"@logic ? {a} : 0", f"@logic ? 0 : {b}")],
                      f"A.istate = {pi}; max_time = 5; time_tick = 1;\n")
    raw, (q, r), rows = _run(tmp_path, bnd, cfg, n)
    var = 2 * pi * (1 - pi) * (k - 1 + math.exp(-k)) / k ** 2
    for w in range(5):
        for name in ("A", "<nil>"):
            p, err = rows[w]["cells"][name]
            assert p == pytest.approx(pi if name == "A" else 1 - pi, rel=0.01)
            assert _var_hat(err, n) == pytest.approx(var, rel=0.04), (w, name)


def test_discrete_time_one_step_per_window(tmp_path):
    n = 20000
    bnd, cfg = _model(tmp_path, "s", [("A", "!A", "@logic ? 1 : 0", "@logic ? 0 : 1.5"),
                                      ("B", "A", "@logic ? 2 : 0", "@logic ? 0 : 0.5")],
                      "A.istate = 0.5; B.istate = 0.5; max_time = 6; time_tick = 1; discrete_time = 1;\n")
    raw, (q, r), rows = _run(tmp_path, bnd, cfg, n)
    assert raw["W"] and set(q) == set(raw["W"])
    for key, s1 in raw["W"].items():
        assert q[key] == s1
