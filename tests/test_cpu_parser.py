"""Model front-end: the four reference models, precedence, errors.

Pins the parser against facts fixed by the model files themselves (node and
observed-node counts and constants, read off the .bnd/.cfg themselves)
and the expression semantics against Python's own evaluation of the same
formulas with C precedence."""
import itertools
import re

import numpy as np
import pytest

import maboss.gpu_amd as mg
from conftest import model_paths

EXPECT = {
    # name: (nodes, observed names, discrete, max_time, tick, sample_count, seed, windows)
    "cellcycle": (10, ["CycD", "CycE", "CycA", "CycB"], 1, 5.0, 0.2, 1000000, 100, 25),
    "sizek": (87, ["CyclinE", "CyclinA", "CyclinB", "Casp3"], 0, 100.0, 5.0, 100000, 0, 20),
    "metastasis": (32, ["Metastasis", "Migration", "Invasion", "Apoptosis", "CellCycleArrest"], 0, 100.0, 0.5,
                   1000, 0, 200),
    "Montagud2022_Prostate_Cancer": (133, ["Apoptosis", "Metastasis", "Proliferation"], 0, 50.0, 0.1, 1000000,
                                     100, 500),
}


@pytest.mark.parametrize("name", list(EXPECT))
def test_reference_models(name):
    m = mg.Model.load(*model_paths(name))
    n, obs, disc, mt, tick, sc, seed, win = EXPECT[name]
    i = m.info
    assert i.n_nodes == n and i.n_words == (n + 31) // 32
    names = m.node_names
    assert [names[j] for j in m.observed] == obs
    assert i.discrete_time == disc
    assert i.max_time == pytest.approx(mt) and i.time_tick == np.float32(tick)
    assert i.sample_count == sc and i.seed == seed and i.windows == win
    assert i.n_params == (2 if name == "cellcycle" else 2 * n)


def _cellcycle_logic():
    txt = open(model_paths("cellcycle")[0]).read()
    nodes = re.findall(r"node\s+(\w+)\s*\{\s*logic\s*=\s*([^;]+);", txt)
    return nodes


def test_cellcycle_rates_all_states():
    """Every node's rate in all 1024 states, against a Python evaluation of
    the logic text and rate = 0 if logic == state else (on ? down : up)."""
    m = mg.Model.load(*model_paths("cellcycle"))
    logic = _cellcycle_logic()
    names = [n for n, _ in logic]
    assert names == m.node_names
    slow = {"CycE", "CycA", "CycB", "E2F", "Cdc20", "UbcH10"}  # rate_up = $slow
    pyexpr = [re.sub(r"!", " not ", re.sub(r"&", " and ", re.sub(r"\|", " or ", f))) for _, f in logic]
    for s in range(1 << 10):
        env = {n: bool((s >> i) & 1) for i, n in enumerate(names)}
        for i, n in enumerate(names):
            lg = bool(eval(pyexpr[i], {}, env))
            on = env[n]
            expect = 0.0 if lg == on else (10.0 if on else (1.0 if n in slow else 10.0))
            assert m.rate(i, [s]) == expect, (s, n)


CASES = [
    # (expression over nodes A, B, C as rate_up, python equivalent)
    ("1 + 2 * 3", lambda a, b, c: 7.0),
    ("-2 * 3 + 10", lambda a, b, c: 4.0),
    ("(A | B & !A) ? 2 : 3", lambda a, b, c: 2.0 if (a or (b and not a)) else 3.0),
    ("A ? 1 : B ? 2 : 3", lambda a, b, c: 1.0 if a else (2.0 if b else 3.0)),
    ("A ^ B ? 5 : 6", lambda a, b, c: 5.0 if (a != b) else 6.0),
    ("A XOR B OR C ? 1 : 0", lambda a, b, c: 1.0 if (a != (b or c)) else 0.0),
    ("(A + B + C) >= 2 ? 4 : 0", lambda a, b, c: 4.0 if (a + b + c) >= 2 else 0.0),
    ("A + B * C", lambda a, b, c: float(a + b * c)),
    ("1 < 2 == 1 ? 8 : 9", lambda a, b, c: 8.0),
    ("NOT A AND B ? 1.5 : .25", lambda a, b, c: 1.5 if ((not a) and b) else 0.25),
    ("A && !B || C ? 1 : 0", lambda a, b, c: 1.0 if ((a and not b) or c) else 0.0),
    ("$k / 4 + 1e-1", lambda a, b, c: np.float32(np.float32(3.0) / np.float32(4.0) + np.float32(0.1))),
    ("A == B ? 2 : 0", lambda a, b, c: 2.0 if a == b else 0.0),
    ("(A != (B | C)) * 3", lambda a, b, c: 3.0 * (a != (b or c))),
]


@pytest.mark.parametrize("expr,fn", CASES)
def test_precedence_and_semantics(expr, fn):
    bnd = ("node A { rate_up = %s; rate_down = 0; }\n"
           "node B { rate_up = 0; rate_down = 0; }\nnode C { rate_up = 0; rate_down = 0; }\n") % expr
    m = mg.Model.from_strings(bnd, "$k = 3;")
    for a, b, c in itertools.product([0, 1], repeat=3):
        if a:
            continue  # rate_up applies when A is off
        s = a | (b << 1) | (c << 2)
        assert m.rate(0, [s]) == np.float32(fn(a, b, c)), (expr, a, b, c)


def test_bracket_istate_and_exact_seed():
    m = mg.Model.from_strings("node A { rate_up = 1; rate_down = 1; }\nnode B { rate_up = 1; rate_down = 1; }",
                              "[A].istate = 3[1], 1[0];\n[B].istate = 3 [0] , 1 [1];\n"
                              "seed_pseudorandom = 1099511627777;\nsample_count = 123456789;")
    assert m.info.seed == 1099511627777
    assert m.info.sample_count == 123456789


def test_comments_and_case():
    m = mg.Model.from_strings("/* block */ NODE A { rate_up = TRUE ? 2 : 3; // c\n rate_down = false; }", "")
    assert m.rate(0, [0]) == 2.0 and m.rate(0, [1]) == 0.0


ERRORS = [
    ("node A { rate_up = 1 rate_down = 1; }", "", 3),            # missing ;
    ("node A { rate_up = 1; }", "", 3),                          # missing rate_down
    ("node A { rate_up = 1; rate_down = 1; } node A { rate_up = 1; rate_down = 1; }", "", 3),
    ("nod A { rate_up = 1; rate_down = 1; }", "", 3),            # keyword
    ("node A { rate_up = B; rate_down = 1; }", "", 4),           # unknown node
    ("node A { rate_up = $x; rate_down = 1; }", "", 4),          # undefined variable
    ("node A { rate_up = @rate_up; rate_down = 1; }", "", 4),    # cyclic alias
    ("node A { rate_up = @nope; rate_down = 1; }", "", 4),       # missing attribute
    ("node A { rate_up = 1; rate_down = 1; }", "B.istate = 1;", 4),
    ("node A { rate_up = 1; rate_down = 1; }", "x = A;", 3),     # node in a constant
    ("node A { rate_up = 1 # 2; rate_down = 1; }", "", 3),       # invalid character
    ("node A { rate_up = 1; rate_down = 1; }", "[A].istate = 1[1], 0[1];", 3),
    ("node A { rate_up = 1; rate_down = 1; }", "time_tick = 0;", 4),
]


@pytest.mark.parametrize("bnd,cfg,code", ERRORS)
def test_errors(bnd, cfg, code):
    with pytest.raises(mg.MbgError) as e:
        mg.Model.from_strings(bnd, cfg)
    assert e.value.code == code, e.value.message
    assert e.value.message


def test_io_errors(tmp_path):
    with pytest.raises(mg.MbgError) as e:
        mg.Model.load(str(tmp_path / "none.bnd"), str(tmp_path / "none.cfg"))
    assert e.value.code == 2
    p = tmp_path / "u16.bnd"
    p.write_bytes("node A { rate_up = 1; rate_down = 1; }".encode("utf-16"))
    with pytest.raises(mg.MbgError) as e:
        mg.Model.load(str(p), str(p))
    assert e.value.code == 2


def test_overrides():
    m = mg.Model.load(*model_paths("cellcycle"))
    m.set_constant("sample_count", 777)
    m.set_variable("fast", 4.5)
    assert m.info.sample_count == 777
    assert dict(zip(m.param_names, m.param_values))["$fast"] == np.float32(4.5)
    with pytest.raises(mg.MbgError):
        m.set_constant("max_time", -1)
