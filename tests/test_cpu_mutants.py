"""Mutant nodes in the .cfg: `X.mutant = ON;` / `X.mutant = OFF;` (also 1 / 0).

No counterpart in the reference (README.md:33 lists mutants as a TODO); the
semantics are those of the batched form's mutant masks: a node forced on
never switches off and starts on (its rate_down becomes 0, its istate 1), a
node forced off the reverse.  A mutant .cfg must therefore give exactly what
the same model edited by hand gives: the oracle's exact integer dump, and the
engine's generated code, are compared between the two."""
import pytest

import maboss.gpu_amd as mg
from conftest import read_raw, run_oracle
from maboss.gpu_amd._capi import MbgError

NODES = [
    ("A", "!C", 1.5, 0.75),
    ("B", "A & !D", 2.0, 0.5),
    ("C", "B | D", 0.625, 1.25),
    ("D", "(A ^ C) | !B", 1.0, 3.0),
    ("E", "C & D", 0.875, 0.25),
]


def _bnd(edit=None):
    lines = []
    for name, logic, u, d in NODES:
        up = f"@logic ? {u} : 0"
        dn = f"@logic ? 0 : {d}"
        if edit and edit[0] == name:
            if edit[1] > 0:
                dn = "0"
            else:
                up = "0"
        lines.append(f"Node {name} {{ logic = {logic}; rate_up = {up}; rate_down = {dn}; }}")
    return "\n".join(lines) + "\n"


CFG = "sample_count = 3000; max_time = 8; time_tick = 0.5; E.is_internal = 1; A.istate = 0.3; C.istate = 1;\n"


def _write(tmp_path, tag, bnd, cfg):
    b, c = tmp_path / f"{tag}.bnd", tmp_path / f"{tag}.cfg"
    b.write_text(bnd)
    c.write_text(cfg)
    return str(b), str(c)


@pytest.mark.parametrize("node,value,direction", [("B", "ON", 1), ("C", "OFF", -1), ("A", "on", 1), ("D", "0", -1),
                                                  ("E", "1", 1)])
def test_oracle_mutant_equals_edited_model(tmp_path, node, value, direction):
    mb, mc = _write(tmp_path, "m", _bnd(), CFG + f"{node}.mutant = {value};\n")
    istate = "1" if direction > 0 else "0"
    eb, ec = _write(tmp_path, "e", _bnd((node, direction)), CFG + f"{node}.istate = {istate};\n")
    run_oracle(mb, mc, "--raw", tmp_path / "m.raw")
    run_oracle(eb, ec, "--raw", tmp_path / "e.raw")
    a, b = read_raw(tmp_path / "m.raw"), read_raw(tmp_path / "e.raw")
    assert a == b
    # and differs from the run without the mutant
    pb, pc = _write(tmp_path, "p", _bnd(), CFG)
    run_oracle(pb, pc, "--raw", tmp_path / "p.raw")
    assert a != read_raw(tmp_path / "p.raw")


def test_oracle_mutant_overrides_istate_in_either_order(tmp_path):
    eb, ec = _write(tmp_path, "e", _bnd(("C", -1)), CFG + "C.istate = 0;\n")
    run_oracle(eb, ec, "--raw", tmp_path / "e.raw")
    for k, text in enumerate(["C.mutant = OFF; C.istate = 1;\n", "C.istate = 1; C.mutant = OFF;\n"]):
        mb, mc = _write(tmp_path, f"m{k}", _bnd(), CFG + text)
        run_oracle(mb, mc, "--raw", tmp_path / f"m{k}.raw")
        assert read_raw(tmp_path / f"m{k}.raw") == read_raw(tmp_path / "e.raw")


@pytest.mark.parametrize("node,value,direction", [("B", "ON", 1), ("C", "OFF", -1), ("D", "0", -1)])
def test_engine_mutant_generates_the_edited_models_code(tmp_path, node, value, direction):
    istate = "1" if direction > 0 else "0"
    m = mg.Model.load(*_write(tmp_path, "m", _bnd(), CFG + f"{node}.mutant = {value};\n"))
    e = mg.Model.load(*_write(tmp_path, "e", _bnd((node, direction)), CFG + f"{node}.istate = {istate};\n"))
    assert m.generate_code() == e.generate_code()
    plain = mg.Model.load(*_write(tmp_path, "p", _bnd(), CFG))
    assert plain.generate_code() != m.generate_code()


@pytest.mark.parametrize("text", ["B.mutant = 2;", "B.mutant = X;", "Z.mutant = ON;"])
def test_bad_mutant_statements_are_rejected(tmp_path, text):
    with pytest.raises(MbgError):
        mg.Model.load(*_write(tmp_path, "m", _bnd(), CFG + text + "\n"))
