"""Reader for the reference's stdout report (.res captures, and the engine's
and oracle's --print output): final points, fixed points, per-window
entropy, transition entropy and probabilities.  States are keyed by the
frozenset of their node names; the captures join names with "-", the current
format with " -- "."""
from __future__ import annotations

import math
import re


def _key(name: str):
    name = name.strip()
    if name == "<nil>":
        return frozenset()
    parts = name.split(" -- ") if " -- " in name else name.split("-")
    return frozenset(p.strip() for p in parts)


def parse_report(path: str) -> dict:
    out = {"final": {}, "fixed": {}, "windows": []}
    section = None
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if not line:
                continue
            if line == "final points:":
                section = out["final"]
                continue
            if line == "fixed points:":
                section = out["fixed"]
                continue
            m = re.match(r"window \(([^,]+), ([^\]]+)\]", line)
            if m:
                w = {"a": float(m.group(1)), "b": float(m.group(2)), "probs": {}}
                out["windows"].append(w)
                section = w["probs"]
                continue
            if line.startswith("entropy: "):
                out["windows"][-1]["H"] = float(line.split(": ")[1])
                continue
            if line.startswith("transition entropy: "):
                out["windows"][-1]["TH"] = float(line.split(": ")[1])
                continue
            p, name = line.split(" ", 1)
            section[_key(name)] = section.get(_key(name), 0.0) + float(p)
    return out


def zscores(ref: dict, new: dict, n_ref: int, n_new: int):
    """z-score per state of two probability tables (binomial, pooled p)."""
    zs = []
    for k in set(ref) | set(new):
        a, b = ref.get(k, 0.0), new.get(k, 0.0)
        p = (a * n_ref + b * n_new) / (n_ref + n_new)
        if p <= 0:
            continue
        sd = math.sqrt(max(p * (1 - p), 0.0) * (1.0 / n_ref + 1.0 / n_new))
        sd = max(sd, 1.0 / (n_ref + n_new))
        zs.append((abs(a - b) / sd, k, a, b))
    return zs
