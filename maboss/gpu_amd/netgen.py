"""Seeded synthetic Boolean networks for the benchmark workloads.

Writes MaBoSS ``.bnd`` / ``.cfg`` text.  Every random choice comes from a
splitmix64 stream seeded with the workload's seed (Steele, Lea, Flood,
OOPSLA 2014), and normals from the Box-Muller transform of two of its
uniforms, so the networks are the same on every machine and Python version.
Rates are drawn lognormal(mu, sigma) = exp(mu + sigma*Z), rounded to float32
and printed with 9 significant digits (exact float32 round trip).

Workloads (the sizes of the benchmark configurations):
  * ``plain_small``  - 10 nodes, seed 0, 2-3 parents, AND/OR/NOT formulas, all
    nodes observed, lognormal(0, 0.5), 1e4 trajectories, max_time 5, tick 0.2;
  * ``plain_100``    - 100 nodes, seed 1, in-degree 1-4, 50 internal nodes,
    lognormal(0, 1), 1e6 trajectories, max_time 100, tick 1 (+ discrete run);
  * ``scale_1000``   - 1000 nodes, seed 2, scale-free out-degree (exponent
    2.5), in-degree <= 4, threshold or AND/OR formulas, 20 observed nodes,
    lognormal(0, 1), istate 0.3, 1e6 trajectories, max_time 50, tick 0.5;
  * ``batched_50``   - 50-node base network, seed 3 (variants are built by
    the batched form);
  * ``grad_30``      - 30 nodes, seed 4, 60 rate parameters lognormal(0, 0.5),
    5 observed target nodes, 1e6 trajectories, max_time 20, tick 0.5.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform on (0, 1) with 53 bits."""
        return ((self.next_u64() >> 11) + 0.5) / float(1 << 53)

    def randint(self, lo: int, hi: int) -> int:
        """Integer in [lo, hi] (inclusive)."""
        return lo + (self.next_u64() % (hi - lo + 1))

    def normal(self) -> float:
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def lognormal(self, mu: float, sigma: float) -> float:
        return f32(math.exp(mu + sigma * self.normal()))

    def sample(self, population: list, k: int) -> list:
        pool = list(population)
        out = []
        for _ in range(k):
            j = self.randint(0, len(pool) - 1)
            out.append(pool.pop(j))
        return out


def f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", x))[0]


def fmt(x: float) -> str:
    return f"{x:.9g}"


@dataclass
class Network:
    names: list[str]
    logic: list[str]
    up: list[float]
    down: list[float]
    internal: list[bool]
    istate: list[float]
    constants: dict = field(default_factory=dict)

    def bnd(self) -> str:
        out = []
        for i, n in enumerate(self.names):
            out.append(f"Node {n} {{\n\tlogic = {self.logic[i]};\n"
                       f"\trate_up = @logic ? $u_{n} : 0;\n\trate_down = @logic ? 0 : $d_{n};\n}}\n")
        return "\n".join(out)

    def cfg(self) -> str:
        out = []
        for i, n in enumerate(self.names):
            out.append(f"$u_{n} = {fmt(self.up[i])};")
            out.append(f"$d_{n} = {fmt(self.down[i])};")
        for i, n in enumerate(self.names):
            out.append(f"{n}.istate = {fmt(self.istate[i])};")
            out.append(f"{n}.is_internal = {1 if self.internal[i] else 0};")
        for k, v in self.constants.items():
            out.append(f"{k} = {v};")
        return "\n".join(out) + "\n"

    def write(self, prefix: str) -> tuple[str, str]:
        d = os.path.dirname(prefix)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(prefix + ".bnd", "w") as f:
            f.write(self.bnd())
        with open(prefix + ".cfg", "w") as f:
            f.write(self.cfg())
        return prefix + ".bnd", prefix + ".cfg"


def _lit(rng: SplitMix64, name: str) -> str:
    return ("!" + name) if rng.next_u64() & 1 else name


def _and_or(rng: SplitMix64, parents: list[str]) -> str:
    expr = _lit(rng, parents[0])
    for p in parents[1:]:
        op = " & " if rng.next_u64() & 1 else " | "
        expr = f"({expr}{op}{_lit(rng, p)})"
    return expr


def _threshold(rng: SplitMix64, parents: list[str]) -> str:
    lits = " + ".join(_lit(rng, p) for p in parents)
    t = rng.randint(1, len(parents))
    return f"({lits}) >= {t}"


def random_network(seed: int, n: int, indeg: tuple[int, int], n_internal: int, mu: float, sigma: float,
                   istate: float = 0.5, prefix: str = "N") -> Network:
    """Random network: uniform in-degree in indeg, AND/OR/NOT formulas."""
    rng = SplitMix64(seed)
    names = [f"{prefix}{i}" for i in range(n)]
    logic, up, down = [], [], []
    for i in range(n):
        k = rng.randint(indeg[0], min(indeg[1], n - 1))
        parents = rng.sample([j for j in range(n) if j != i], k)
        logic.append(_and_or(rng, [names[j] for j in parents]))
    for i in range(n):
        up.append(rng.lognormal(mu, sigma))
        down.append(rng.lognormal(mu, sigma))
    internal_set = set(rng.sample(list(range(n)), n_internal))
    internal = [i in internal_set for i in range(n)]
    return Network(names, logic, up, down, internal, [istate] * n)


def scale_free_network(seed: int, n: int, exponent: float, max_in: int, n_observed: int, mu: float,
                       sigma: float, istate: float) -> Network:
    """Scale-free out-degree (Chung-Lu style weights w_j ~ j^(-1/(exponent-1))),
    in-degree uniform in 1..max_in; formulas are threshold or AND/OR."""
    rng = SplitMix64(seed)
    names = [f"S{i}" for i in range(n)]
    alpha = 1.0 / (exponent - 1.0)
    weights = [(j + 1) ** (-alpha) for j in range(n)]
    order = rng.sample(list(range(n)), n)  # random hub placement
    w = [0.0] * n
    for rank, j in enumerate(order):
        w[j] = weights[rank]
    cum = []
    acc = 0.0
    for x in w:
        acc += x
        cum.append(acc)

    def pick_parent(exclude: set) -> int:
        while True:
            r = rng.uniform() * acc
            lo, hi = 0, n - 1
            while lo < hi:
                mid = (lo + hi) // 2
                if cum[mid] < r:
                    lo = mid + 1
                else:
                    hi = mid
            if lo not in exclude:
                return lo

    logic, up, down = [], [], []
    for i in range(n):
        k = rng.randint(1, max_in)
        parents: list[int] = []
        excl = {i}
        for _ in range(k):
            p = pick_parent(excl)
            excl.add(p)
            parents.append(p)
        pn = [names[j] for j in parents]
        logic.append(_threshold(rng, pn) if (rng.next_u64() & 1) and k > 1 else _and_or(rng, pn))
    for i in range(n):
        up.append(rng.lognormal(mu, sigma))
        down.append(rng.lognormal(mu, sigma))
    observed = set(rng.sample(list(range(n)), n_observed))
    internal = [i not in observed for i in range(n)]
    return Network(names, logic, up, down, internal, [istate] * n)


WORKLOADS = ("plain_small", "plain_100", "plain_100_discrete", "scale_1000", "batched_50", "grad_30")


def workload(name: str) -> Network:
    """The benchmark network of a workload name (see module docstring)."""
    if name == "plain_small":
        net = random_network(0, 10, (2, 3), 0, 0.0, 0.5)
        net.constants = dict(sample_count=10000, max_time=5, time_tick=0.2, discrete_time=0, seed_pseudorandom=0)
    elif name in ("plain_100", "plain_100_discrete"):
        net = random_network(1, 100, (1, 4), 50, 0.0, 1.0)
        net.constants = dict(sample_count=1000000, max_time=100, time_tick=1,
                             discrete_time=1 if name.endswith("discrete") else 0, seed_pseudorandom=1)
    elif name == "scale_1000":
        net = scale_free_network(2, 1000, 2.5, 4, 20, 0.0, 1.0, 0.3)
        net.constants = dict(sample_count=1000000, max_time=50, time_tick=0.5, discrete_time=0, seed_pseudorandom=2)
    elif name == "batched_50":
        net = random_network(3, 50, (1, 4), 40, 0.0, 1.0)
        net.constants = dict(sample_count=10000, max_time=100, time_tick=1, discrete_time=0, seed_pseudorandom=3)
    elif name == "grad_30":
        net = random_network(4, 30, (1, 3), 25, 0.0, 0.5)
        net.constants = dict(sample_count=1000000, max_time=20, time_tick=0.5, discrete_time=0, seed_pseudorandom=4)
    else:
        raise KeyError(name)
    return net


def write_workload(name: str, directory: str) -> tuple[str, str]:
    return workload(name).write(os.path.join(directory, name))


def batched_variants(param_values, n_nodes: int, count: int = 256, seed: int = 3) -> list[dict]:
    """The variants of the batched workload over the ``batched_50`` network:
    every rate multiplied by lognormal(0, 0.5), 0-3 mutant nodes (each forced
    on or off with probability 1/2), a trajectory count log-uniform in
    [1e3, 1e5] and max_time uniform in [10, 200].  Drawn from a splitmix64
    stream of its own (the network's seed with the top bit set), so they are
    the same on every machine; returns dicts for ``Program.simulate_batched``."""
    rng = SplitMix64(seed | (1 << 63))
    out = []
    for _ in range(count):
        rates = [f32(v * f32(rng.lognormal(0.0, 0.5))) for v in param_values]
        mut = [0] * n_nodes
        for i in rng.sample(list(range(n_nodes)), rng.randint(0, 3)):
            mut[i] = 1 if rng.uniform() < 0.5 else -1
        n = int(round(math.exp(math.log(1e3) + rng.uniform() * (math.log(1e5) - math.log(1e3)))))
        max_time = f32(10.0 + 190.0 * rng.uniform())
        out.append({"n_trajectories": n, "rates": rates, "mutations": mut, "max_time": max_time})
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("name", choices=WORKLOADS)
    ap.add_argument("out_dir")
    a = ap.parse_args()
    print(*write_workload(a.name, a.out_dir))
