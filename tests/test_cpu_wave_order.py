"""The oracle's wave-kernel summation order (--order wave) against its
sequential order, where the mathematics makes them agree exactly.

With every rate a multiple of 1/8 (at most 4) and at most 1500 nodes, every
partial sum is a multiple of 1/8 below 2^13, so it is exact in FP32 whatever
the order of the additions; u2 * total is one rounding of the same value in
both orders, and the remainder rr - P[lane - 1] the wave order forms is a
multiple of ulp(rr) no larger than rr, hence exact too.  In exact arithmetic
the first node whose prefix sum reaches rr is the same node whichever way the
sums are grouped, so the wave order (lane slices of ceil(N/64) nodes, the
64-lane scan, then the slice's running sums) must return the same total and
the same flip node as the sequential order (src/jit_kernels/simulation.cu:
9-21) on every state.  A slice-indexing or scan mistake (a node in the wrong
lane, a lane added twice or not at all, a wrong row broadcast) changes the
total or the selected node for some state and fails here.  Non-dyadic rates
then bound the order's rounding: totals within a few ulps of the sequential
ones, entropies within 1e-5 relative."""
import numpy as np
import pytest

from conftest import ORACLE, pipe


def _dyadic_model(tmp_path, n_nodes, seed, dyadic=True):
    rng = np.random.default_rng(seed)
    lines = []
    for i in range(n_nodes):
        k = int(rng.integers(1, 5))
        par = rng.choice([j for j in range(n_nodes) if j != i], size=k, replace=False)
        lits = [("!" if rng.random() < 0.5 else "") + f"X{p}" for p in par]
        expr = lits[0]
        for lit in lits[1:]:
            expr = f"({expr} {'&' if rng.random() < 0.5 else '|'} {lit})"
        if dyadic:
            # multiples of 1/8 in [0, 4], about one in eight zero
            u, d = (int(x) / 8 for x in rng.integers(0, 33, size=2))
            u, d = (0.0 if rng.random() < 0.12 else u), (0.0 if rng.random() < 0.12 else d)
        else:
            u, d = (float(np.float32(x)) for x in np.exp(rng.normal(0, 1, size=2)))
        lines.append(f"Node X{i} {{ logic = {expr}; rate_up = @logic ? {u!r} : 0; rate_down = @logic ? 0 : {d!r}; }}")
    bnd = tmp_path / f"d{n_nodes}_{seed}.bnd"
    cfg = tmp_path / f"d{n_nodes}_{seed}.cfg"
    bnd.write_text("\n".join(lines) + "\n")
    internal = "".join(f"X{i}.is_internal = 1;\n" for i in range(n_nodes) if i % 3 != 0)
    cfg.write_text("sample_count = 10; max_time = 1; time_tick = 1;\n" + internal)
    return str(bnd), str(cfg)


def _states(n_nodes, count, seed):
    rng = np.random.default_rng(seed)
    W = (n_nodes + 31) // 32
    st = rng.integers(0, 2**32, size=(count, W), dtype=np.uint64).astype(np.uint32)
    if n_nodes % 32:
        st[:, -1] &= np.uint32((1 << (n_nodes % 32)) - 1)
    # a few sparse states, where long runs of zero rates sit between the nonzero ones
    st[:50] &= rng.integers(0, 2**32, size=(50, W), dtype=np.uint64).astype(np.uint32)
    u2 = rng.uniform(0, 1, size=count).astype(np.float32)
    u2[:4] = [np.float32(1.0), np.float32(2**-24), np.float32(0.5), np.float32(1.0 - 2**-24)]
    return st, u2


def _eval(bnd, cfg, st, u2, order):
    text = "".join(" ".join(f"{w:x}" for w in row) + f" {u.view(np.uint32):08x}\n" for row, u in zip(st, u2))
    out = pipe(ORACLE, text, "--eval-states", "--order", order, bnd, cfg, timeout=600)
    rows = [line.split() for line in out]
    assert len(rows) == len(st)
    return rows


def _f(hexbits):
    return float(np.array([int(hexbits, 16)], dtype=np.uint32).view(np.float32)[0])


@pytest.mark.parametrize("n_nodes", [65, 130, 1000, 1500])
def test_dyadic_rates_wave_order_selects_like_sequential(tmp_path, n_nodes):
    bnd, cfg = _dyadic_model(tmp_path, n_nodes, seed=n_nodes)
    st, u2 = _states(n_nodes, 10000, seed=n_nodes + 1)
    seq = _eval(bnd, cfg, st, u2, "seq")
    wave = _eval(bnd, cfg, st, u2, "wave")
    nonzero = 0
    for a, b in zip(seq, wave):
        # total (bits), flip node and observed index are identical
        assert (a[0], a[2], a[3]) == (b[0], b[2], b[3]), (a, b)
        nonzero += _f(a[0]) > 0
    assert nonzero > 9000
    # the selections are spread over the nodes, not stuck at one end
    flips = np.array([int(r[2]) for r in seq])
    assert len(set(flips.tolist())) > min(n_nodes // 2, 500)


@pytest.mark.parametrize("n_nodes", [130, 1500])
def test_lognormal_rates_wave_order_rounding_bounded(tmp_path, n_nodes):
    bnd, cfg = _dyadic_model(tmp_path, n_nodes, seed=7 * n_nodes, dyadic=False)
    st, u2 = _states(n_nodes, 3000, seed=3)
    seq = _eval(bnd, cfg, st, u2, "seq")
    wave = _eval(bnd, cfg, st, u2, "wave")
    same_flip = 0
    for a, b in zip(seq, wave):
        ta, tb = _f(a[0]), _f(b[0])
        # N additions in either order: a few ulps of the total apart
        assert abs(ta - tb) <= 4 * n_nodes * np.spacing(np.float32(ta)), (ta, tb)
        ha, hb = _f(a[1]), _f(b[1])
        assert abs(ha - hb) <= 1e-5 * max(abs(ha), 1e-3), (ha, hb)
        assert a[3] == b[3]
        same_flip += a[2] == b[2]
    # the flip can differ only when u2 * total falls within rounding of a prefix sum
    assert same_flip >= 0.995 * len(seq)
