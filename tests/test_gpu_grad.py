"""Gradient form on the GPU.

* mbg_backward against the oracle's double-precision estimator on the same
  trajectories (the oracle evaluates sum_w Gw * S(e_w) directly, the kernel
  its suffix form, so agreement also checks that algebra); both accumulate
  in double, so every component agrees to 1e-6 relative;
* against closed forms (a two-state toggle and pure decay), statistically;
* the torch.autograd.Function end to end, against mbg_backward and against
  central finite differences with common random numbers (step 1e-3 of each
  rate).
"""
import math
import re

import numpy as np
import pytest

import maboss.gpu_amd as mg
from maboss.gpu_amd import grad as mgrad
from maboss.gpu_amd import netgen
from maboss.gpu_amd._capi import lib
from conftest import model_paths, run_oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOGGLE_BND = """Node A {
	logic = !A;
	rate_up = @logic ? $u : 0;
	rate_down = @logic ? 0 : $d;
}
"""
DECAY_BND = """Node A {
	logic = !A;
	rate_up = @logic ? 0 : 0;
	rate_down = @logic ? 0 : $d;
}
"""


def cfg_text(rates, istate, max_time, tick, n=1000, seed=7):
    lines = [f"${k} = {v!r};" for k, v in rates.items()]
    lines += [f"A.istate = {istate};", f"sample_count = {n};", f"max_time = {max_time};", f"time_tick = {tick};",
              "discrete_time = 0;", f"seed_pseudorandom = {seed};"]
    return "\n".join(lines) + "\n"


@pytest.fixture(scope="module")
def g30(tmp_path_factory):
    d = tmp_path_factory.mktemp("g30")
    bnd, cfg = netgen.write_workload("grad_30", str(d))
    return bnd, cfg, mg.Model.load(bnd, cfg).compile()


def oracle_grad(tmp_path, bnd, cfg, G, n, first_id=0, tag="o"):
    gpath = tmp_path / f"{tag}.g.bin"
    np.ascontiguousarray(G, dtype=np.float64).tofile(gpath)
    out = tmp_path / f"{tag}.grad.txt"
    run_oracle(bnd, cfg, "--n", n, "--first-id", first_id, "--grad-in", gpath, "--grad-out", out)
    d = {}
    for line in open(out):
        k, v = line.split()
        d[k] = float(v)
    return d


def gpu_grad(prog, G, n, first_id=0, rates=None, info=False):
    g = torch.from_numpy(np.ascontiguousarray(G, dtype=np.float64)).cuda()
    r = None if rates is None else torch.as_tensor(rates, dtype=torch.float32).cuda()
    out = mgrad.backward(prog, g, rates=r, n_trajectories=n, first_id=first_id, info=info)
    if info:
        return out[0].cpu().numpy(), out[1]
    return out.cpu().numpy()


def test_backward_matches_oracle(tmp_path, g30):
    bnd, cfg, prog = g30
    m = prog.model.info
    rng = np.random.default_rng(3)
    G = rng.normal(size=(m.windows, 1 << m.n_observed))
    n = 20000
    print(f"[gpu] grad_30 backward n={n}", flush=True)
    got, gi = gpu_grad(prog, G, n, info=True)
    want = oracle_grad(tmp_path, bnd, cfg, G, n)
    names = prog.model.param_names
    assert gi.transitions == int(want["transitions"])
    o = np.array([want[k] for k in names])
    scale = np.abs(o).max()
    assert scale > 0
    err = np.abs(got - o)
    # Both accumulate every score in double on the same trajectories; they
    # differ only by rounding (suffix form vs direct form, atomic order).
    # Every component within 1e-6 relative, with an absolute floor of
    # 1e-10 of the largest component for components that average to ~0.
    assert np.all(err <= 1e-6 * np.abs(o) + 1e-10 * scale), (err / np.maximum(np.abs(o), 1e-300)).max()


SHARED_BND = """Node A { logic = !C; rate_up = @logic ? $k : 0; rate_down = @logic ? 0 : $d; }
Node B { logic = A & !D; rate_up = @logic ? $k : 0; rate_down = @logic ? 0 : $d; }
Node C { logic = B | E; rate_up = @logic ? $k : 0; rate_down = @logic ? 0 : $k; }
Node D { logic = !A; rate_up = @logic ? $u : 0; rate_down = @logic ? 0 : $d; }
Node E { logic = C & D; rate_up = @logic ? 1.5 : 0; rate_down = @logic ? 0 : $u; }
Node F { logic = E | !B; rate_up = @logic ? $u : 0; rate_down = @logic ? 0 : 0.75; }
"""
SHARED_CFG = """$k = 1.25; $d = 0.5; $u = 2.0;
E.is_internal = 1; F.is_internal = 1; A.istate = 0.3; C.istate = 0.6;
sample_count = 20000; max_time = 12; time_tick = 0.5; discrete_time = 0; seed_pseudorandom = 11;
"""


def test_backward_shared_parameters_match_oracle(tmp_path):
    """Parameters each used by several node directions ($k by three up rates
    and one down rate, $d by three down rates, $u by up and down rates of
    different nodes), with parameter-free rates beside them: every component
    equals the oracle's to the same bar as grad_30."""
    bnd, cfg = tmp_path / "s.bnd", tmp_path / "s.cfg"
    bnd.write_text(SHARED_BND)
    cfg.write_text(SHARED_CFG)
    prog = mg.Model.load(str(bnd), str(cfg)).compile()
    m = prog.model.info
    G = np.random.default_rng(5).normal(size=(m.windows, 1 << m.n_observed))
    n = 20000
    print(f"[gpu] shared-parameter gradient n={n}", flush=True)
    got = gpu_grad(prog, G, n)
    want = oracle_grad(tmp_path, str(bnd), str(cfg), G, n, tag="sh")
    o = np.array([want[k] for k in prog.model.param_names])
    scale = np.abs(o).max()
    assert scale > 0 and len(o) == 3
    err = np.abs(got - o)
    assert np.all(err <= 1e-6 * np.abs(o) + 1e-10 * scale), (err / np.maximum(np.abs(o), 1e-300)).max()


def test_backward_two_words_ninety_parameters(tmp_path):
    """A 45-node network (2 state words) with 90 rate parameters: the kernel
    compiled for 1 wave per SIMD (above 64 parameters) and the enabled-word
    events of the second word; every component equals the oracle's."""
    net = netgen.random_network(6, 45, (1, 3), 40, 0.0, 0.5)
    net.constants = dict(sample_count=20000, max_time=10, time_tick=0.5, discrete_time=0, seed_pseudorandom=6)
    bnd, cfg = net.write(str(tmp_path / "g45"))
    prog = mg.Model.load(bnd, cfg).compile()
    m = prog.model.info
    assert m.n_words == 2 and m.n_params == 90
    G = np.random.default_rng(9).normal(size=(m.windows, 1 << m.n_observed))
    n = 20000
    print(f"[gpu] 45-node gradient n={n}", flush=True)
    got = gpu_grad(prog, G, n)
    want = oracle_grad(tmp_path, bnd, cfg, G, n, tag="g45")
    o = np.array([want[k] for k in prog.model.param_names])
    scale = np.abs(o).max()
    assert scale > 0
    err = np.abs(got - o)
    assert np.all(err <= 1e-6 * np.abs(o) + 1e-10 * scale), (err / np.maximum(np.abs(o), 1e-300)).max()


def test_backward_runtime_rates_and_shards(tmp_path, g30):
    """Rates passed at run time equal rates edited into the model; shards add up."""
    bnd, cfg, prog = g30
    m = prog.model.info
    names = prog.model.param_names
    base = prog.model.param_values
    rates = (base * np.float32(1.3)).astype(np.float32)
    text = open(cfg).read()
    for k, nm in enumerate(names):
        text, c = re.subn(rf"^\{nm}\s*=.*;$", f"{nm} = {float(rates[k]):.9g};", text, flags=re.M)
        assert c == 1
    cfg2 = tmp_path / "scaled.cfg"
    cfg2.write_text(text)
    G = np.random.default_rng(4).normal(size=(m.windows, 1 << m.n_observed))
    n = 12000
    got = gpu_grad(prog, G, n, rates=rates)
    want = oracle_grad(tmp_path, bnd, str(cfg2), G, n, tag="s")
    o = np.array([want[k] for k in names])
    assert np.all(np.abs(got - o) <= 1e-6 * np.abs(o) + 1e-10 * np.abs(o).max())
    # ids [0, 5000) + [5000, 12000), each normalised by its own count
    a = gpu_grad(prog, G, 5000, rates=rates)
    b = gpu_grad(prog, G, 7000, first_id=5000, rates=rates)
    comb = (5000 * a + 7000 * b) / 12000
    assert np.abs(comb - got).max() <= 1e-9 * np.abs(got).max()


def _batches(prog, G, n_batch, k_batches):
    vals = np.array([gpu_grad(prog, G, n_batch, first_id=i * n_batch) for i in range(k_batches)])
    return vals.mean(0), vals.std(0, ddof=1) / math.sqrt(k_batches)


def test_toggle_closed_form(tmp_path):
    """Two-state chain, rates u (off->on) and d (on->off): P(on, t) = pi + (p0 - pi) e^{-(u+d) t}."""
    u, d, p0, T, tick = 1.25, 0.5, 0.3, 4.0, 0.5
    prog = mg.Model.from_strings(TOGGLE_BND, cfg_text({"u": u, "d": d}, p0, T, tick)).compile()
    W = int(round(T / tick))
    G = np.zeros((W, 2))
    G[:, 1] = 1.0  # L = sum_w P_w(on)

    def L(u, d):
        lam = u + d
        pi = u / lam
        return sum(pi + (p0 - pi) * (math.exp(-lam * w * tick) - math.exp(-lam * (w + 1) * tick)) / (lam * tick)
                   for w in range(W))

    h = 1e-6
    exact = np.array([(L(u + h, d) - L(u - h, d)) / (2 * h), (L(u, d + h) - L(u, d - h)) / (2 * h)])
    mean, se = _batches(prog, G, 250000, 8)
    order = [prog.model.param_names.index(k) for k in ("$u", "$d")]
    z = (mean[order] - exact) / se[order]
    assert np.all(np.abs(z) < 4.5), (mean[order], exact, se[order])
    assert np.all(se[order] < 0.02 * np.abs(exact))


def test_decay_closed_form():
    """Pure decay from on: P(on, t) = e^{-d t}; the off state is a fixed point."""
    d, T, tick = 0.8, 3.0, 0.25
    prog = mg.Model.from_strings(DECAY_BND, cfg_text({"d": d}, 1, T, tick)).compile()
    W = int(round(T / tick))
    G = np.zeros((W, 2))
    G[:, 1] = np.linspace(1.0, 2.0, W)  # weighted sum of P_w(on)

    def L(d):
        return sum(G[w, 1] * (math.exp(-d * w * tick) - math.exp(-d * (w + 1) * tick)) / (d * tick) for w in range(W))

    exact = (L(d + 1e-6) - L(d - 1e-6)) / 2e-6
    mean, se = _batches(prog, G, 250000, 8)
    assert abs(mean[0] - exact) < 4.5 * se[0], (mean[0], exact, se[0])


def test_autograd_function_and_crn_finite_differences(g30):
    bnd, cfg, prog = g30
    m = prog.model.info
    base = torch.tensor(prog.model.param_values, dtype=torch.float64)
    gen = torch.Generator().manual_seed(9)
    theta2 = base * torch.exp(0.3 * torch.randn(base.shape, generator=gen, dtype=torch.float64))
    n = 200000
    with torch.no_grad():
        target = mgrad.observed_marginals(mgrad.window_probabilities(prog, theta2, n, seed=11), m.n_observed)

    def loss_of(theta, seed):
        P = mgrad.window_probabilities(prog, theta, n, seed=seed)
        marg = mgrad.observed_marginals(P, m.n_observed)
        return ((marg - target) ** 2).sum(), P, marg

    # autograd == mbg_backward with the chained dL/dP
    theta = base.clone().requires_grad_(True)
    loss, P, marg = loss_of(theta, 1)
    loss.backward()
    lr = theta.grad.clone()
    x = torch.arange(1 << m.n_observed, device=P.device)
    bits = ((x[:, None] >> torch.arange(m.n_observed, device=P.device)[None, :]) & 1).double()
    gP = (2 * (marg.detach() - target)) @ bits.T
    direct = mgrad.backward(prog, gP.contiguous(), rates=base.float().cuda(), n_trajectories=n, seed=1)
    assert torch.allclose(lr, direct.cpu(), rtol=1e-9, atol=1e-12 * float(lr.abs().max()))

    # LR vs central differences, common random numbers (same seed and ids), over 6 seeds
    k = torch.argsort(lr.abs(), descending=True)[:3]
    lrs, fds = [], []
    for seed in range(6):
        th = base.clone().requires_grad_(True)
        l, _, _ = loss_of(th, 100 + seed)
        l.backward()
        lrs.append(th.grad[k].numpy())
        row = []
        for j in k.tolist():
            h = 1e-3 * float(base[j])
            with torch.no_grad():
                tp = base.clone()
                tp[j] += h
                tm = base.clone()
                tm[j] -= h
                row.append((loss_of(tp, 100 + seed)[0].item() - loss_of(tm, 100 + seed)[0].item()) / (2 * h))
        fds.append(row)
    lrs, fds = np.array(lrs), np.array(fds)
    diff = lrs - fds
    se = diff.std(0, ddof=1) / math.sqrt(len(diff))
    assert np.all(np.abs(diff.mean(0)) < 4.5 * se + 0.02 * np.abs(lrs.mean(0))), (lrs.mean(0), fds.mean(0), se)


def _dense(res, windows, n_obs):
    Q = np.zeros((windows, 1 << n_obs))
    for w in range(windows):
        win = res.window(w)
        Q[w, win.states.astype(np.int64)] = win.probs
    return Q


def test_window_probs_device_equal_result_probs(g30):
    """The stream-ordered device probabilities (the autograd forward) equal the
    host result's probabilities bit for bit: model rates, device rates, and a
    discrete-time model."""
    bnd, cfg, prog = g30
    m = prog.model.info
    base = prog.model.param_values
    for rates in (None, (base * np.float32(1.2)).astype(np.float32)):
        print(f"[gpu] device probabilities grad_30 rates={'model' if rates is None else 'scaled'}", flush=True)
        r = None if rates is None else torch.as_tensor(rates).cuda()
        P = mgrad.window_probs_device(prog, r, n_trajectories=20000, seed=5).cpu().numpy()
        res = prog.simulate(n_trajectories=20000, seed=5, rates=rates)
        assert np.array_equal(P, _dense(res, m.windows, m.n_observed))
    cc = mg.Model.load(*model_paths("cellcycle")).compile()
    ci = cc.model.info
    assert ci.discrete_time == 1
    P = mgrad.window_probs_device(cc, None, n_trajectories=5000).cpu().numpy()
    assert np.array_equal(P, _dense(cc.simulate(n_trajectories=5000), ci.windows, ci.n_observed))


def test_gradient_form_refusals(g30):
    bnd, cfg, prog = g30
    m = prog.model.info
    with pytest.raises(ValueError):
        mgrad.backward(prog, torch.zeros(3, dtype=torch.float64).cuda())
    cc = mg.Model.load(*model_paths("cellcycle")).compile()
    assert lib().mbg_grad_check(cc._h) == 5  # discrete time: MBG_ERR_LIMIT
    assert lib().mbg_grad_check(prog._h) == 0
