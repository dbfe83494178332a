"""mbg_backward_noalloc (the gradient entry point that allocates nothing)
against mbg_backward, through the C-ABI.

Both run the same kernel on the same trajectories; the cross-trajectory sums
are double-precision atomics whose order depends on scheduling, so the two
agree to rounding: |a - b| <= 1e-9 * max|b| per component.  Two calls in a
row, the second on a smaller model (fewer nodes, parameters, windows and
trajectories) with the model's own rates, so it also checks that nothing of
the first call's contents in the shared scratch leaks into the second."""
import ctypes

import numpy as np
import pytest

import maboss.gpu_amd as mg
from maboss.gpu_amd import grad as mgrad
from maboss.gpu_amd import netgen
from maboss.gpu_amd._capi import GradParams, lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-9  # relative to the largest component

SMALL_BND = """Node A {
	logic = !A | B;
	rate_up = @logic ? $ua : 0;
	rate_down = @logic ? 0 : $da;
}
Node B {
	logic = A;
	rate_up = @logic ? $ub : 0;
	rate_down = @logic ? 0 : $db;
}
"""
SMALL_CFG = """$ua = 1.5; $da = 0.7; $ub = 0.9; $db = 2.1;
sample_count = 5000; max_time = 6; time_tick = 0.5; discrete_time = 0; seed_pseudorandom = 3;
"""


def noalloc(prog, G, rates, n):
    out = torch.empty(prog.model.info.n_params, dtype=torch.float64, device="cuda")
    gp = GradParams()
    lib().mbg_grad_params_default(ctypes.byref(gp))
    gp.n_trajectories = n
    s = torch.cuda.current_stream().cuda_stream
    rc = lib().mbg_backward_noalloc(prog._h, ctypes.byref(gp),
                                    ctypes.c_void_p(rates.data_ptr()) if rates is not None else None,
                                    ctypes.c_void_p(G.data_ptr()), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s))
    return rc, out


def check_close(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    scale = max(float(np.abs(b).max()), 1e-300)
    assert np.all(np.abs(a - b) <= TOL * scale), (a, b)
    assert np.abs(b).max() > 0


def test_noalloc_backward_matches_backward(tmp_path):
    d = tmp_path / "g30"
    d.mkdir()
    bnd, cfg = netgen.write_workload("grad_30", str(d))
    big = mg.Model.load(bnd, cfg).compile()
    small = mg.Model.from_strings(SMALL_BND, SMALL_CFG).compile()

    # before mbg_grad_check the gradient kernel is not loaded: refused, nothing launched
    m = big.model.info
    G = torch.from_numpy(np.random.default_rng(1).normal(size=(m.windows, 1 << m.n_observed))).cuda()
    rates = torch.as_tensor(big.model.param_values, dtype=torch.float32).cuda()
    rc, _ = noalloc(big, G, rates, 100)
    assert rc == 1, lib().mbg_last_error()

    for prog in (big, small):
        assert lib().mbg_grad_check(prog._h) == 0, lib().mbg_last_error()

    # call 1: the 30-node model, 60 parameters, rates given on the device
    print("[gpu] noalloc call 1: grad_30, 20000 trajectories", flush=True)
    rc, a = noalloc(big, G, rates, 20000)
    assert rc == 0, lib().mbg_last_error()
    b = mgrad.backward(big, G, rates=rates, n_trajectories=20000)
    torch.cuda.synchronize()
    check_close(a, b)

    # call 2: a smaller input -- 2 nodes, 4 parameters, 5000 trajectories, the model's rates
    print("[gpu] noalloc call 2: 2-node model, 5000 trajectories", flush=True)
    ms = small.model.info
    G2 = torch.from_numpy(np.random.default_rng(2).normal(size=(ms.windows, 1 << ms.n_observed))).cuda()
    rc, a2 = noalloc(small, G2, None, 5000)
    assert rc == 0, lib().mbg_last_error()
    b2 = mgrad.backward(small, G2, rates=None, n_trajectories=5000)
    torch.cuda.synchronize()
    check_close(a2, b2)
