"""Benchmark configurations that the oracle finishes in seconds, run through
the C ABI on the GPU and compared with the oracle step by step (FP64, 1e-12)."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.test_gpu_parity import TOL, rel, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def test_config1_1d_gaussian_kick_200_steps(torch):
    """1D SLDA, 64 points, 128 states, Gaussian V_ext of 0.1 eF switched on
    at t = 0, 200 AB5AM5 steps with a measurement every 50: densities (rho,
    nu, tau, j), energies and particle number against the oracle."""
    say("case: config 1 (1D, 64 points, 128 states, 200 steps) vs oracle")
    from wslda_amd import problems, td
    s = problems.config1()
    p = s["params"]
    cfg = po.Cfg()
    C.memmove(C.addressof(cfg), C.addressof(p), C.sizeof(cfg))
    P = po.Problem(cfg, s["wf"], s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    n = s["wf"].shape[0] // 2
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    R = td.TDRun(ctx, s["wf"], s["pots"])
    P.initial_measure()
    R.initial_measure()
    for it in range(200):
        grad = (it % 50 == 49)
        P.step(it, grad_last=grad)
        R.step(it, grad_last=grad)
        if grad:
            d = R.dens.cpu().numpy()
            assert rel(d, P.dens) < TOL, it
            e, er = R.energy(it + 1), P.energy(it + 1)
            for k in (po.EKIN, po.EPOT, po.EPAIR, po.EPOTEXT):
                assert abs(e[k] - er[k]) <= TOL * max(abs(er[k]), abs(er[:7]).max()), (it, k)
            assert abs(e[7] + e[8] - er[7] - er[8]) <= TOL * abs(er[7] + er[8])
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    ctx.close()
