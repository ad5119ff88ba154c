"""Taylor self-start (pca.c:1086-1392, pca_kernels.cu:1032-1085) on the GPU.

taylor_expansion_contribution is checked against its closed form on
plane-wave eigenstates (each term is (-i lam dt)^k / k! psi); tdw_selfstart
against the oracle's restatement (itself pinned to the exact propagator of a
linear problem in tests/test_oracle_selfstart.py) at 1e-12, and then through
ABM steps that continue from the self-started history."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss
from tests.test_gpu_parity import TOL, rel, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def test_taylor_terms_closed_form(torch):
    """taylor_expansion_contribution (pca_kernels.cu:1032-1085) on plane-wave
    eigenstates of the uniform solution, orders 1..5, H applied without the
    <H> subtraction (useqpe = 0):
    - each call gives contr = (-i dt / k) H contr_prev and upd += contr
      exactly (the recursion, to round-off);
    - the k-th term is (-i lam dt)^k / k! psi with lam the state's Rayleigh
      quotient (equal to the solver's E_n): to 1e-12 for k = 1, then within
      1e-12 (K / lam)^(k - 1) with K / lam <= 100 -- round-off of a
      finite-precision eigenstate lies along high-momentum directions, which
      H amplifies by the kinetic scale K over lam at each order."""
    from wslda_amd import problems, td
    say("case: taylor_expansion_contribution, plane-wave eigenstates, orders 1..5")
    s = problems.uniform_start(3, (8, 8, 8), n_states=40)
    p = s["params"]
    n = s["wf"].shape[0] // 2
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(s["fbeta"])
    L = td.lib()
    L.tdw_bind(ctx.h)
    try:
        ptr = lambda t: C.c_void_p(t.data_ptr())
        psi = torch.as_tensor(s["wf"]).to("cuda:0")
        pots = torch.as_tensor(s["pots"]).to("cuda:0")
        dens = torch.zeros(16 * ctx.N, dtype=torch.float64, device="cuda:0")
        zero = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        lap = torch.zeros_like(psi)
        hpsi = torch.zeros_like(psi)
        contr = psi.clone()
        upd = psi.clone()
        dt = 0.04
        En = torch.as_tensor(s["En"]).to("cuda:0")
        expect_upd = psi.clone()
        for k in range(1, 6):
            td.check(L.compute_laplace(2 * n, ptr(contr), ptr(lap), 256), "compute_laplace")
            td.check(L.apply_hamiltonian(0, n, ptr(contr), ptr(hpsi), None, None, None, ptr(lap), None, ptr(dens),
                                         ptr(pots), C.c_double(0.0), ptr(zero), C.c_double(1.0), 256), "apply_h")
            ctx.sync()
            upd_prev = upd.clone()
            td.check(L.taylor_expansion_contribution(k, C.c_double(dt), n, ptr(hpsi), ptr(upd), ptr(contr), 256),
                     "taylor")
            ctx.sync()
            # the recursion itself
            step = hpsi * (-1j * dt / k)
            assert rel(contr.cpu().numpy(), step.cpu().numpy()) < 1e-15, k
            assert rel(upd.cpu().numpy(), (upd_prev + contr).cpu().numpy()) < 1e-15, k
            if k == 1:
                # per state (u row a and v row n + a together)
                num = (psi[:n].conj() * hpsi[:n]).sum(dim=1).real + (psi[n:].conj() * hpsi[n:]).sum(dim=1).real
                den = (psi[:n].abs() ** 2).sum(dim=1) + (psi[n:].abs() ** 2).sum(dim=1)
                lam = num / den
                say("  max |lam - E_n| = %.3e" % (lam - En).abs().max().item())
                assert (lam - En).abs().max().item() < 1e-9
            fac = ((-1j * dt * lam) ** k / math.factorial(k))
            want = psi * torch.cat([fac, fac])[:, None]
            expect_upd = expect_upd + want
            e_contr = rel(contr.cpu().numpy(), want.cpu().numpy())
            e_upd = rel(upd.cpu().numpy(), expect_upd.cpu().numpy())
            say("  order %d: term %.2e, update %.2e" % (k, e_contr, e_upd))
            assert e_contr < 1e-12 * 100.0 ** (k - 1), k
            assert e_upd < 1e-12, k
    finally:
        L.tdw_bind(None)
    ctx.close()


def run_pair(torch, cfg, wf, fbeta, kt, cnt, pots, fused, steps=3):
    from wslda_amd import td
    P = po.Problem(cfg, wf, fbeta, kt=kt, cnt=cnt, pots=pots)
    n = P.n
    ctx = td.TDContext(tdgauss.to_params(cfg), n, 0)
    ctx.set_states(fbeta, kt, cnt)
    R = td.TDRun(ctx, wf, pots, fused=fused)
    P.initial_measure()
    R.initial_measure()
    P.selfstart()
    R.selfstart()
    k = {34: 3, 45: 4, 55: 5}[cfg.scheme]
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    assert rel(R.dens.cpu().numpy(), P.dens) < TOL
    assert rel(R.pots.cpu().numpy(), P.pots) < TOL
    hg, ho = R.history(), P.history()
    for j in range(k):
        assert rel(hg[j].cpu().numpy(), ho[j]) < 1e-11, j
    for it in range(steps):
        grad = it == steps - 1
        P.step(it, grad_last=grad)
        R.step(it, grad_last=grad)
    R.finalize()
    ctx.sync()
    assert rel(R.dens.cpu().numpy(), P.dens) < TOL
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    ctx.close()


@pytest.mark.parametrize("dim,scheme", [(1, 45), (1, 34), (2, 55), (3, 55)])
def test_selfstart_td_gauss_vs_oracle(torch, dim, scheme):
    """td-gauss (SLDA, V_ext switched on at the start so the states move)
    on 1D / 2D / 3D lattices: the self-started state, history, densities and
    potentials, and three ABM steps after it, against the oracle."""
    say("case: selfstart td-gauss dim=%d scheme=%d" % (dim, scheme))
    s = tdgauss.setup(dim)
    cfg = s["cfg"]
    cfg.scheme = scheme
    cfg.vext.t_start, cfg.vext.t_switch = -1.0, 0.5
    run_pair(torch, cfg, s["wf"], s["fbeta"], s["kt"], s["cnt"], s["pots"], fused=(dim == 3))


def test_selfstart_aslda_and_filter_vs_oracle(torch):
    """ASLDA (varying mass: gradients every application, lap(alpha psi)) and
    the mean-field filter at hkf_mode 2, which every apply_hamiltonian call of
    the self-start applies again in place, on the 1D td-gauss problem."""
    say("case: selfstart ASLDA 1D with hkf_mode 2")
    s = tdgauss.setup(1)
    cfg = s["cfg"]
    cfg.functional = po.ASLDA
    cfg.mf_correction = 0
    cfg.hkf_mode, cfg.hkf_mu, cfg.hkf_T = 2, 3.0, 0.5
    cfg.vext.t_start, cfg.vext.t_switch = -1.0, 0.5
    run_pair(torch, cfg, s["wf"], s["fbeta"], s["kt"], s["cnt"], s["pots"], fused=False)
