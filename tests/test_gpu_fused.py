"""The fused PECE step (six passes per step) against the oracle and against
the modular step, FP64, to the 1e-12 bar."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss
from tests.test_gpu_parity import gpu_run, rel, say, TOL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def dens_check(d, ref, N):
    for sl, name in [(slice(0, 2 * N), "nu"), (slice(2 * N, 3 * N), "rho_a"), (slice(3 * N, 4 * N), "tau_a"),
                     (slice(7 * N, 8 * N), "rho_b"), (slice(8 * N, 9 * N), "tau_b")]:
        if np.abs(ref[sl]).max() == 0.0:
            assert np.abs(d[sl]).max() == 0.0, name
        else:
            assert rel(d[sl], ref[sl]) < TOL, name


@pytest.mark.parametrize("scheme", [po.AB3AM4, po.AB4AM5, po.AB5AM5])
def test_fused_vs_oracle_td_gauss_3d(torch, scheme):
    say("case: fused step vs oracle, td-gauss 3D 12^3, scheme %d" % scheme)
    from wslda_amd import td
    s = tdgauss.setup(3, scheme=scheme)
    it0 = int(round(6.0 / s["eF"] / s["cfg"].dt))
    P = po.Problem(s["cfg"], s["wf"], s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    P.initial_measure()
    ctx, R = gpu_run(torch, s, fused=True)
    assert td.lib().tdw_fused_supported(ctx.h) == 1
    R.initial_measure()
    nsteps = 7
    N = ctx.N
    for k in range(nsteps):
        grad = (k == 3)            # a measurement step in the middle
        P.step(it0 + k, grad_last=grad)
        R.step(it0 + k, grad_last=grad)
        if k in (2, 3, 6):
            ctx.sync()
            dens_check(R.dens.cpu().numpy(), P.dens, N)
    R.finalize()
    ctx.sync()
    w = R.wf.cpu().numpy()
    d = np.abs(w - P.wf)
    i, j = np.unravel_index(d.argmax(), d.shape)
    assert rel(w, P.wf) < 1e-11, ("max diff at state-row %d point %d: %r vs %r" % (i, j, w[i, j], P.wf[i, j]))
    hist_ref = P.history()
    hist = R.history()
    k = {po.AB3AM4: 3, po.AB4AM5: 4, po.AB5AM5: 5}[scheme]
    # f = (H - <H>) psi of near-eigenstates is a small difference of O(E psi)
    # terms: compare on the scale of H psi (max|psi| times the largest
    # kinetic energy on the lattice), not of f itself
    emax = 3 * 0.5 * (np.pi / s["cfg"].dx) ** 2
    scale = np.abs(P.wf).max() * emax
    for j in range(k - 1):
        assert np.abs(hist[j].cpu().numpy() - hist_ref[j]).max() < 1e-12 * scale, j
    e_ref = P.energy(it0 + nsteps)
    e = R.energy(it0 + nsteps)
    assert abs(e[7] + e[8] - e_ref[7] - e_ref[8]) <= TOL * abs(e_ref[7] + e_ref[8])
    ctx.close()


def test_fused_vs_modular_24cubed(torch):
    """The benchmark lattice (24^3) on a slice of its states: fused and
    modular steps agree, including a measurement step."""
    say("case: fused vs modular, 24^3, 512 states")
    from wslda_amd import problems, td
    w = problems.config2(torch, first=512)
    p = w["params"]
    n = w["wf"].shape[0] // 2
    ctxs, runs = [], []
    for fused in (False, True):
        ctx = td.TDContext(p, n, 0)
        ctx.set_states(w["fbeta"], w["kt"], w["cnt"])
        R = td.TDRun(ctx, w["wf"].clone(), w["pots"], fused=fused)
        R.initial_measure()
        ctxs.append(ctx)
        runs.append(R)
    assert td.lib().tdw_fused_supported(ctxs[1].h) == 1
    for it in range(12):
        for R in runs:
            R.step(it, grad_last=(it == 5))
    runs[1].finalize()
    for c in ctxs:
        c.sync()
    N = ctxs[0].N
    dens_check(runs[1].dens.cpu().numpy(), runs[0].dens.cpu().numpy(), N)
    assert rel(runs[1].wf.cpu().numpy(), runs[0].wf.cpu().numpy()) < 1e-11
    assert rel(runs[1].qpe.cpu().numpy(), runs[0].qpe.cpu().numpy()) < 1e-11
    for c in ctxs:
        c.close()


def test_timing_report(torch):
    say("case: per-kernel timing report")
    from wslda_amd import problems, td
    w = problems.config2(torch, first=256)
    ctx = td.TDContext(w["params"], 256, 0)
    ctx.set_states(w["fbeta"])
    R = td.TDRun(ctx, w["wf"], w["pots"], fused=True)
    R.initial_measure()
    ctx.timing(True)
    for it in range(3):
        R.step(it)
    rep = ctx.timing_report()
    ctx.timing(False)
    assert "f_pred" in rep and "f_hyz" in rep and "f_corr" in rep
    assert rep["f_pred"][1] == 3 and rep["f_hyz"][1] == 6
    assert all(v[0] > 0 for v in rep.values())
    ctx.close()
