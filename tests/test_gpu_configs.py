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


def oracle_of(w):
    """The oracle's copy of a workload built on the GPU (host arrays)."""
    p = w["params"]
    cfg = po.Cfg()
    C.memmove(C.addressof(cfg), C.addressof(p), C.sizeof(cfg))
    wf = w["wf"].cpu().numpy() if hasattr(w["wf"], "cpu") else w["wf"]
    return cfg, po.Problem(cfg, wf, w["fbeta"], kt=w.get("kt"), cnt=w.get("cnt"), pots=w["pots"])


def check_step_against_oracle(R, P, it, what):
    d = R.dens.cpu().numpy()
    assert rel(d, P.dens) < TOL, (what, it, rel(d, P.dens))
    assert rel(R.pots.cpu().numpy()[:4 * P.N], P.pots[:4 * P.N]) < TOL, (what, it)


@pytest.mark.parametrize("fused", [False, True])
def test_config2_24cubed_slice_vs_oracle(torch, fused):
    """The 24^3 SLDA workload (AB5AM5, seeded 1e-6 noise) on the first 96
    states of its state set: 6 PECE steps with gradient steps at 2 and 5,
    modular and fused, against the oracle at 1e-12 (densities with tau and
    j, potentials, <H>, the states and the energies)."""
    say("case: config 2 (24^3, first 96 states), fused=%d, 6 steps vs oracle" % fused)
    from wslda_amd import problems, td
    w = problems.config2(torch, first=96)
    p = w["params"]
    n = w["wf"].shape[0] // 2
    cfg, P = oracle_of(w)
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(w["fbeta"])
    R = td.TDRun(ctx, w["wf"], w["pots"], fused=fused)
    P.initial_measure()
    R.initial_measure()
    check_step_against_oracle(R, P, -1, "initial")
    for it in range(6):
        grad = it in (2, 5)
        P.step(it, grad_last=grad)
        R.step(it, grad_last=grad)
        if fused:
            R.finalize()
        check_step_against_oracle(R, P, it, "step")
        assert rel(R.qpe.cpu().numpy(), P.qpe) < TOL, it
        if grad:
            e, er = R.energy(it + 1), P.energy(it + 1)
            assert abs(e[:7].sum() - er[:7].sum()) <= TOL * abs(er[:7].sum()), it
            assert abs(e[7] + e[8] - er[7] - er[8]) <= TOL * abs(er[7] + er[8]), it
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    ctx.close()


def test_config2_full_start_is_stationary(torch):
    """All 16308 states of the 24^3 uniform solution (no noise): the start is
    the solution's own state set, so (H - <H>) psi = 0 and 20 fused PECE
    steps keep the energy to |dE|/E <= 1e-12 and the particle number to
    1e-12; <H> equals the states' BdG energies."""
    say("case: config 2 full state set (16308 states, no noise), 20 fused steps, stationarity")
    from wslda_amd import problems, td
    w = problems.config2(torch, noise=0.0)
    p = w["params"]
    n = w["wf"].shape[0] // 2
    assert n == 16308
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(w["fbeta"])
    R = td.TDRun(ctx, w["wf"], w["pots"], fused=True)
    R.initial_measure()
    e0 = R.energy(0)
    E = [e0[:7].sum()]
    for it in range(20):
        R.step(it, grad_last=(it % 10 == 9))
        if it % 10 == 9:
            e = R.energy(it + 1)
            E.append(e[:7].sum())
            say("  step %d: E = %.15f, |dE|/E = %.2e, dN/N = %.2e" % (
                it + 1, E[-1], abs(E[-1] - E[0]) / abs(E[0]), abs(e[7] + e[8] - e0[7] - e0[8]) / (e0[7] + e0[8])))
            assert abs(E[-1] - E[0]) <= 1e-12 * abs(E[0])
            assert abs(e[7] + e[8] - e0[7] - e0[8]) <= 1e-12 * (e0[7] + e0[8])
    R.finalize()
    ctx.sync()
    # <H>_n of a plane-wave BdG state is its quasiparticle energy E_n
    En = np.abs(np.asarray(w["E"]))
    qpe = R.qpe.cpu().numpy()
    assert np.abs(np.abs(qpe) - En).max() <= 1e-10 * En.max()
    ctx.close()


def test_config3_trap_vortex_slice_vs_oracle(torch):
    """The 32^3 trapped-vortex workload (trap 0.1 eF, vortex phase on u,
    seeded noise in a Gaussian envelope) on the first 48 of its states, FP64:
    5 fused PECE steps with gradient steps at 1 and 4 against the oracle at
    1e-12."""
    say("case: config 3 (32^3 trap + vortex, 48 states), FP64 fused, 5 steps vs oracle")
    from wslda_amd import problems, td
    w = problems.config3(torch, precision=td.FP64, first=48)
    p = w["params"]
    n = w["wf"].shape[0] // 2
    cfg, P = oracle_of(w)
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(w["fbeta"])
    R = td.TDRun(ctx, w["wf"], w["pots"], fused=True)
    P.initial_measure()
    R.initial_measure()
    for it in range(5):
        grad = it in (1, 4)
        P.step(it, grad_last=grad)
        R.step(it, grad_last=grad)
        R.finalize()
        check_step_against_oracle(R, P, it, "step")
        if grad:
            e, er = R.energy(it + 1), P.energy(it + 1)
            for k in (po.EKIN, po.EPOT, po.EPAIR, po.EPOTEXT):
                assert abs(e[k] - er[k]) <= TOL * max(abs(er[k]), abs(er[:7]).max()), (it, k)
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    ctx.close()


def test_config4_adjoint_64x64_vs_finite_differences(torch):
    """The adjoint workload's own lattice (2D 64x64, 8-point transverse box,
    AB5AM5, two seeded Gaussian beams, N inside a disk of radius 10 dx) on
    256 states and 24 steps with checkpoints every 6: the gradient in all 8
    beam parameters (amplitude, x, y, width) against central differences
    with h = 1e-6, to 1e-5 of the largest component."""
    say("case: config 4 lattice (2D 64x64, 256 states, 24 steps, K = 6), adjoint vs FD")
    from wslda_amd import adjoint, problems
    s = problems.config4(n_states=256)
    p = s["params"]
    theta = s["theta"]
    g = adjoint.disk_weights(p, s["center"], s["radius"])
    prob = adjoint.AdjointProblem(p, s["wf"], s["pots"], s["fbeta"], s["kt"], s["cnt"], nsteps=24, K=6,
                                  dens_grad=g, taped_segments=1)
    prob.set_vext(adjoint.beams_vext(theta))
    prob.forward()
    grad = adjoint.beams_grad(theta, prob.backward())
    h = 1e-6
    fd = np.zeros(8)
    for k in range(8):
        vals = []
        for sg in (1, -1):
            th = theta.copy()
            th[k] += sg * h
            prob.set_vext(adjoint.beams_vext(th))
            prob.forward(keep_checkpoints=False)
            vals.append(prob.observable())
        fd[k] = (vals[0] - vals[1]) / (2 * h)
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    say("  adjoint %s\n  fd      %s\n  max rel err %.2e" % (np.array2string(grad, precision=6),
                                                           np.array2string(fd, precision=6), err))
    assert err <= 1e-5
    prob.close()
