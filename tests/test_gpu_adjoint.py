"""Adjoint gradients against central finite differences of the forward solver.

The observable is the particle number inside a disk (2D) / ball (3D) /
interval (1D) after tens of PECE steps with two Gaussian beams switched on;
the gradient is with respect to every Gaussian parameter (amplitude, centre,
exponent coefficients).  Bar: |g_adjoint - g_fd| <= 1e-5 max|g_fd|, with
central differences of step h = 1e-6 in each parameter."""
import numpy as np
import pytest

from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-5


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


CASES = [
    # dim, shape, states, scheme, steps, K, segments taped by the forward run
    (2, (16, 16, 8), 64, 55, 30, 8, 2),
    (3, (8, 8, 8), 48, 45, 20, 6, 1),
    (1, (32, 8, 8), 48, 34, 24, 10, 1),
]


def setup_problem(dim, shape, nst, scheme, steps, K, taped=1, tape_history=False):
    from wslda_amd import adjoint, problems
    s = problems.uniform_start(dim, shape, n_states=nst, scheme=scheme)
    p = s["params"]
    eF = 0.5
    L = shape[0] * p.dx
    # two beams, off-centre, a disk / ball / interval around the origin
    theta = np.array([0.15 * eF, 0.12 * L, -0.08 * L, 2.0, -0.1 * eF, -0.1 * L, 0.15 * L, 2.5])
    if dim == 1:
        theta[[2, 6]] = 0.0
    g = adjoint.disk_weights(p, (0.0, 0.0, 0.0), 0.3 * L)
    prob = adjoint.AdjointProblem(p, s["wf"], s["pots"], s["fbeta"], s["kt"], s["cnt"], nsteps=steps, K=K,
                                  dens_grad=g, taped_segments=taped, tape_history=tape_history)
    return prob, theta


def gauss_slots(theta, dim):
    """Gaussian parameters of the beams as the library holds them."""
    from wslda_amd import adjoint
    v = adjoint.beams_vext(theta)
    if dim < 3:
        for k in range(v.ngauss):
            v.wz[k] = 0.0
    if dim < 2:
        for k in range(v.ngauss):
            v.wy[k] = 0.0
    return v


def observable_at(prob, v):
    prob.set_vext(v)
    prob.forward(keep_checkpoints=False)
    return prob.observable()


@pytest.mark.parametrize("dim,shape,nst,scheme,steps,K,taped", CASES)
def test_adjoint_vs_finite_differences(torch, dim, shape, nst, scheme, steps, K, taped):
    say("case: adjoint vs FD, dim=%d shape=%s states=%d scheme=%d steps=%d K=%d taped=%d" % (
        dim, shape, nst, scheme, steps, K, taped))
    prob, theta = setup_problem(dim, shape, nst, scheme, steps, K, taped)
    v0 = gauss_slots(theta, dim)
    prob.set_vext(v0)
    prob.forward()
    O0 = prob.observable()
    g = prob.backward()
    # central differences on every active parameter slot
    names = ["amp", "cx", "cy", "cz", "wx", "wy", "wz"]
    active = [0, 1, 4] + ([2, 5] if dim >= 2 else [])
    worst = 0.0
    gmax = 0.0
    rows = []
    for k in range(2):
        for q in active:
            h = 1e-6
            vals = []
            for sgn in (+1, -1):
                v = gauss_slots(theta, dim)
                arr = getattr(v, names[q])
                arr[k] += sgn * h
                vals.append(observable_at(prob, v))
            fd = (vals[0] - vals[1]) / (2 * h)
            rows.append((k, names[q], g[k, q], fd))
            gmax = max(gmax, abs(fd))
    for k, nm, ga, fd in rows:
        say("  beam %d %-3s adjoint %+.10e  fd %+.10e" % (k, nm, ga, fd))
        worst = max(worst, abs(ga - fd))
    say("  O = %.12e, max |g_adj - g_fd| / max|g_fd| = %.2e" % (O0, worst / gmax))
    assert gmax > 0
    assert worst <= GRAD_TOL * gmax
    # the run is unchanged by the backward sweep
    prob.set_vext(v0)
    assert abs(observable_at(prob, v0) - O0) <= 1e-14 * abs(O0)
    prob.close()


def test_autograd_function_matches_adjoint(torch):
    """torch.autograd through vext_observable returns the adjoint gradient
    mapped to the beam parameters (amplitude, x, y, width)."""
    say("case: autograd Function, 2D 16x16")
    from wslda_amd import adjoint
    prob, theta = setup_problem(2, (16, 16, 8), 64, 55, 20, 7)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    O = adjoint.vext_observable(th, prob)
    O.backward()
    g_auto = th.grad.numpy().copy()
    # by hand: Gaussian-slot gradient, chain rule to widths
    prob.set_vext(adjoint.beams_vext(theta))
    prob.forward()
    g = adjoint.beams_grad(theta, prob.backward())
    assert np.abs(g_auto - g).max() <= 1e-12 * np.abs(g).max()
    # width derivative against a finite difference in theta itself
    h = 1e-6
    tp, tm = theta.copy(), theta.copy()
    tp[3] += h
    tm[3] -= h
    fd = (adjoint.vext_observable(torch.tensor(tp), prob).item() - adjoint.vext_observable(torch.tensor(tm), prob).item()) / (2 * h)
    assert abs(fd - g_auto[3]) <= GRAD_TOL * np.abs(g_auto).max()
    prob.close()


@pytest.mark.parametrize("dim,shape,nst,scheme,steps,K,taped", CASES)
def test_light_tape_matches_full_tape(torch, dim, shape, nst, scheme, steps, K, taped):
    """A tape of y alone, whose history entries the backward sweep rebuilds
    (tdw_adj_history_entry, one H application per step, across segment
    boundaries too), gives the gradient of the tape that keeps the entries:
    the rebuild runs the H application the taped step ran, on the same
    inputs.  One more segment is taped than in the FD cases, so that a taped
    segment starts from the tape of the one before it, and a second run tapes
    every segment (no recompute at all)."""
    for tp in (taped + 1, 99):
        say("case: light vs full tape, dim=%d shape=%s scheme=%d steps=%d K=%d taped=%d" % (
            dim, shape, scheme, steps, K, tp))
        out = {}
        for full in (True, False):
            prob, theta = setup_problem(dim, shape, nst, scheme, steps, K, tp, tape_history=full)
            prob.set_vext(gauss_slots(theta, dim))
            prob.forward()
            g = prob.backward()
            out[full] = (g, prob.lambda_y0.clone(), len(prob.ckpts))
            prob.close()
        g1, g0 = out[True][0], out[False][0]
        l1, l0 = out[True][1], out[False][1]
        dg = np.abs(g1 - g0).max() / np.abs(g1).max()
        dl = float((l1 - l0).abs().max() / l1.abs().max())
        say("  max rel diff: gradient %.2e, dL/dy0 %.2e; checkpoints kept %d (full) / %d (light)" % (
            dg, dl, out[True][2], out[False][2]))
        assert dg <= 1e-13 and dl <= 1e-13
        # the light tape keeps no checkpoint for a taped segment after another
        assert out[False][2] < out[True][2]


def test_fused_taped_step_matches_modular_step(torch):
    """On 3D lattices of the fused step the taped forward step runs the fused
    FP64 passes (finalised at each step's end).  From the same start, its
    tape agrees with the modular step (tdw_step) step by step to 1e-12: y
    after the step, the new history entry, the densities and potentials of
    the corrected state and its <H>."""
    from wslda_amd import td
    say("case: fused taped step vs modular step, 3D 8^3")
    steps, K = 6, 6
    prob, theta = setup_problem(3, (8, 8, 8), 48, 45, steps, K, 99, tape_history=True)
    prob.set_vext(gauss_slots(theta, 3))
    prob.forward()
    ty, tf, tscal, tgrid, _ = prob._tape(prob.taped[0])
    R = td.TDRun(prob.ctx, prob.wf0.clone(), prob.pots0.clone().cpu().numpy())
    R.initial_measure()
    n = prob.n
    worst = 0.0
    for k in range(steps):
        R.step(prob.it0 + k, grad_last=False)
        prob.ctx.sync()
        errs = [float((ty[k + 1] - R.wf).abs().max() / R.wf.abs().max()),
                float((tf[k] - R.history()[0]).abs().max() / R.history()[0].abs().max()),
                float((tgrid[k, 2] - R.dens).abs().max() / R.dens.abs().max()),
                float((tgrid[k, 3] - R.pots).abs().max() / R.pots.abs().max()),
                float((tscal[k, 3 * n:] - R.qpe).abs().max() / R.qpe.abs().max())]
        worst = max(worst, max(errs))
    say("  max rel diff over %d steps (y, f_new, dens, pots, <H>): %.2e" % (steps, worst))
    assert worst <= 1e-12
    prob.close()
