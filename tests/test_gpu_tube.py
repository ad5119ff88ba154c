"""The elongated-tube Josephson junction (32 x 32 x 96 lattice, Gaussian
barrier of 0.3 eF and width 3 dx across the tube, phase step of pi imprinted
on u across it; wslda_amd.problems.config5) on the GPU:

- the fused FP64 step (x-pencils of 32, 32 x 96 yz planes in LDS) and the
  modular step against the oracle at 1e-12 on a slice of the states;
- the split-precision step against FP64 on 512 states: every density field
  (the current j = j_a + j_b against its own maximum) within 1e-6, energy
  within 1e-8;
- the adjoint gradient of the final-time current through the barrier plane
  with respect to the barrier's height and width against central finite
  differences (h = 1e-6) to 1e-5."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.test_gpu_configs import check_step_against_oracle, oracle_of
from tests.test_gpu_parity import TOL, rel, say
from tests.test_gpu_split import DENS_TOL, ENERGY_TOL, field_errors, make_runs, spin_current_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


@pytest.mark.parametrize("fused", [True, False])
def test_tube_slice_vs_oracle(torch, fused):
    """24 states of the tube workload, 4 PECE steps with gradient steps at 2
    and 4, fused or modular, against the oracle."""
    say("case: tube 32x32x96, 24 states, 4 steps, fused=%s vs oracle" % fused)
    from wslda_amd import problems, td
    w = problems.config5(torch, precision=td.FP64, first=24)
    p = w["params"]
    n = w["wf"].shape[0] // 2
    cfg, P = oracle_of(w)
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(w["fbeta"])
    assert td.lib().tdw_fused_supported(ctx.h) == 1
    R = td.TDRun(ctx, w["wf"], w["pots"], fused=fused)
    P.initial_measure()
    R.initial_measure()
    for it in range(4):
        grad = it in (1, 3)
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


def test_tube_split_vs_fp64(torch):
    """512 states of the tube workload, 40 steps with gradient steps every 20:
    split precision against FP64 (the phase step drives a current through the
    barrier)."""
    say("case: tube 32x32x96, split vs fp64, 512 states, 40 steps")
    from wslda_amd import problems, td
    w = problems.config5(torch, precision=td.FP64, first=512)
    p64 = w["params"]
    psp = td.Params.from_buffer_copy(p64)
    psp.precision = td.SPLIT
    runs = make_runs(torch, p64, psp, w["wf"], w["fbeta"], w["pots"])
    N = runs[0][0].N
    for it in range(40):
        grad = (it % 20 == 19)
        for ctx, R in runs:
            R.step(it, grad_last=grad)
        if grad:
            d = [R.dens.cpu().numpy() for _, R in runs]
            err = field_errors(d[1], d[0], N)
            jmax = np.abs(d[0][4 * N:7 * N] + d[0][9 * N:12 * N]).max()
            say("  step %d: |j|_inf = %.3e, %s, per spin %s"
                % (it + 1, jmax, {a: "%.2e" % b for a, b in err.items()},
                   {a: "%.2e" % b for a, b in spin_current_errors(d[1], d[0], N).items()}))
            assert jmax > 1e-4 * d[0][2 * N:3 * N].max()
            assert max(err.values()) <= DENS_TOL, err
            e = [R.energy(it + 1) for _, R in runs]
            et = [x[:7].sum() for x in e]
            assert abs(et[1] - et[0]) <= ENERGY_TOL * abs(et[0]), et
    for ctx, _ in runs:
        ctx.close()


def test_tube_barrier_gradient_vs_finite_differences(torch):
    """The final-time current through the barrier plane (z = 0) after 24
    FP64 steps from the tube start (64 states, checkpoints every 6): its
    gradient in (height, width) of the barrier by the adjoint sweep against
    central differences with h = 1e-6, to 1e-5 of the largest component."""
    say("case: tube 32x32x96, 64 states, 24 steps, K = 6: barrier gradient of the current vs FD")
    from wslda_amd import adjoint, problems, td
    w = problems.config5(torch, precision=td.FP64, first=64)
    p = w["params"]
    theta = np.array([0.3 * w["eF"], 3.0])
    prob = adjoint.AdjointProblem(p, w["wf"], w["pots"], w["fbeta"], nsteps=24, K=6, taped_segments=1)
    prob.state_observable = adjoint.PlaneCurrent(prob.ctx, w["fbeta"], axis=2)
    prob.set_vext(adjoint.barrier_vext(theta))
    prob.forward()
    O = prob.observable()
    grad = adjoint.barrier_grad(theta, prob.backward())
    h = 1e-6
    fd = np.zeros(2)
    for k in range(2):
        vals = []
        for sg in (1, -1):
            th = theta.copy()
            th[k] += sg * h
            prob.set_vext(adjoint.barrier_vext(th))
            prob.forward(keep_checkpoints=False)
            vals.append(prob.observable())
        fd[k] = (vals[0] - vals[1]) / (2 * h)
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    say("  current %.6e\n  adjoint %s\n  fd      %s\n  max rel err %.2e" % (
        O, np.array2string(grad, precision=8), np.array2string(fd, precision=8), err))
    assert abs(O) > 0.0
    assert err <= 1e-5
    prob.close()


def test_tube_fused_and_split_forward_gradient(torch):
    """The same barrier gradient with the untaped segments of the forward
    run on the fused step: FP64 (against the modular forward, 1e-8) and
    split precision with FP32-history checkpoints (against the modular FP64
    gradient, 1e-5; the backward sweep is FP64 in all three)."""
    say("case: tube 32x32x96, 64 states, 24 steps, K = 6: modular / fused FP64 / split forward, adjoint gradient")
    from wslda_amd import adjoint, problems, td
    w = problems.config5(torch, precision=td.FP64, first=64)
    theta = np.array([0.3 * w["eF"], 3.0])
    res = {}
    for name, fp in (("modular", None), ("fused", td.FP64), ("split", td.SPLIT)):
        p = td.Params.from_buffer_copy(w["params"])
        prob = adjoint.AdjointProblem(p, w["wf"], w["pots"], w["fbeta"], nsteps=24, K=6, taped_segments=1,
                                      forward_precision=fp)
        prob.state_observable = adjoint.PlaneCurrent(prob.ctx, w["fbeta"], axis=2)
        prob.set_vext(adjoint.barrier_vext(theta))
        prob.forward()
        if fp == td.SPLIT:
            # the checkpoints keep the FP32 history
            assert all(t.dtype == torch.complex64 for t in prob.ckpts[1][2])
        O = prob.observable()
        res[name] = (O, adjoint.barrier_grad(theta, prob.backward()))
        prob.close()
        say("  %-8s O = %.10e  grad %s" % (name, O, np.array2string(res[name][1], precision=9)))
    O0, g0 = res["modular"]
    for name, tol in (("fused", 1e-8), ("split", 1e-5)):
        O, g = res[name]
        eo, eg = abs(O - O0) / abs(O0), np.abs(g - g0).max() / np.abs(g0).max()
        say("  %s vs modular: O %.2e, gradient %.2e" % (name, eo, eg))
        assert eo <= tol and eg <= tol, (name, eo, eg)
