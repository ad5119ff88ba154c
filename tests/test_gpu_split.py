"""Split-precision fused step against the FP64 fused step (itself held to the
oracle at 1e-12 by test_gpu_fused).

The split form runs the FFTs and the kinetic operator in FP32 and stores the
ABM history in FP32; the state, every ABM update of it, the densities, norms
and <H> are FP64.  Bars: relative error <= 1e-6 for every density field
(max-abs difference over max-abs value; |nu| through the complex difference)
and <= 1e-8 relative for the total energy and particle number, after tens of
steps with an external potential switching on."""
import numpy as np
import pytest

from tests import tdgauss
from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu

DENS_TOL = 1e-6
ENERGY_TOL = 1e-8


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def field_errors(d, ref, N, currents=True):
    """Relative error per density field of the gate (rho_a, rho_b, |nu|,
    tau_a, tau_b, j): max-abs difference over the field's own max-abs value
    (|nu| through the complex difference; the current j = j_a + j_b, the
    total current, as one vector field).  currents=False leaves j out, for
    workloads whose currents are zero up to noise (a uniform state, or before
    an external potential has switched on), where a ratio to their own
    maximum measures nothing."""
    out = {}
    nu, nur = d[0:2 * N:2] + 1j * d[1:2 * N:2], ref[0:2 * N:2] + 1j * ref[1:2 * N:2]
    out["nu"] = np.abs(nu - nur).max() / np.abs(nur).max()
    names = {"rho_a": 2, "tau_a": 3, "rho_b": 7, "tau_b": 8}
    for k, i in names.items():
        r = ref[i * N:(i + 1) * N]
        out[k] = np.abs(d[i * N:(i + 1) * N] - r).max() / np.abs(r).max()
    if currents:
        j, jr = d[4 * N:7 * N] + d[9 * N:12 * N], ref[4 * N:7 * N] + ref[9 * N:12 * N]
        out["j"] = np.abs(j - jr).max() / np.abs(jr).max()
    return out


def spin_current_errors(d, ref, N):
    """j_a and j_b each against its own maximum (reported beside the gate)."""
    out = {}
    for k, i0 in (("j_a", 4), ("j_b", 9)):
        r = ref[i0 * N:(i0 + 3) * N]
        out[k] = np.abs(d[i0 * N:(i0 + 3) * N] - r).max() / np.abs(r).max()
    return out


def make_runs(torch, p_fp64, p_split, wf, fbeta, pots, kt=None, cnt=None):
    from wslda_amd import td
    n = wf.shape[0] // 2
    runs = []
    for p in (p_fp64, p_split):
        ctx = td.TDContext(p, n, 0)
        ctx.set_states(fbeta, kt, cnt)
        R = td.TDRun(ctx, wf.clone() if hasattr(wf, "clone") else wf, pots, fused=True)
        assert td.lib().tdw_fused_supported(ctx.h) == 1
        R.initial_measure()
        runs.append((ctx, R))
    return runs


def test_split_vs_fp64_td_gauss_3d(torch):
    """td-gauss 12^3 (AB4AM5), 30 steps from t eF = 5.5, gradient steps every
    10: densities, energy and N of the split run against FP64.  The smooth
    switch of V_ext is still ~1e-10 of its amplitude over these steps, so the
    currents are noise and are gated on the vortex workload below."""
    say("case: split vs fp64, td-gauss 3D 12^3, 30 steps")
    from wslda_amd import td
    s = tdgauss.setup(3)
    it0 = int(round(5.5 / s["eF"] / s["cfg"].dt))
    runs = make_runs(torch, tdgauss.to_params(s["cfg"], td.FP64), tdgauss.to_params(s["cfg"], td.SPLIT),
                     s["wf"], s["fbeta"], s["pots"])
    N = runs[0][0].N
    for k in range(30):
        grad = (k % 10 == 9)
        for ctx, R in runs:
            R.step(it0 + k, grad_last=grad)
        if grad:
            d = [R.dens.cpu().numpy() for _, R in runs]
            err = field_errors(d[1], d[0], N, currents=False)
            say("  step %d: %s" % (k, {a: "%.2e" % b for a, b in err.items()}))
            assert max(err.values()) <= DENS_TOL, err
            e = [R.energy(it0 + k + 1) for _, R in runs]
            et = [x[:7].sum() for x in e]
            assert abs(et[1] - et[0]) <= ENERGY_TOL * abs(et[0]), et
            assert abs(e[1][7] + e[1][8] - e[0][7] - e[0][8]) <= ENERGY_TOL * abs(e[0][7] + e[0][8])
    for ctx, _ in runs:
        ctx.close()


def test_split_vs_fp64_24cubed(torch):
    """The benchmark lattice on 1024 of its states, 40 steps: densities,
    energy, normalisation and the FP32 history against FP64."""
    say("case: split vs fp64, 24^3, 1024 states, 40 steps")
    from wslda_amd import problems, td
    w = problems.config2(torch, first=1024)
    p64 = w["params"]
    psp = td.Params.from_buffer_copy(p64)
    psp.precision = td.SPLIT
    runs = make_runs(torch, p64, psp, w["wf"], w["fbeta"], w["pots"])
    N = runs[0][0].N
    for it in range(40):
        for ctx, R in runs:
            R.step(it, grad_last=(it % 20 == 19))
    d = [R.dens.cpu().numpy() for _, R in runs]
    err = field_errors(d[1], d[0], N, currents=False)
    say("  %s" % {a: "%.2e" % b for a, b in err.items()})
    assert max(err.values()) <= DENS_TOL, err
    e = [R.energy(40) for _, R in runs]
    assert abs(e[1][:7].sum() - e[0][:7].sum()) <= ENERGY_TOL * abs(e[0][:7].sum())
    # states are normalised after a gradient step (norm sums in FP64)
    ys = runs[1][1].wf
    nrm = (ys.abs() ** 2).sum(dim=1).cpu().numpy()
    n = nrm.size // 2
    dv = p64.dx * p64.dy * p64.dz
    assert np.abs((nrm[:n] + nrm[n:]) * dv - 1.0).max() < 1e-12
    # history: FP32 entries against FP64 ones, on the scale of H psi
    h64 = runs[0][1].history()
    hsp = runs[1][1].history()
    emax = 3 * 0.5 * (np.pi / p64.dx) ** 2
    scale = runs[0][1].wf.abs().max().item() * emax
    for j in range(5):
        diff = (hsp[j].to(torch.complex128) - h64[j]).abs().max().item()
        assert diff <= 1e-5 * scale, (j, diff, scale)
    for ctx, _ in runs:
        ctx.close()


def test_split_rejects_unfused_paths(torch):
    """The split form exists only as the fused 3D step: the modular step and
    lattices without fused kernels report TDW_EINVAL instead of running FP64
    on FP32 history."""
    say("case: split mode rejects modular / unsupported steps")
    from wslda_amd import td
    s = tdgauss.setup(3)
    ctx = td.TDContext(tdgauss.to_params(s["cfg"], td.SPLIT), s["wf"].shape[0] // 2, 0)
    ctx.set_states(s["fbeta"])
    R = td.TDRun(ctx, s["wf"], s["pots"], fused=True)
    R.initial_measure()
    import ctypes as C
    rc = td.lib().tdw_step(ctx.h, C.byref(R.st), 0, 0)
    assert rc == td.TDW_EINVAL
    ctx.close()
    s1 = tdgauss.setup(1)
    ctx = td.TDContext(tdgauss.to_params(s1["cfg"], td.SPLIT), s1["wf"].shape[0] // 2, 0)
    ctx.set_states(s1["fbeta"], s1["kt"], s1["cnt"])
    R = td.TDRun(ctx, s1["wf"], s1["pots"], fused=True)
    R.initial_measure()
    assert td.lib().tdw_fused_supported(ctx.h) == 0
    rc = td.lib().tdw_step_fused(ctx.h, C.byref(R.st), 0, 0)
    assert rc == td.TDW_EINVAL
    ctx.close()


def test_split_vs_fp64_trap_vortex_currents(torch):
    """The 32^3 trapped-vortex workload (trap 0.1 eF, vortex phase on u) on
    its first 1024 states, 40 steps with gradient steps every 20:
    every density field of the split run, the current j = j_a + j_b included
    and measured against its own maximum (the vortex carries real currents),
    within 1e-6 of the FP64 run; total energy within 1e-8.  j_a and j_b
    against their own maxima are printed beside the gate: the vortex phase
    sits on the u components, so |j_b| is much smaller than |j_a| and its
    own-maximum ratio is the larger one."""
    say("case: split vs fp64, 32^3 trap + vortex, 1024 states, 40 steps, j against |j|_inf")
    from wslda_amd import problems, td
    w = problems.config3(torch, precision=td.FP64, first=1024)
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
            jmax = np.abs(d[0][4 * N:7 * N]).max()
            jbmax = np.abs(d[0][9 * N:12 * N]).max()
            say("  step %d: |j_a|_inf = %.3e, |j_b|_inf = %.3e, %s, per spin %s"
                % (it + 1, jmax, jbmax, {a: "%.2e" % b for a, b in err.items()},
                   {a: "%.2e" % b for a, b in spin_current_errors(d[1], d[0], N).items()}))
            assert jmax > 1e-4 * d[0][2 * N:3 * N].max()   # the currents are real
            assert max(err.values()) <= DENS_TOL, err
            e = [R.energy(it + 1) for _, R in runs]
            et = [x[:7].sum() for x in e]
            assert abs(et[1] - et[0]) <= ENERGY_TOL * abs(et[0]), et
    for ctx, _ in runs:
        ctx.close()
