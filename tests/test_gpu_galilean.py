"""Galilean boost of the uniform ASLDA / SLDAE / SLDA solutions through the
HIP library: the closed forms of tests/galilean.py at 1e-12 (densities with
gradients, the potentials with their current terms, H psi with the A.grad and
effective-mass terms, energies), and PECE steps of the boosted ASLDA state
from its exact history keep the flow.  3D 12^3 runs the two-pass ASLDA
kernels (gradients formed in LDS)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import galilean as G
from tests import tdgauss
from tests.test_gpu_parity import say
from tests.test_oracle_galilean import NX, solution

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def gpu_measure(torch, ctx, wf, pots):
    """Densities (with gradients), potentials, H psi without the <H>
    subtraction, and energies through the library."""
    n = wf.shape[0] // 2
    W = torch.as_tensor(wf).to("cuda:0")
    d = [torch.zeros_like(W) for _ in range(4)]
    ctx.compute_derivatives(W, *d)
    D = torch.zeros(12 * ctx.N, dtype=torch.float64, device="cuda:0")
    ctx.calculate_densities(W, d[0], d[1], d[2], d[3], D, 1)
    ctx.density_caculate_tau(D)
    P = torch.as_tensor(pots).to("cuda:0")
    ctx.compute_potentials(0, D, P)
    H = torch.zeros_like(W)
    zero = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    ctx.apply_hamiltonian_full(0, W, H, d[0], d[1], d[2], d[3], None, P, useqpe=zero)
    ctx.sync()
    lu, lv = G.component_eigenvalues(wf, H.cpu().numpy(), n)
    e = ctx.compute_energy(0, D, P)
    return dict(dens=D.cpu().numpy(), pots=P.cpu().numpy(), lu=lu, lv=lv, E=e[:7].sum(), Npart=e[7] + e[8])


@pytest.mark.parametrize("functional,sclgth", [(po.ASLDA, 0.0), (po.SLDAE, -1.1), (po.SLDA, 0.0)])
def test_boost_closed_forms_gpu(torch, functional, sclgth):
    from wslda_amd import td
    say("case: Galilean boost closed forms on the GPU, functional %d, 12^3" % functional)
    cfg, u, wf, fb = solution(functional, sclgth)
    N = NX ** 3
    n = wf.shape[0] // 2
    m = (1, -1, 1)
    phase, q = G.lattice_phase(NX, NX, NX, cfg.dx, m)
    kvec = G.state_momenta(wf, n, (NX, NX, NX), cfg.dx)
    ctx0 = td.TDContext(tdgauss.to_params(cfg), n, 0)
    ctx0.set_states(fb)
    rest = gpu_measure(torch, ctx0, wf, G.rest_pots(u, N))
    ctx0.close()
    wb, pb, dmu = G.boost(u, wf, G.rest_pots(u, N), phase, q, N)
    cb = po.Cfg.from_buffer_copy(cfg)
    cb.mu_a += dmu
    cb.mu_b += dmu
    ctx1 = td.TDContext(tdgauss.to_params(cb), n, 0)
    ctx1.set_states(fb)
    bst = gpu_measure(torch, ctx1, wb, pb)
    ctx1.close()
    err = G.check_closed_forms(rest, bst, u, q, phase, kvec, N)
    say("  %s" % {k: "%.1e" % v for k, v in err.items()})
    bad = {k: v for k, v in err.items() if not v < 1e-12}
    assert not bad, err


@pytest.mark.parametrize("fused", [False, True])
def test_boosted_aslda_gpu_steps_keep_flow(torch, fused):
    """Three ASLDA PECE steps of the boosted state from its exact history
    on the GPU (the modular step, and the folded step of tdw_step_fused):
    rho, tau, j, |nu| and the energy stay put to 1e-11."""
    from wslda_amd import td
    say("case: boosted ASLDA 12^3, 3 PECE steps from the exact history, fused=%d" % fused)
    cfg, u, wf, fb = solution(po.ASLDA)
    N = NX ** 3
    n = wf.shape[0] // 2
    phase, q = G.lattice_phase(NX, NX, NX, cfg.dx, (1, 0, 0))
    wb, pb, dmu = G.boost(u, wf, G.rest_pots(u, N), phase, q, N)
    cfg.mu_a += dmu
    cfg.mu_b += dmu
    ctx = td.TDContext(tdgauss.to_params(cfg), n, 0)
    ctx.set_states(fb)
    m = gpu_measure(torch, ctx, wb, pb)
    R = td.TDRun(ctx, wb, m["pots"], fused=fused)
    R.initial_measure()
    d0 = R.dens.cpu().numpy()
    e0 = R.energy(0)[:7].sum()
    U, V = wb[:n], wb[n:]
    wu, wv = np.sum(np.abs(U) ** 2, 1), np.sum(np.abs(V) ** 2, 1)
    e = (m["lu"] * wu + m["lv"] * wv) / (wu + wv)
    ou, ov = m["lu"] - e, m["lv"] - e
    hist = R.history()
    for j in range(5):
        t = -j * cfg.dt
        h = np.concatenate([ou[:, None] * U * np.exp(-1j * ou * t)[:, None],
                            ov[:, None] * V * np.exp(-1j * ov * t)[:, None]])
        hist[j].copy_(torch.as_tensor(h))
    for it in range(3):
        R.step(it, grad_last=True)
    ctx.sync()
    d1 = R.dens.cpu().numpy()
    for lo, hi in ((2, 4), (7, 9), (4, 7), (9, 12)):
        assert np.abs(d1[lo * N:hi * N] - d0[lo * N:hi * N]).max() < 1e-11 * np.abs(d0[lo * N:hi * N]).max(), (lo, hi)
    nu0 = np.abs(d0[0:2 * N:2] + 1j * d0[1:2 * N:2])
    nu1 = np.abs(d1[0:2 * N:2] + 1j * d1[1:2 * N:2])
    assert np.abs(nu1 - nu0).max() < 1e-11 * nu0.max()
    assert abs(R.energy(3)[:7].sum() - e0) < 1e-11 * abs(e0)
    ctx.close()
