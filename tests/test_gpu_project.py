"""Project builds on the GPU: a problem's predefines.h and problem-definition.h
compiled into the library (tests/projects/*, built by __graft_entry__.build)
drive the engine as the reference's callbacks drive its kernels
(pca_kernels.cu:28-42, 90-92, 137-185, 274-275; tdwslda_functionals.h:72-73,
397; tdwslda_energy_generic.h:99-101; pca.c:1443-1447, 1515-1531).

- td_gauss_1d: the td-gauss-slda testsuite problem written against the
  callback interface; its V_ext, the run it drives and the whole golden
  trajectory equal the built-in problem's and the oracle's.
- hooks_3d: every callback switched on (V_ext, Delta_ext, a two-valued
  scattering length, modify_potentials / _energies / _densities) on the
  td-gauss-bdg lattice; potentials, H and energies against the oracle with
  the same modifications made by hand, and the whole step against the
  drop-in call sequence with modify_densities made by the test."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss
from tests.test_gpu_parity import TOL, rel, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_1D = os.path.join(ROOT, "tests", "projects", "td_gauss_1d", "libwslda_td.so")
LIB_3D = os.path.join(ROOT, "tests", "projects", "hooks_3d", "libwslda_td.so")
# hooks_3d parameters (see its problem-definition.h)
HOOKS = [0.05, -0.02, 0.3, 0.01, -0.02, 0.25, 0.04, 0.7, 1.0e-3]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def nptr(a):
    return a.ctypes.data_as(C.c_void_p)


def td_gauss_1d_project(torch):
    """The td_gauss_1d project context and the built-in one on the same
    start; the project's parameters come from its own process_params."""
    from wslda_amd import td
    s = tdgauss.setup(1)
    cfg = s["cfg"]
    L = td.lib(LIB_1D)
    prm = np.zeros(td.MAX_USER_PARAMS)
    prm[:8] = tdgauss.golden()["runs"]["ss0-1d"]["params"]
    kF = C.c_double(s["kF"])
    mu = np.array([cfg.mu_a, cfg.mu_b])
    L.process_params.restype = None
    L.process_params(nptr(prm), C.byref(kF), nptr(mu), C.c_size_t(0), None)
    pp = tdgauss.to_params(cfg)
    pp.vext = td.Vext()   # a project ignores the built-in V_ext; left empty here
    n = s["wf"].shape[0] // 2
    ctxp = td.TDContext(pp, n, 0, libpath=LIB_1D)
    td.check(L.tdw_set_user_params(ctxp.h, nptr(prm)), "tdw_set_user_params")
    ctxb = td.TDContext(tdgauss.to_params(cfg), n, 0)
    for ctx in (ctxp, ctxb):
        ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    return s, ctxp, ctxb


def test_td_gauss_1d_project_vext_and_steps(torch):
    """The problem's v_ext equals the built-in Gaussian with its smooth switch
    at times before, during and after the switch-on; 286 PECE steps from
    inside the switch-on window agree with the built-in run and the oracle."""
    from wslda_amd import td
    say("case: project td_gauss_1d: V_ext and 286 steps against built-in and oracle")
    s, ctxp, ctxb = td_gauss_1d_project(torch)
    N = ctxp.N
    eF, dt = s["eF"], s["cfg"].dt
    it0 = int(round(6.0 / eF / dt))
    for it in (0, it0, int(round(12.0 / eF / dt)), int(round(30.0 / eF / dt)), int(round(54.0 / eF / dt))):
        v = [torch.zeros(2 * N, dtype=torch.float64, device="cuda:0") for _ in range(2)]
        for ctx, vv in ((ctxp, v[0]), (ctxb, v[1])):
            td.check(ctx.L.tdw_vext_eval(ctx.h, it, ptr(vv), C.c_void_p(vv.data_ptr() + 8 * N)), "tdw_vext_eval")
            ctx.sync()
        a, b = v[0].cpu().numpy(), v[1].cpu().numpy()
        if np.abs(b).max() == 0.0:
            assert np.abs(a).max() == 0.0, it
        else:
            assert rel(a, b) < 1e-13, (it, rel(a, b))
    P = po.Problem(s["cfg"], s["wf"], s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    Rp = td.TDRun(ctxp, s["wf"], s["pots"])
    Rb = td.TDRun(ctxb, s["wf"], s["pots"])
    for R in (Rp, Rb):
        R.initial_measure()
    P.initial_measure()
    nsteps = 286
    for k in range(nsteps):
        last = k == nsteps - 1
        for R in (Rp, Rb):
            R.step(it0 + k, grad_last=last)
        P.step(it0 + k, grad_last=last)
    ctxp.sync()
    ctxb.sync()
    dp, db = Rp.dens.cpu().numpy(), Rb.dens.cpu().numpy()
    assert rel(dp, db) < TOL
    for sl in (slice(0, 4 * N), slice(7 * N, 9 * N)):
        assert rel(dp[sl], P.dens[sl]) < TOL
    ep, eb, eo = Rp.energy(it0 + nsteps), Rb.energy(it0 + nsteps), P.energy(it0 + nsteps)
    assert abs(ep[:7].sum() - eb[:7].sum()) <= TOL * abs(eb[:7].sum())
    assert abs(ep[:7].sum() - eo[:7].sum()) <= TOL * abs(eo[:7].sum())
    assert abs(ep[td.EPOTEXT] - eo[td.EPOTEXT]) <= TOL * abs(eo[:7]).max()
    ctxp.close()
    ctxb.close()


def test_td_gauss_1d_project_golden(torch):
    """The testsuite problem compiled as a project reproduces the reference's
    whole 28,600-step golden log (10 printed digits), as the built-in problem
    does."""
    say("case: project td_gauss_1d: whole golden trajectory")
    from wslda_amd import td
    s, ctxp, ctxb = td_gauss_1d_project(torch)
    ctxb.close()
    R = td.TDRun(ctxp, s["wf"], s["pots"])
    R.initial_measure()
    ref = tdgauss.golden()["runs"]["ss0-1d"]["rows"]
    it = 0
    bad = []
    for m in range(101):
        if m > 0:
            for k in range(s["timesteps"]):
                R.step(it, grad_last=(k == s["timesteps"] - 1))
                it += 1
        row = tdgauss.wlog_row(R.energy(it), s["Effg"])
        for col, val in ((2, row[0]), (3, row[1]), (4, row[2]), (5, row[3]), (6, row[4]), (7, row[5]),
                         (8, row[6])):
            if not tdgauss.digits_match(val, ref[m][col]):
                bad.append((m, col, val, ref[m][col]))
        if abs(row[8] - ref[m][10]) > 2e-11:
            bad.append((m, 10, row[8], ref[m][10]))
    assert not bad, bad[:10]
    ctxp.close()


def hooks_setup(torch):
    from wslda_amd import td
    s = tdgauss.setup(3, functional=po.BDG)
    prm = np.zeros(32)
    prm[:len(HOOKS)] = HOOKS
    n = s["wf"].shape[0] // 2
    ctx = td.TDContext(tdgauss.to_params(s["cfg"]), n, 0, libpath=LIB_3D)
    td.check(ctx.L.tdw_set_user_params(ctx.h, nptr(prm)), "tdw_set_user_params")
    ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    return s, prm, ctx


def oracle_cfg(cfg, sclgth=None):
    """The oracle's configuration with the hooks problem's V_ext (a Gaussian
    in x, opposite sign for spin b) and, optionally, another scattering
    length."""
    c = po.Cfg.from_buffer_copy(cfg)
    c.vext = po.gauss_vext([(HOOKS[0], (0.0, 0.0, 0.0), (HOOKS[1], 0.0, 0.0))], flip_b=1)
    if sclgth is not None:
        c.sclgth = sclgth
    return c


def coords(cfg):
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return (ix.ravel(), (cfg.dx * (ix - nx // 2)).ravel(), (cfg.dy * (iy - ny // 2)).ravel(),
            (cfg.dz * (iz - nz // 2)).ravel())


def delta_ext(cfg, delta):
    _, _, y, _ = coords(cfg)
    ph = 2.0 * math.pi * y / (cfg.ny * cfg.dy)
    return HOOKS[2] * delta + (HOOKS[3] + 1j * HOOKS[4]) * np.exp(1j * ph)


def oracle_potentials(cfg, it, dens, pots_prev):
    """tdwo_potentials with the problem's V_ext and the scattering length of
    each half of the lattice, then modify_potentials by hand."""
    N = po.npts(cfg)
    ix, _, _, z = coords(cfg)
    out = []
    for a in (cfg.sclgth, cfg.sclgth * (1.0 + HOOKS[5])):
        c = oracle_cfg(cfg, a)
        p = pots_prev.copy()
        po.lib().tdwo_potentials(C.byref(c), it, po.ptr(dens), po.ptr(p))
        out.append(p)
    upper = ix >= cfg.nx // 2
    pots = out[0].copy()
    for blk in range(12):
        sl = slice(blk * N, (blk + 1) * N)
        pots[sl] = np.where(upper, out[1][sl], out[0][sl])
    # Delta is stored interleaved (re, im) in blocks 2, 3
    d0, d1 = out[0][2 * N:4 * N].reshape(N, 2), out[1][2 * N:4 * N].reshape(N, 2)
    pots[2 * N:4 * N] = np.where(upper[:, None], d1, d0).ravel()
    pots[N:2 * N] += HOOKS[6] * np.cos(2.0 * math.pi * z / (cfg.nz * cfg.dz))
    return pots


def test_hooks_potentials_h_energy(torch):
    """Potentials (V_ext, per-point scattering length, modify_potentials), H
    (Delta + Delta_ext inside the application only) and the energies
    (Delta_ext energy, modify_energies) against the oracle."""
    from wslda_amd import td
    say("case: project hooks_3d: potentials, H, energies against the oracle")
    s, prm, ctx = hooks_setup(torch)
    cfg = s["cfg"]
    N = ctx.N
    n = ctx.nwf
    R = td.TDRun(ctx, s["wf"], s["pots"])
    R.initial_measure()
    ctx.sync()
    dens = R.dens.cpu().numpy()
    it = 17
    pots_prev = R.pots.cpu().numpy().copy()
    ctx.compute_potentials(it, R.dens, R.pots)
    ctx.sync()
    pots = R.pots.cpu().numpy()
    ref = oracle_potentials(cfg, it, dens, pots_prev)
    assert rel(pots[:4 * N], ref[:4 * N]) < TOL
    # H with Delta + Delta_ext; the stored potentials are left as they are
    wf = R.wf
    lap = torch.zeros_like(wf)
    out = torch.zeros_like(wf)
    qpe = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    ctx.compute_laplace(wf, lap)
    ctx.apply_hamiltonian(it, wf, out, lap, R.pots, None, qpe)
    ctx.sync()
    assert np.array_equal(R.pots.cpu().numpy(), pots)
    d = pots[2 * N:4 * N:2] + 1j * pots[2 * N + 1:4 * N:2]
    dh = d + delta_ext(cfg, d)
    pots_h = pots.copy()
    pots_h[2 * N:4 * N:2], pots_h[2 * N + 1:4 * N:2] = dh.real, dh.imag
    wfh = wf.cpu().numpy()
    lap_o = np.zeros_like(wfh)
    out_o = np.zeros_like(wfh)
    qpe_o = np.zeros(n)
    L = po.lib()
    L.tdwo_laplace(C.byref(cfg), 2 * n, po.ptr(wfh), po.ptr(lap_o))
    L.tdwo_apply_h(C.byref(cfg), n, po.ptr(wfh), po.ptr(out_o), po.ptr(lap_o), po.ptr(pots_h), None, None,
                   po.ptr(qpe_o))
    assert rel(out.cpu().numpy(), out_o) < TOL
    assert rel(qpe.cpu().numpy(), qpe_o) < TOL
    # energies
    e = ctx.compute_energy(it, R.dens, R.pots)
    eo = np.zeros(td.NITEMS)
    c = oracle_cfg(cfg)
    L.tdwo_energy(C.byref(c), it, po.ptr(dens), po.ptr(pots), po.ptr(eo))
    dV = cfg.dx * cfg.dy * cfg.dz
    nu = dens[0:2 * N:2] + 1j * dens[1:2 * N:2]
    eo[td.EPAIREXT] = (np.conj(nu) * delta_ext(cfg, d)).real.sum() * (-2.0) * dV
    eo[td.EPOT] += HOOKS[7] * dens[2 * N:3 * N].sum() * dV
    say("  energies %s" % e[:7])
    assert abs(e[td.EPAIREXT]) > 1e-6 * abs(e[:7]).max()
    for k in range(td.NITEMS):
        assert abs(e[k] - eo[k]) <= TOL * max(abs(eo[:7]).max(), abs(eo[k])), (k, e[k], eo[k])
    ctx.close()


def test_hooks_step_matches_dropin_sequence(torch):
    """tdw_step of the project equals the reference's drop-in call sequence
    (pca.c:1401-1570) with modify_densities done by the test between the
    densities and the potentials, as pca.c calls it; the problem's host hook
    ran twice per step with the step's iteration number."""
    from wslda_amd import td
    say("case: project hooks_3d: tdw_step against the drop-in sequence")
    s, prm, ctx = hooks_setup(torch)
    _, _, ctx2 = hooks_setup(torch)
    L = ctx.L
    calls = C.c_int.in_dll(L, "hooks_calls")
    last_it = C.c_int.in_dll(L, "hooks_last_it")
    c0 = calls.value
    R = td.TDRun(ctx, s["wf"], s["pots"])
    R2 = td.TDRun(ctx2, s["wf"], s["pots"])
    n, N = ctx.nwf, ctx.N
    R.initial_measure()
    R2.initial_measure()
    for it in range(3):
        R.step(it, grad_last=(it == 2))
    ctx.sync()
    assert calls.value - c0 == 6 and last_it.value == 3
    td.check(L.tdw_bind(ctx2.h), "tdw_bind")
    work = torch.zeros(max(n, 16), dtype=torch.float64, device="cuda:0")
    ws = C.c_size_t()
    td.check(L.create_cufftPlans(100000, n, C.byref(ws)), "create_cufftPlans")
    td.check(L.set_workspace_for_cufftPlan(ptr(work)), "set_workspace")
    f = list(R2.f)
    fbeta_t = torch.as_tensor(s["fbeta"]).to("cuda:0")

    def modify_densities():
        ctx2.sync()
        R2.dens[7 * N:8 * N] *= 1.0 + HOOKS[8]
        torch.cuda.synchronize()

    torch.cuda.synchronize()
    for it in range(3):
        td.check(L.abm_step1(n, ptr(R2.wf), ptr(f[0]), ptr(f[1]), ptr(f[2]), ptr(f[3]), ptr(f[4]), 45, 512), "abm1")
        td.check(L.normalize_wf(n, ptr(R2.wf), 512), "normalize_wf")
        td.check(L.compute_laplace(2 * n, ptr(R2.wf), ptr(R2.lap), 512), "compute_laplace")
        td.check(L.calculate_densities(n, ptr(R2.wf), ptr(R2.ddx), ptr(R2.ddy), ptr(R2.ddz), ptr(R2.lap),
                                       ptr(fbeta_t), None, ptr(R2.dens), 0, 512), "calculate_densities")
        modify_densities()
        td.check(L.compute_potentials(it + 1, ptr(R2.dens), ptr(R2.pots), 1.0, 512), "compute_potentials")
        td.check(L.apply_hamiltonian(it + 1, n, ptr(R2.wf), ptr(R2.lap), ptr(R2.ddx), ptr(R2.ddy), ptr(R2.ddz),
                                     ptr(R2.lap), None, ptr(R2.dens), ptr(R2.pots), 0.0, None, 0.0, 512), "H")
        td.check(L.abm_step4(n, ptr(R2.lap), ptr(R2.wf), ptr(f[3]), 45, 512), "abm4")
        td.check(L.normalize_wf(n, ptr(R2.wf), 512), "normalize_wf")
        grad = it == 2
        if grad:
            td.check(L.compute_derivatives(2 * n, ptr(R2.wf), ptr(R2.ddx), ptr(R2.ddy), ptr(R2.ddz), ptr(R2.lap),
                                           512), "compute_derivatives")
        else:
            td.check(L.compute_laplace(2 * n, ptr(R2.wf), ptr(R2.lap), 512), "compute_laplace")
        td.check(L.calculate_densities(n, ptr(R2.wf), ptr(R2.ddx), ptr(R2.ddy), ptr(R2.ddz), ptr(R2.lap),
                                       ptr(fbeta_t), None, ptr(R2.dens), int(grad), 512), "calculate_densities")
        if grad:
            td.check(L.density_caculate_tau(ptr(R2.dens), 512), "tau")
        modify_densities()
        td.check(L.compute_potentials(it + 1, ptr(R2.dens), ptr(R2.pots), 1.0, 512), "compute_potentials")
        f = [f[3], f[0], f[1], f[2], f[4]]
        td.check(L.apply_hamiltonian(it + 1, n, ptr(R2.wf), ptr(f[0]), ptr(R2.ddx), ptr(R2.ddy), ptr(R2.ddz),
                                     ptr(R2.lap), None, ptr(R2.dens), ptr(R2.pots), 0.0, None, 0.0, 512), "H")
    td.check(L.compute_energy(3, ptr(R2.dens), ptr(R2.pots), ptr(work), 512), "compute_energy")
    torch.cuda.synchronize()
    ctx2.sync()
    assert calls.value - c0 == 6   # the drop-in calls leave the host hook to the driver
    assert rel(R2.dens.cpu().numpy(), R.dens.cpu().numpy()) < 1e-14
    assert rel(R2.wf.cpu().numpy(), R.wf.cpu().numpy()) < 1e-14
    e = R.energy(3)
    assert np.allclose(work[:11].cpu().numpy(), e, rtol=1e-13, atol=0)
    L.tdw_bind(None)
    ctx.close()
    ctx2.close()
