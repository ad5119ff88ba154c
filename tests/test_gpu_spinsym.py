"""Spin-symmetric mode (SPINSYMMETRY_MODE) on the HIP path against the
oracle's spin-symmetric mode (itself pinned by the doubled-spectrum
identities of test_oracle_spinsym.py).

Bar: relative error <= 1e-12 for densities, the uniform solution, PECE steps
(FP64 modular and fused, ASLDA two-pass); the split-precision run within
1e-6 of FP64 in the densities.
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss
from tests.test_oracle_spinsym import random_states

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


@pytest.mark.parametrize("dim,shape", [(1, (32, 4, 4)), (2, (12, 16, 6)), (3, (8, 12, 6)), (3, (16, 16, 16))])
@pytest.mark.parametrize("grad", [True, False])
def test_densities(torch, dim, shape, grad):
    """calculate_densities (+ tau) over stored derivatives, and for the 16^3
    lattice also with the gradients formed inside the density passes."""
    from wslda_amd import td
    say("case: spin-symmetric densities dim=%d shape=%s grad=%d" % (dim, shape, grad))
    wf, fb, kt, cnt = random_states(dim, shape, 40, 3 + dim)
    n = wf.shape[0] // 2
    cfg = po.make_cfg(dim, *shape, dx=0.9, spinsymmetry=1)
    N = po.npts(cfg)
    P = po.Problem(cfg, wf, fb, kt=kt, cnt=cnt)
    L = po.lib()
    if grad:
        L.tdwo_derivatives(C.byref(cfg), 2 * n, po.ptr(P.wf), po.ptr(P.ddx), po.ptr(P.ddy), po.ptr(P.ddz),
                           po.ptr(P.lap))
    dref = np.zeros(12 * N)
    g = (lambda a: po.ptr(a) if grad else None)
    L.tdwo_densities(C.byref(cfg), n, po.ptr(P.wf), g(P.ddx), g(P.ddy), g(P.ddz), g(P.lap), po.ptr(P.fbeta),
                     None if kt is None else po.ptr(P.kt), None if cnt is None else po.ptr(P.cnt),
                     po.ptr(dref), int(grad))
    if grad:
        L.tdwo_density_tau(C.byref(cfg), po.ptr(dref))
    ctx = td.TDContext(tdgauss.to_params(cfg), n, 0)
    ctx.set_states(fb, kt, cnt)
    W = torch.as_tensor(wf).to("cuda:0")
    D = [torch.zeros_like(W) for _ in range(4)]
    if grad:
        ctx.compute_derivatives(W, D[0], D[1], D[2], D[3])
    d = torch.zeros(12 * N, dtype=torch.float64, device="cuda:0")
    ctx.calculate_densities(W, *(D if grad else [None] * 4), d, int(grad))
    if grad:
        ctx.density_caculate_tau(d)
    outs = [d]
    if grad and dim == 3 and shape == (16, 16, 16):
        d2 = torch.zeros_like(d)
        ctx.calculate_densities(W, None, None, None, None, d2, 1)   # in-pass gradients
        ctx.density_caculate_tau(d2)
        outs.append(d2)
    ctx.sync()
    for x in outs:
        a = x.cpu().numpy()
        for sl, name in [(slice(0, 2 * N), "nu"), (slice(2 * N, 7 * N), "spin a"), (slice(7 * N, 12 * N), "spin b")]:
            assert rel(a[sl], dref[sl]) < TOL, name
        assert np.array_equal(a[2 * N:7 * N], a[7 * N:12 * N])
    ctx.close()


def test_symmetrize_densities_device(torch):
    """The drop-in copies rho_a, tau_a, j_a over rho_b, tau_b, j_b
    (pca_dens.cu:343-359) and leaves nu alone."""
    from wslda_amd import td
    say("case: symmetrize_densities_device")
    cfg = po.make_cfg(3, 8, 12, 6, dx=0.9, spinsymmetry=1)
    N = po.npts(cfg)
    ctx = td.TDContext(tdgauss.to_params(cfg), 4, 0)
    h = np.random.default_rng(2).standard_normal(12 * N)
    d = torch.as_tensor(h).to("cuda:0")
    L = td.lib()
    L.tdw_bind(ctx.h)
    td.check(L.symmetrize_densities_device(C.c_void_p(d.data_ptr())), "symmetrize_densities_device")
    ctx.sync()
    L.tdw_bind(None)
    out = d.cpu().numpy()
    want = h.copy()
    want[7 * N:12 * N] = h[2 * N:7 * N]
    assert np.array_equal(out, want)
    ctx.close()


def test_uniform_solution_and_states(torch):
    """Library uniform solver and plane-wave states in spin-symmetric mode
    against the oracle's."""
    from wslda_amd import td
    say("case: spin-symmetric uniform solution")
    NX, DX = 10, 1.2
    LXYZ = (NX * DX) ** 3
    n0 = math.ceil(LXYZ / (6 * math.pi ** 2)) / LXYZ
    for functional in (po.SLDA, po.ASLDA, po.BDG):
        cfg = po.make_cfg(3, NX, NX, NX, dx=DX, functional=functional, spinsymmetry=1,
                          sclgth=-0.8 if functional == po.BDG else 0.0)
        mr = 0 if functional == po.ASLDA else 1
        u_ref, rc = po.uniform_solve(cfg, n0, n0, 0.999999 * math.pi / DX, 0.05, eps=1e-11, mu_rule=mr)
        assert rc == 0
        p = tdgauss.to_params(cfg)
        u, rc = td.uniform_solve(p, n0, n0, 0.999999 * math.pi / DX, 0.05, eps=1e-11, mu_rule=mr)
        assert rc == 0
        for k in ("n0_a", "n0_b", "mu_a", "mu_b", "delta", "nu", "tau_a", "tau_b", "S", "ekin", "epot", "epair"):
            x, y = getattr(u, k), getattr(u_ref, k)
            assert abs(x - y) <= TOL * max(abs(y), 1e-300), (functional, k, x, y)
        assert u.nwf == u_ref.nwf
        wf_ref, En_ref, _, _ = po.uniform_states(cfg, u_ref)
        wf, En, _, _ = td.uniform_states(p, u)
        assert rel(np.asarray(En), En_ref) < TOL
        assert rel(np.asarray(wf), wf_ref) < TOL


def ss_setup(dim, functional=po.SLDA, NX=12):
    """The td-gauss problem, made spin-symmetric: N_a = N_b, a V_ext that
    does not flip its sign for spin b, the E > 0 half of the spectrum."""
    g = tdgauss.golden(po.SLDA)["inputs"]
    DXv = g["DX"]
    LXYZ = (NX * DXv) ** 3
    n0 = math.ceil(LXYZ / (6.0 * math.pi ** 2)) / LXYZ
    cfg = po.make_cfg(dim, NX, NX, NX, dx=DXv, functional=functional, spinsymmetry=1)
    u, rc = po.uniform_solve(cfg, n0, n0, 0.999999 * math.pi / DXv, g["temperature"], eps=g["init0eps"],
                             scmix=g["init0scmix"], muchange=g["init0muchange"], maxiter=g["init0maxiter"],
                             mu_rule=tdgauss.mu_rule_of(functional))
    assert rc == 0
    n = u.n0_a + u.n0_b
    eF = (3 * math.pi ** 2 * n) ** (2 / 3) / 2
    kF = (3 * math.pi ** 2 * n) ** (1 / 3)
    wf, En, kt, cnt = po.uniform_states(cfg, u)
    fb = np.array([po.fbeta(E, u.beta) for E in En])
    cfg.vext = po.gauss_vext([(-0.5 * eF, (0, 0, 0), (-0.5 / (3.0 / kF) ** 2,) * 3)],
                             switch=(0.0, 50.0 / eF, 5.0 / eF), flip_b=0)
    cfg.mu_a, cfg.mu_b, cfg.ec = u.mu_a, u.mu_b, u.ec
    cfg.dt = g["dt_eF"] / eF
    N = po.npts(cfg)
    pots = np.zeros(12 * N)
    pots[:N] = u.V_a
    pots[N:2 * N] = u.V_b
    pots[2 * N:4 * N:2] = u.delta
    pots[4 * N:5 * N] = u.alph_a
    pots[5 * N:6 * N] = u.alph_b
    rng = np.random.default_rng(9)
    wf = wf * (1.0 + 1e-3 * rng.standard_normal(wf.shape))
    return dict(cfg=cfg, u=u, wf=wf, fbeta=fb, kt=kt, cnt=cnt, pots=pots, eF=eF)


def gpu_run(torch, s, fused=False, precision=0):
    from wslda_amd import td
    ctx = td.TDContext(tdgauss.to_params(s["cfg"], precision), s["wf"].shape[0] // 2, 0)
    ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    return ctx, td.TDRun(ctx, s["wf"], s["pots"], fused=fused)


@pytest.mark.parametrize("dim,functional,fused,nsteps", [(1, po.SLDA, False, 60), (2, po.SLDA, False, 12),
                                                        (3, po.SLDA, False, 4), (3, po.SLDA, True, 4),
                                                        (3, po.ASLDA, False, 4), (3, po.ASLDA, True, 4)])
def test_pece_steps(torch, dim, functional, fused, nsteps):
    say("case: spin-symmetric PECE steps dim=%d functional=%d fused=%d nsteps=%d" % (dim, functional, fused, nsteps))
    s = ss_setup(dim, functional, NX=16 if functional == po.ASLDA else 12)
    it0 = int(round(6.0 / s["eF"] / s["cfg"].dt))
    P = po.Problem(s["cfg"], s["wf"], s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    P.initial_measure()
    ctx, R = gpu_run(torch, s, fused=fused)
    R.initial_measure()
    for k in range(nsteps):
        last = (k == nsteps - 1)
        P.step(it0 + k, grad_last=last)
        R.step(it0 + k, grad_last=last)
    if fused:
        R.finalize()
    ctx.sync()
    N = ctx.N
    d = R.dens.cpu().numpy()
    for sl, name in [(slice(0, 2 * N), "nu"), (slice(2 * N, 3 * N), "rho_a"), (slice(3 * N, 4 * N), "tau_a"),
                     (slice(7 * N, 8 * N), "rho_b"), (slice(8 * N, 9 * N), "tau_b")]:
        assert rel(d[sl], P.dens[sl]) < TOL, name
    assert np.array_equal(d[2 * N:7 * N], d[7 * N:12 * N])
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    e_ref = P.energy(it0 + nsteps)
    e = R.energy(it0 + nsteps)
    assert abs(e[:7].sum() - e_ref[:7].sum()) <= TOL * abs(e_ref[:7].sum())
    assert abs(e[7] + e[8] - e_ref[7] - e_ref[8]) <= TOL * abs(e_ref[7] + e_ref[8])
    ctx.close()


def test_split_precision_steps(torch):
    """The split-precision fused step in spin-symmetric mode stays within
    1e-6 of the FP64 run in the densities."""
    from wslda_amd import td
    say("case: spin-symmetric split-precision steps (3D 12^3, 20 steps)")
    s = ss_setup(3)
    it0 = int(round(6.0 / s["eF"] / s["cfg"].dt))
    ctx, R = gpu_run(torch, s, fused=True)
    ctx2, R2 = gpu_run(torch, s, fused=True, precision=td.SPLIT)
    R.initial_measure()
    R2.initial_measure()
    for k in range(20):
        R.step(it0 + k, grad_last=(k == 19))
        R2.step(it0 + k, grad_last=(k == 19))
    ctx.sync()
    ctx2.sync()
    N = ctx.N
    a, b = R.dens.cpu().numpy(), R2.dens.cpu().numpy()
    for sl, name in [(slice(0, 2 * N), "nu"), (slice(2 * N, 3 * N), "rho_a"), (slice(7 * N, 8 * N), "rho_b")]:
        assert rel(b[sl], a[sl]) < 1e-6, name
    assert np.array_equal(b[2 * N:7 * N], b[7 * N:12 * N])
    ctx.close()
    ctx2.close()
