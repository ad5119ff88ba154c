"""HIP path (through the C ABI) against the CPU FP64 oracle.

Bar: relative error <= 1e-12 (max-abs difference over max-abs value) for
Laplacians, derivatives, densities, potentials, H application, energies and
particle number; FP64 everywhere.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.abs(a - b).max()
    s = np.abs(b).max()
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


def cuda(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a)).to("cuda:0")


def lattice_cfg(dim, shape, dx=1.0):
    nx, ny, nz = shape
    return po.make_cfg(dim, nx, ny, nz, dx=dx)


LATTICES = [(1, (24, 4, 4)), (1, (64, 1, 1)), (2, (12, 16, 6)), (2, (64, 64, 4)),
            (3, (8, 12, 6)), (3, (24, 24, 24)), (3, (16, 8, 12))]


@pytest.mark.parametrize("dim,shape", LATTICES)
def test_laplace_and_derivatives(torch, dim, shape):
    from wslda_amd import td
    say("case: laplace/derivatives dim=%d shape=%s" % (dim, shape))
    cfg = lattice_cfg(dim, shape, dx=1.1)
    N = po.npts(cfg)
    m = 6
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((m, N)) + 1j * rng.standard_normal((m, N))).astype(np.complex128)
    lap_ref = np.zeros_like(x)
    po.lib().tdwo_laplace(C.byref(cfg), m, po.ptr(x), po.ptr(lap_ref))
    dref = [np.zeros_like(x) for _ in range(4)]
    po.lib().tdwo_derivatives(C.byref(cfg), m, po.ptr(x), po.ptr(dref[0]),
                              po.ptr(dref[1]) if dim >= 2 else None,
                              po.ptr(dref[2]) if dim >= 3 else None, po.ptr(dref[3]))
    ctx = td.TDContext(tdgauss.to_params(cfg), 3, 0)
    X = cuda(torch, x)
    L = torch.zeros_like(X)
    ctx.compute_laplace(X, L)
    ctx.sync()
    assert rel(L.cpu().numpy(), lap_ref) < TOL
    # in place
    X2 = X.clone()
    ctx.compute_laplace(X2, X2)
    ctx.sync()
    assert rel(X2.cpu().numpy(), lap_ref) < TOL
    D = [torch.zeros_like(X) for _ in range(4)]
    ctx.compute_derivatives(X, D[0], D[1] if dim >= 2 else None, D[2] if dim >= 3 else None, D[3])
    ctx.sync()
    for k in range(4):
        if k == 1 and dim < 2 or k == 2 and dim < 3:
            continue
        assert rel(D[k].cpu().numpy(), dref[k]) < TOL, k
    ctx.close()


def gpu_run(torch, s, fused=False, precision=0):
    from wslda_amd import td
    ctx = td.TDContext(tdgauss.to_params(s["cfg"], precision), s["wf"].shape[0] // 2, 0)
    ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    R = td.TDRun(ctx, s["wf"], s["pots"], fused=fused)
    return ctx, R


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_building_blocks(torch, dim):
    """Each reference entry point on the td-gauss state, perturbed."""
    say("case: building blocks dim=%d" % dim)
    s = tdgauss.setup(dim)
    cfg = s["cfg"]
    rng = np.random.default_rng(5 + dim)
    wf = s["wf"] * (1.0 + 0.01 * rng.standard_normal(s["wf"].shape)) + 1e-3 * (
        rng.standard_normal(s["wf"].shape) + 1j * rng.standard_normal(s["wf"].shape))
    n, N = wf.shape[0] // 2, wf.shape[1]
    P = po.Problem(cfg, wf.copy(), s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    ctx, R = gpu_run(torch, dict(s, wf=wf))
    it = 1200   # V_ext switched on (t*eF = 4.2..)
    # normalisation
    po.lib().tdwo_normalize(C.byref(cfg), n, po.ptr(P.wf))
    ctx.normalize_wf(R.wf)
    ctx.sync()
    assert rel(R.wf.cpu().numpy(), P.wf) < TOL
    # derivatives + densities (with gradients) + tau
    P.initial_measure()
    R.initial_measure()
    ctx.sync()
    assert rel(R.lap.cpu().numpy(), P.lap) < TOL
    d_gpu = R.dens.cpu().numpy()
    for k, name in [(slice(0, 2 * N), "nu"), (slice(2 * N, 3 * N), "rho_a"), (slice(3 * N, 4 * N), "tau_a"),
                    (slice(4 * N, 5 * N), "j_a_x"), (slice(7 * N, 8 * N), "rho_b"), (slice(8 * N, 9 * N), "tau_b"),
                    (slice(9 * N, 10 * N), "j_b_x")]:
        assert rel(d_gpu[k], P.dens[k]) < TOL or np.abs(P.dens[k]).max() < 1e-14, name
    # densities without gradients
    dl_ref = np.zeros(12 * N)
    po.lib().tdwo_densities(C.byref(cfg), n, po.ptr(P.wf), None, None, None, None, po.ptr(P.fbeta),
                            po.ptr(P.kt), po.ptr(P.cnt), po.ptr(dl_ref), 0)
    dl = torch.zeros(12 * N, dtype=torch.float64, device="cuda:0")
    ctx.calculate_densities(R.wf, None, None, None, None, dl, 0)
    ctx.sync()
    assert rel(dl.cpu().numpy(), dl_ref) < TOL
    # potentials (SLDA, 15-iteration local loop) from the perturbed densities
    po.lib().tdwo_potentials(C.byref(cfg), it, po.ptr(P.dens), po.ptr(P.pots))
    ctx.compute_potentials(it, R.dens, R.pots)
    ctx.sync()
    assert rel(R.pots.cpu().numpy(), P.pots) < TOL
    # H application and <H>
    out_ref = np.zeros_like(P.wf)
    q_ref = np.zeros(n)
    po.lib().tdwo_apply_h(C.byref(cfg), n, po.ptr(P.wf), po.ptr(out_ref), po.ptr(P.lap), po.ptr(P.pots),
                          po.ptr(P.kt), None, po.ptr(q_ref))
    out = torch.zeros_like(R.wf)
    q = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    ctx.apply_hamiltonian(it, R.wf, out, R.lap, R.pots, qpe=q)
    ctx.sync()
    assert rel(q.cpu().numpy(), q_ref) < TOL
    assert rel(out.cpu().numpy(), out_ref) < TOL
    # energy
    e_ref = P.energy(it)
    e = R.energy(it)
    for k in (0, 1, 2, 4, 7, 8):
        assert abs(e[k] - e_ref[k]) <= TOL * max(abs(e_ref[k]), abs(e_ref[0])), k
    # ABM predictor / corrector on random histories
    hist = [(rng.standard_normal(wf.shape) + 1j * rng.standard_normal(wf.shape)) * 0.1 for _ in range(5)]
    for scheme in (po.AB3AM4, po.AB4AM5, po.AB5AM5):
        cfg.scheme = scheme
        y = P.wf.copy()
        hf = [h.copy() for h in hist]
        po.lib().tdwo_abm_step1(C.byref(cfg), n, po.ptr(y), *[po.ptr(h) for h in hf])
        ctx.set_params(tdgauss.to_params(cfg))
        Y = R.wf.clone()
        HF = [cuda(torch, h) for h in hist]
        ctx.abm_step1(Y, HF)
        ctx.sync()
        assert rel(Y.cpu().numpy(), y) < TOL
        for h, H in zip(hf, HF):
            assert rel(H.cpu().numpy(), h) < TOL
        y4 = np.zeros_like(y)
        po.lib().tdwo_abm_step4(C.byref(cfg), n, po.ptr(out_ref), po.ptr(y4), po.ptr(hf[0]))
        Y4 = torch.zeros_like(Y)
        ctx.abm_step4(out, Y4, HF[0])
        ctx.sync()
        assert rel(Y4.cpu().numpy(), y4) < TOL
    ctx.close()


@pytest.mark.parametrize("dim,nsteps", [(1, 286), (2, 40), (3, 4)])
def test_pece_steps(torch, dim, nsteps):
    """Full PECE steps from the td-gauss uniform state with V_ext on."""
    say("case: PECE steps dim=%d nsteps=%d" % (dim, nsteps))
    s = tdgauss.setup(dim)
    it0 = int(round(6.0 / s["eF"] / s["cfg"].dt))    # start inside the switch-on window
    P = po.Problem(s["cfg"], s["wf"], s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    P.initial_measure()
    ctx, R = gpu_run(torch, s)
    R.initial_measure()
    for k in range(nsteps):
        last = (k == nsteps - 1)
        P.step(it0 + k, grad_last=last)
        R.step(it0 + k, grad_last=last)
    ctx.sync()
    N = ctx.N
    d = R.dens.cpu().numpy()
    for sl, name in [(slice(0, 2 * N), "nu"), (slice(2 * N, 3 * N), "rho_a"), (slice(3 * N, 4 * N), "tau_a"),
                     (slice(7 * N, 8 * N), "rho_b"), (slice(8 * N, 9 * N), "tau_b")]:
        assert rel(d[sl], P.dens[sl]) < TOL, name
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    e_ref = P.energy(it0 + nsteps)
    e = R.energy(it0 + nsteps)
    etot_ref, etot = e_ref[:7].sum(), e[:7].sum()
    assert abs(etot - etot_ref) <= TOL * abs(etot_ref)
    assert abs(e[7] + e[8] - e_ref[7] - e_ref[8]) <= TOL * abs(e_ref[7] + e_ref[8])
    ctx.close()


@pytest.mark.parametrize("functional", [po.SLDA, po.BDG])
def test_td_gauss_1d_golden_on_gpu(torch, functional):
    """The whole 28,600-step 1D trajectory on the GPU against the
    reference's golden rows (10 printed digits), SLDA and BdG."""
    say("case: td-gauss 1D golden trajectory on GPU, functional %d" % functional)
    s = tdgauss.setup(1, functional=functional)
    ctx, R = gpu_run(torch, s)
    R.initial_measure()
    ref = tdgauss.golden(functional)["runs"]["ss0-1d"]["rows"]
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
    ctx.close()


def test_dropin_call_sequence(torch):
    """The reference-named entry points, called in pca.c's order on the
    bound context, give the same step as tdw_step."""
    from wslda_amd import td
    say("case: drop-in call sequence (3D td-gauss)")
    s = tdgauss.setup(3)
    ctx, R = gpu_run(torch, s)
    ctx2, R2 = gpu_run(torch, s)
    L = td.lib()
    n = R.n
    N = ctx.N
    R.initial_measure()
    R2.initial_measure()
    for it in range(3):
        R.step(it, grad_last=(it == 2))
    # drop-in sequence on R2's arrays
    L.tdw_bind(ctx2.h)
    work = torch.zeros(max(n, 16), dtype=torch.float64, device="cuda:0")
    ws = C.c_size_t()
    td.check(L.create_cufftPlans(100000, n, C.byref(ws)), "create_cufftPlans")
    td.check(L.set_workspace_for_cufftPlan(C.c_void_p(work.data_ptr())), "set_workspace")
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    f = list(R2.f)
    fbeta_t = torch.as_tensor(s["fbeta"]).to("cuda:0")
    torch.cuda.synchronize()
    for it in range(3):
        td.check(L.abm_step1(n, p(R2.wf), p(f[0]), p(f[1]), p(f[2]), p(f[3]), p(f[4]), 45, 512), "abm_step1")
        td.check(L.normalize_wf(n, p(R2.wf), 512), "normalize_wf")
        td.check(L.compute_laplace(2 * n, p(R2.wf), p(R2.lap), 512), "compute_laplace")
        td.check(L.calculate_densities(n, p(R2.wf), p(R2.ddx), p(R2.ddy), p(R2.ddz), p(R2.lap), p(fbeta_t),
                                       None, p(R2.dens), 0, 512), "calculate_densities")
        td.check(L.compute_potentials(it + 1, p(R2.dens), p(R2.pots), 1.0, 512), "compute_potentials")
        td.check(L.apply_hamiltonian(it + 1, n, p(R2.wf), p(R2.lap), p(R2.ddx), p(R2.ddy), p(R2.ddz), p(R2.lap),
                                     None, p(R2.dens), p(R2.pots), 0.0, None, 0.0, 512), "apply_hamiltonian")
        td.check(L.abm_step4(n, p(R2.lap), p(R2.wf), p(f[3]), 45, 512), "abm_step4")
        td.check(L.normalize_wf(n, p(R2.wf), 512), "normalize_wf")
        grad = it == 2
        if grad:
            td.check(L.compute_derivatives(2 * n, p(R2.wf), p(R2.ddx), p(R2.ddy), p(R2.ddz), p(R2.lap), 512), "deriv")
        else:
            td.check(L.compute_laplace(2 * n, p(R2.wf), p(R2.lap), 512), "compute_laplace")
        td.check(L.calculate_densities(n, p(R2.wf), p(R2.ddx), p(R2.ddy), p(R2.ddz), p(R2.lap), p(fbeta_t),
                                       None, p(R2.dens), int(grad), 512), "calculate_densities")
        if grad:
            td.check(L.density_caculate_tau(p(R2.dens), 512), "tau")
        td.check(L.compute_potentials(it + 1, p(R2.dens), p(R2.pots), 1.0, 512), "compute_potentials")
        f = [f[3], f[0], f[1], f[2], f[4]]
        td.check(L.apply_hamiltonian(it + 1, n, p(R2.wf), p(f[0]), p(R2.ddx), p(R2.ddy), p(R2.ddz), p(R2.lap),
                                     None, p(R2.dens), p(R2.pots), 0.0, None, 0.0, 512), "apply_hamiltonian")
    td.check(L.compute_energy(3, p(R2.dens), p(R2.pots), p(work), 512), "compute_energy")
    torch.cuda.synchronize()
    ctx.sync()
    assert rel(R2.dens.cpu().numpy(), R.dens.cpu().numpy()) < 1e-14
    assert rel(R2.wf.cpu().numpy(), R.wf.cpu().numpy()) < 1e-14
    e = R.energy(3)
    assert np.allclose(work[:11].cpu().numpy(), e, rtol=1e-13, atol=0)
    L.tdw_bind(None)
    ctx.close()
    ctx2.close()
