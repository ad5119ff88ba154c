"""Grid-field drop-ins (real-field gradient / derivatives / divergence,
Fermi-Dirac filters, overlaps, sums, workspace views) on the bound context
against the oracle (pinned by tests/test_oracle_grid.py).  Bar: 1e-12
relative to the field scale."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss
from tests.test_oracle_grid import LATT

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


@pytest.mark.parametrize("dim,shape,sp", LATT)
def test_real_field_derivatives_and_filters(torch, dim, shape, sp):
    from wslda_amd import td
    say("case: grid drop-ins dim=%d shape=%s" % (dim, shape))
    nx, ny, nz = shape
    cfg = po.make_cfg(dim, nx, ny, nz, dx=sp[0], dy=sp[1], dz=sp[2])
    N = po.npts(cfg)
    rng = np.random.default_rng(17 + dim)
    f = [np.ascontiguousarray(rng.standard_normal(N)) for _ in range(3)]
    O = po.lib()
    gref = [np.zeros(N) for _ in range(3)]
    O.tdwo_gradient_real(C.byref(cfg), po.ptr(f[0]), po.ptr(gref[0]), po.ptr(gref[1]), po.ptr(gref[2]))
    dref = [np.zeros(N) for _ in range(3)]
    O.tdwo_dvec_real(C.byref(cfg), po.ptr(f[0]), po.ptr(f[1]), po.ptr(f[2]), po.ptr(dref[0]), po.ptr(dref[1]),
                     po.ptr(dref[2]))
    divref = np.zeros(N)
    O.tdwo_div_real(C.byref(cfg), po.ptr(f[0]), po.ptr(f[1]), po.ptr(f[2]), po.ptr(divref))
    mu, T = 1.7, 0.4
    fl = np.ascontiguousarray(np.stack([f[0], f[1]]))
    flref = np.zeros(2 * N)
    O.tdwo_hf_filter_d(C.byref(cfg), 2, po.ptr(fl), po.ptr(flref), mu, T)
    z = np.ascontiguousarray(rng.standard_normal(N) + 1j * rng.standard_normal(N))
    zref = np.zeros_like(z)
    O.tdwo_hf_filter_c(C.byref(cfg), po.ptr(z), po.ptr(zref), mu, T)

    ctx = td.TDContext(tdgauss.to_params(cfg), 4, 0)
    L = td.lib()
    L.tdw_bind(ctx.h)
    try:
        dev = lambda a: torch.as_tensor(a).to("cuda:0")
        p = lambda t: C.c_void_p(t.data_ptr())
        F = [dev(x) for x in f]
        G = [torch.zeros(N, dtype=torch.float64, device="cuda:0") for _ in range(3)]
        td.check(L.compute_gradient_real_f(p(F[0]), p(G[0]), p(G[1]), p(G[2]), 256), "gradient")
        D = [x.clone() for x in F]
        # in place: outputs alias the inputs
        td.check(L.compute_derivative_real_vector_f(p(D[0]), p(D[1]), p(D[2]), p(D[0]), p(D[1]), p(D[2]), 256),
                 "dvec")
        V = [x.clone() for x in F]
        td.check(L.compute_divergence_real_vector_f(p(V[0]), p(V[1]), p(V[2]), p(V[0]), 256), "div")
        FL = dev(fl)
        td.check(L.high_frequency_filter_massive_d(2, p(FL), p(FL), mu, T, 256), "filter massive")
        F1 = F[0].clone()
        td.check(L.high_frequency_filter_d(p(F1), p(F1), mu, T, 256), "filter d")
        Z = dev(z)
        ZO = torch.zeros_like(Z)
        td.check(L.high_frequency_filter_c(p(Z), p(ZO), mu, T, 256), "filter c")
        ctx.sync()
    finally:
        L.tdw_bind(None)
    for d in range(dim):
        assert rel(G[d].cpu().numpy(), gref[d]) < TOL, ("gradient", d)
        assert rel(D[d].cpu().numpy(), dref[d]) < TOL, ("dvec", d)
    for d in range(dim, 3):   # axes beyond the lattice's are left alone
        assert np.array_equal(D[d].cpu().numpy(), f[d])
    assert rel(V[0].cpu().numpy(), divref) < TOL
    assert rel(FL.cpu().numpy().ravel(), flref) < TOL
    assert rel(F1.cpu().numpy(), flref[:N]) < TOL
    assert rel(ZO.cpu().numpy(), zref) < TOL
    ctx.close()


def test_overlap_sum_workspace(torch):
    from wslda_amd import td
    say("case: compute_ovelap, local_reductionR, mm_get_*")
    cfg = po.make_cfg(3, 8, 12, 6, dx=0.9, dy=1.1, dz=0.7)
    N = po.npts(cfg)
    n = 7
    rng = np.random.default_rng(4)
    w1 = rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))
    w2 = rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))
    re_ref, im_ref = np.zeros(n), np.zeros(n)
    po.lib().tdwo_overlap(C.byref(cfg), n, po.ptr(w1), po.ptr(w2), po.ptr(re_ref), po.ptr(im_ref))
    a = rng.standard_normal(100003)
    ctx = td.TDContext(tdgauss.to_params(cfg), n, 0)
    L = td.lib()
    L.tdw_bind(ctx.h)
    try:
        p = lambda t: C.c_void_p(t.data_ptr())
        W1, W2 = torch.as_tensor(w1).to("cuda:0"), torch.as_tensor(w2).to("cuda:0")
        re = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        im = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        td.check(L.compute_ovelap(n, p(W1), p(W2), p(re), p(im), None, 256), "compute_ovelap")
        re2 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        im2 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        td.check(L.tdw_overlap(ctx.h, n, p(W1), p(W2), p(re2), p(im2)), "tdw_overlap")
        A = torch.as_tensor(a).to("cuda:0")
        td.check(L.local_reductionR(p(A), a.size, p(A), 512, 0), "local_reductionR")
        assert L.local_reductionR(p(A), a.size, p(A), 512, 1) == td.TDW_EINVAL
        work = torch.zeros(32 * N + 16, dtype=torch.float64, device="cuda:0")
        td.check(L.set_workspace_for_cufftPlan(p(work)), "set_workspace")
        base = work.data_ptr()
        assert L.mm_get_pointer_to_kernels_workspace(N) == base
        assert L.mm_get_pointer_to_total_workspace(N) == base
        assert L.mm_get_pointer_to_derivatives_workspace(N) == base + 8 * 24 * N
        assert L.mm_get_pointer_to_fftONE_workspace(N) == base + 8 * 32 * N
        assert L.mm_get_size_of_total_workspace(N) == 8 * 32 * N
        # the work size create_cufftPlans reports covers the kernels and
        # derivatives zones and the fftONE zone behind them (2 N complex)
        ws = C.c_size_t(0)
        td.check(L.create_cufftPlans(2 * n, n, C.byref(ws)), "create_cufftPlans")
        assert ws.value >= L.mm_get_size_of_total_workspace(N) + 32 * N
        assert ws.value >= 8 * max(n, 11)
        td.check(L.memcopy_extra_data(64, p(work)), "memcopy_extra_data")
        ctx.sync()
    finally:
        L.tdw_bind(None)
    assert rel(re.cpu().numpy(), re_ref) < TOL
    # the drop-in copies the real part into overlap_im, as the reference does
    assert np.array_equal(im.cpu().numpy(), re.cpu().numpy())
    assert rel(re2.cpu().numpy(), re_ref) < TOL
    assert rel(im2.cpu().numpy(), im_ref) < TOL
    assert abs(A[0].item() - a.sum()) <= 1e-12 * np.abs(a).sum()
    ctx.close()


HKF_CASES = [(1, (32, 4, 4), 1), (1, (32, 4, 4), 2), (2, (12, 16, 6), 2), (3, (8, 12, 6), 2)]


@pytest.mark.parametrize("dim,shape,mode", HKF_CASES)
def test_apply_hamiltonian_hkf_filter(torch, dim, shape, mode):
    """Drop-in apply_hamiltonian with hkf_mode 1 / 2: V_a, V_b (and Delta for
    mode 2) are filtered in place with the Fermi-Dirac filter before H is
    applied (pca_kernels.cu:733-764).  Against the oracle's filter followed by
    its H: the filtered potentials and H psi - <H> psi at 1e-12."""
    from wslda_amd import td
    say("case: apply_hamiltonian with hkf_mode=%d, dim=%d shape=%s" % (mode, dim, shape))
    mu, T = 1.3, 0.25
    cfg = po.make_cfg(dim, *shape, dx=1.1, hkf=(mode, mu, T))
    cfg.mu_a, cfg.mu_b = 0.3, 0.25
    N = po.npts(cfg)
    n = 5
    rng = np.random.default_rng(17)
    wf = rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))
    pots = np.zeros(12 * N)
    pots[:2 * N] = rng.standard_normal(2 * N)
    pots[2 * N:4 * N] = rng.standard_normal(2 * N)
    pots[4 * N:6 * N] = 1.0
    kt = None
    if dim == 1:
        kt = rng.standard_normal(2 * n) * 0.3
    elif dim == 2:
        kt = rng.standard_normal(n) * 0.3
    lap = np.zeros_like(wf)
    po.lib().tdwo_laplace(C.byref(cfg), 2 * n, po.ptr(wf), po.ptr(lap))
    pref = pots.copy()
    po.lib().tdwo_hf_filter_d(C.byref(cfg), 2, po.ptr(pref), po.ptr(pref), mu, T)
    if mode >= 2:
        po.lib().tdwo_hf_filter_c(C.byref(cfg), po.ptr(pref[2 * N:4 * N]), po.ptr(pref[2 * N:4 * N]), mu, T)
    href = np.zeros_like(wf)
    qref = np.zeros(n)
    po.lib().tdwo_apply_h(C.byref(cfg), n, po.ptr(wf), po.ptr(href), po.ptr(lap), po.ptr(pref), po.ptr(kt), None,
                          po.ptr(qref))
    ctx = td.TDContext(tdgauss.to_params(cfg), n, 0)
    ctx.set_states(np.full(n, 0.5), kt, np.ones(n, dtype=np.int32) if dim == 1 else None)
    L = td.lib()
    L.tdw_bind(ctx.h)
    try:
        p = lambda t: C.c_void_p(t.data_ptr())
        W = torch.as_tensor(wf).to("cuda:0")
        LP = torch.as_tensor(lap).to("cuda:0")
        H = torch.zeros_like(W)
        P = torch.as_tensor(pots).to("cuda:0")
        D = torch.zeros(16 * N, dtype=torch.float64, device="cuda:0")
        KT = torch.as_tensor(kt).to("cuda:0") if kt is not None else None
        if dim == 3:
            rc = L.apply_hamiltonian(0, n, p(W), p(H), None, None, None, p(LP), None, p(D), p(P), C.c_double(0.0),
                                     None, C.c_double(1.0), 256)
        elif dim == 2:
            rc = L.apply_hamiltonian_2d(0, n, p(W), p(H), None, None, p(KT), p(LP), None, p(D), p(P),
                                        C.c_double(0.0), None, C.c_double(1.0), 256)
        else:
            rc = L.apply_hamiltonian_1d(0, n, p(W), p(H), None, p(KT), p(LP), None, p(D), p(P), C.c_double(0.0),
                                        None, C.c_double(1.0), 256)
        td.check(rc, "apply_hamiltonian")
        ctx.sync()
    finally:
        L.tdw_bind(None)
    pg = P.cpu().numpy()
    assert rel(pg[:4 * N], pref[:4 * N]) < TOL
    if mode < 2:
        assert np.array_equal(pg[2 * N:4 * N], pots[2 * N:4 * N])
    assert rel(H.cpu().numpy(), href) < TOL
    ctx.close()


@pytest.mark.parametrize("dim,shape,functional", [(1, (32, 4, 4), 111), (3, (8, 8, 8), 111), (3, (12, 12, 12), 112)])
def test_steps_with_hkf_filter(torch, dim, shape, functional):
    """PECE steps with hkf_mode = 2 (V, Delta and, for ASLDA, alpha and A
    filtered after every computation of the potentials) against the oracle's
    step with the same filtering: densities, potentials and states at 1e-12
    after 4 steps (the 3D SLDA case through the fused step)."""
    from wslda_amd import problems, td
    say("case: PECE steps with hkf_mode=2, dim=%d shape=%s functional=%d" % (dim, shape, functional))
    s = problems.uniform_start(dim, shape, n_states=24)
    p = s["params"]
    p.functional = functional
    if functional == td.ASLDA:
        p.mf_correction = 0
    p.hkf_mode, p.hkf_mu, p.hkf_T = 2, 2.0, 0.3
    L = shape[0] * p.dx
    p.vext = td.gauss_vext([(0.05, (0.1 * L, 0.0, 0.0), (-0.05, -0.05 if dim >= 2 else 0.0,
                                                          -0.05 if dim == 3 else 0.0))])
    cfg = po.Cfg()
    C.memmove(C.addressof(cfg), C.addressof(p), C.sizeof(cfg))
    wf = s["wf"] * (1.0 + 0.01 * np.random.default_rng(5).standard_normal(s["wf"].shape))
    P = po.Problem(cfg, wf, s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    n = wf.shape[0] // 2
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    fused = dim == 3 and functional == td.SLDA
    R = td.TDRun(ctx, wf, s["pots"], fused=fused)
    P.initial_measure()
    R.initial_measure()
    for it in range(4):
        P.step(it, grad_last=(it == 3))
        R.step(it, grad_last=(it == 3))
    R.finalize()
    ctx.sync()
    assert rel(R.dens.cpu().numpy(), P.dens) < TOL
    assert rel(R.pots.cpu().numpy(), P.pots) < TOL
    assert rel(R.wf.cpu().numpy(), P.wf) < 1e-11
    ctx.close()
