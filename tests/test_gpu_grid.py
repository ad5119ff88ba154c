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
        td.check(L.memcopy_extra_data(64, p(work)), "memcopy_extra_data")
        ctx.sync()
    finally:
        L.tdw_bind(None)
    assert rel(re.cpu().numpy(), re_ref) < TOL
    assert rel(im.cpu().numpy(), im_ref) < TOL
    assert abs(A[0].item() - a.sum()) <= 1e-12 * np.abs(a).sum()
    ctx.close()
