"""Quantum friction on the HIP path (calculate_quantum_friction_densities and
its 1D / 2D forms, apply_hamiltonian with qfswitch > 0) through the C ABI
against the oracle (pinned by tests/test_oracle_qf.py).  Bar: 1e-12."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss
from tests.test_oracle_qf import band_limited_states, kt_cnt

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


LATT = [(1, (32, 6, 4), 0.8), (2, (16, 12, 6), 0.9), (3, (12, 8, 12), 0.7)]


@pytest.mark.parametrize("dim,shape,dx", LATT)
@pytest.mark.parametrize("ss", [0, 1])
def test_qf_densities(torch, dim, shape, dx, ss):
    from wslda_amd import td
    say("case: quantum-friction densities dim=%d ss=%d" % (dim, ss))
    rng = np.random.default_rng(31 + dim + 5 * ss)
    n = 24
    cfg = po.make_cfg(dim, *shape, dx=dx, spinsymmetry=ss)
    N = po.npts(cfg)
    wf = band_limited_states(dim, shape, n, rng, kmax_frac=0.45) + 0.02 * (
        rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N)))
    fb = rng.uniform(0, 1, n)
    kt, cnt = kt_cnt(dim, n, rng, dx)
    P = po.Problem(cfg, wf, fb, kt=kt, cnt=cnt)
    O = po.lib()
    O.tdwo_derivatives(C.byref(cfg), 2 * n, po.ptr(P.wf), po.ptr(P.ddx), po.ptr(P.ddy), po.ptr(P.ddz),
                       po.ptr(P.lap))
    qref = np.zeros(4 * N)
    O.tdwo_qf_densities(C.byref(cfg), n, po.ptr(P.wf), po.ptr(P.lap), po.ptr(P.fbeta),
                        None if kt is None else po.ptr(P.kt), None if cnt is None else po.ptr(P.cnt), po.ptr(qref))
    ctx = td.TDContext(tdgauss.to_params(cfg), n, 0)
    ctx.set_states(fb, kt, cnt)
    L = td.lib()
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to("cuda:0")
    p = lambda t: C.c_void_p(t.data_ptr())
    W, LAP, FB = dev(P.wf), dev(P.lap), dev(fb)
    out = [torch.zeros(4 * N, dtype=torch.float64, device="cuda:0") for _ in range(2)]
    L.tdw_bind(ctx.h)
    try:
        td.check(L.calculate_quantum_friction_densities(n, p(W), p(LAP), p(FB), None, p(out[0]), 256), "qf")
        if dim == 1:
            KT, CN = dev(kt), dev(cnt)
            td.check(L.calculate_quantum_friction_densities_1d(n, p(W), p(LAP), p(KT), p(FB), None, p(CN),
                                                               p(out[1]), 256), "qf_1d")
        elif dim == 2:
            KT = dev(kt)
            td.check(L.calculate_quantum_friction_densities_2d(n, p(W), p(LAP), p(KT), p(FB), None, p(out[1]), 256),
                     "qf_2d")
        else:
            out[1] = out[0]
        ctx.sync()
    finally:
        L.tdw_bind(None)
    for o in out:
        a = o.cpu().numpy()
        assert rel(a[:2 * N], qref[:2 * N]) < TOL
        assert rel(a[2 * N:], qref[2 * N:]) < TOL
    ctx.close()


@pytest.mark.parametrize("dim", [1, 3])
def test_apply_hamiltonian_with_friction(torch, dim):
    """apply_hamiltonian(qfswitch > 0) = H with the friction terms in V_a,
    V_b, Delta; the potentials come back unchanged and the friction slots
    hold the unmodified fields (pca_kernels.cu:714-733, 835-845)."""
    from wslda_amd import td
    say("case: apply_hamiltonian with quantum friction dim=%d" % dim)
    s = tdgauss.setup(dim)
    cfg = s["cfg"]
    cfg.kF, cfg.qfalpha, cfg.qfbeta, cfg.qfgamma = s["kF"], 0.3, 0.2, 0.1
    rng = np.random.default_rng(8 + dim)
    wf = s["wf"] * (1.0 + 0.01 * rng.standard_normal(s["wf"].shape)) + 1e-3 * (
        rng.standard_normal(s["wf"].shape) + 1j * rng.standard_normal(s["wf"].shape))
    n, N = wf.shape[0] // 2, wf.shape[1]
    P = po.Problem(cfg, wf, s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    P.initial_measure()
    O = po.lib()
    it = 1200
    O.tdwo_potentials(C.byref(cfg), it, po.ptr(P.dens), po.ptr(P.pots))
    dens = np.zeros(16 * N)
    dens[:12 * N] = P.dens
    O.tdwo_qf_densities(C.byref(cfg), n, po.ptr(P.wf), po.ptr(P.lap), po.ptr(P.fbeta),
                        None if P.kt is None else po.ptr(P.kt), None if P.cnt is None else po.ptr(P.cnt),
                        po.ptr(dens[12 * N:]))
    qf_in = dens[12 * N:].copy()
    pots0 = P.pots.copy()
    qfswitch, pccoeff = 0.8, 1.3
    pots_ref, qf_ref = pots0.copy(), qf_in.copy()
    O.tdwo_qf_add(C.byref(cfg), po.ptr(qf_ref), po.ptr(pots_ref), qfswitch, pccoeff)
    out_ref = np.zeros_like(P.wf)
    q_ref = np.zeros(n)
    O.tdwo_apply_h(C.byref(cfg), n, po.ptr(P.wf), po.ptr(out_ref), po.ptr(P.lap), po.ptr(pots_ref),
                   None if P.kt is None else po.ptr(P.kt), None, po.ptr(q_ref))
    O.tdwo_qf_remove(C.byref(cfg), po.ptr(qf_ref), po.ptr(pots_ref))
    assert np.array_equal(pots_ref, pots0)

    ctx = td.TDContext(tdgauss.to_params(cfg), n, 0)
    ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    L = td.lib()
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to("cuda:0")
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    W, LAP, D, PT = dev(P.wf), dev(P.lap), dev(dens), dev(pots0)
    OUT = torch.zeros_like(W)
    work = torch.zeros(max(n, 16), dtype=torch.float64, device="cuda:0")
    L.tdw_bind(ctx.h)
    try:
        td.check(L.set_workspace_for_cufftPlan(p(work)), "workspace")
        td.check(L.memcopy_const(cfg.mu_a, cfg.mu_b, cfg.ec, cfg.t0, cfg.dt, cfg.kF), "memcopy_const")
        if dim == 1:
            KT = dev(s["kt"])
            td.check(L.apply_hamiltonian_1d(it, n, p(W), p(OUT), None, p(KT), p(LAP), None, p(D), p(PT), qfswitch,
                                            None, pccoeff, 256), "apply_hamiltonian_1d")
        else:
            td.check(L.apply_hamiltonian(it, n, p(W), p(OUT), None, None, None, p(LAP), None, p(D), p(PT), qfswitch,
                                         None, pccoeff, 256), "apply_hamiltonian")
        ctx.sync()
    finally:
        L.tdw_bind(None)
    assert rel(OUT.cpu().numpy(), out_ref) < TOL
    assert rel(work[:n].cpu().numpy(), q_ref) < TOL
    assert np.array_equal(PT.cpu().numpy(), pots0)
    assert np.array_equal(D.cpu().numpy()[12 * N:], pots0[:4 * N])
    # the friction changed H (the test is not the plain H)
    out_plain = np.zeros_like(P.wf)
    O.tdwo_apply_h(C.byref(cfg), n, po.ptr(P.wf), po.ptr(out_plain), po.ptr(P.lap), po.ptr(pots0),
                   None if P.kt is None else po.ptr(P.kt), None, po.ptr(np.zeros(n)))
    assert rel(out_plain, out_ref) > 1e-6
    ctx.close()
