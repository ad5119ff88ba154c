"""The oracle against what mathematics fixes, for the parts the td-gauss
goldens do not isolate: FFT against a brute-force DFT, spectral derivatives of
lattice plane waves, the ABM predictor-corrector's order of accuracy, the
uniform BdG state's densities / potentials / stationarity, and the product's
host-side uniform solver against the oracle's.  Each check is chosen so that a
plausible slip (a wrong sign, index, coefficient or missing term) fails it.
CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as po

PI = math.pi


def _fft(cfg, z, sign):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    po.lib().tdwo_fft(C.byref(cfg), po.ptr(z), sign)
    return z


def _dft_matrix(n, sign):
    j = np.arange(n)
    return np.exp(sign * 2j * PI * np.outer(j, j) / n)


@pytest.mark.parametrize("dim,shape", [(1, (6, 1, 1)), (1, (10, 1, 1)), (1, (24, 1, 1)), (1, (15, 1, 1)),
                                       (2, (12, 8, 1)), (3, (4, 6, 5)), (3, (8, 3, 2))])
def test_fft_equals_brute_force_dft(dim, shape):
    """Forward (sign -1) and inverse (+1) unnormalised transforms equal the
    DFT written as a matrix product along every axis (C order, z fastest)."""
    nx, ny, nz = shape
    cfg = po.make_cfg(dim, nx, ny, nz)
    dims = [nx, ny, nz][:dim]
    rng = np.random.default_rng(11)
    z = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    for sign in (-1, +1):
        ref = z.copy()
        for ax, n in enumerate(dims):
            ref = np.moveaxis(np.tensordot(_dft_matrix(n, sign), np.moveaxis(ref, ax, 0), axes=(1, 0)), 0, ax)
        got = _fft(cfg, z.ravel(), sign).reshape(dims)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def _kgrid(n, L):
    i = np.arange(n)
    return np.where(i < n // 2, 2 * PI / L * i, 2 * PI / L * (i - n))


@pytest.mark.parametrize("shape,dx", [((8, 6, 10), 1.0), ((6, 8, 4), 0.7)])
def test_spectral_derivatives_of_plane_waves(shape, dx):
    """lap e^{ik.r} = -k^2 e^{ik.r};  d_a e^{ik.r} = i k_a e^{ik.r}, zero when
    k_a is the axis' Nyquist momentum (pca_derivative.cu:169-171)."""
    nx, ny, nz = shape
    cfg = po.make_cfg(3, nx, ny, nz, dx=dx)
    N = nx * ny * nz
    x = np.arange(nx)[:, None, None] * dx
    y = np.arange(ny)[None, :, None] * dx
    z = np.arange(nz)[None, None, :] * dx
    kxs, kys, kzs = _kgrid(nx, nx * dx), _kgrid(ny, ny * dx), _kgrid(nz, nz * dx)
    # include each axis' Nyquist index and a negative one
    picks = [(1, 2, 3), (nx // 2, 1, nz - 1), (nx - 1, ny // 2, nz // 2), (0, 0, 0)]
    wf = np.zeros((len(picks), N), dtype=np.complex128)
    exp_lap = np.zeros_like(wf)
    exp_d = [np.zeros_like(wf) for _ in range(3)]
    for s, (i, j, k) in enumerate(picks):
        kv = (kxs[i], kys[j], kzs[k])
        e = np.exp(1j * (kv[0] * x + kv[1] * y + kv[2] * z)).ravel()
        wf[s] = e
        exp_lap[s] = -(kv[0] ** 2 + kv[1] ** 2 + kv[2] ** 2) * e
        for a, (idx, n) in enumerate(zip((i, j, k), (nx, ny, nz))):
            exp_d[a][s] = 0.0 if idx == n // 2 else 1j * kv[a] * e
    lap = np.zeros_like(wf)
    po.lib().tdwo_laplace(C.byref(cfg), len(picks), po.ptr(wf), po.ptr(lap))
    scale = (PI / dx) ** 2 * 3
    assert np.abs(lap - exp_lap).max() < 1e-12 * scale
    d = [np.zeros_like(wf) for _ in range(3)]
    lap2 = np.zeros_like(wf)
    po.lib().tdwo_derivatives(C.byref(cfg), len(picks), po.ptr(wf), po.ptr(d[0]), po.ptr(d[1]), po.ptr(d[2]),
                              po.ptr(lap2))
    assert np.abs(lap2 - exp_lap).max() < 1e-12 * scale
    for a in range(3):
        assert np.abs(d[a] - exp_d[a]).max() < 1e-12 * (PI / dx), a


def test_laplace_real_of_cosines():
    nx, ny, nz = 8, 6, 10
    cfg = po.make_cfg(3, nx, ny, nz, dx=0.9)
    x = np.arange(nx)[:, None, None] * 0.9
    y = np.arange(ny)[None, :, None] * 0.9
    z = np.arange(nz)[None, None, :] * 0.9
    k = (2 * PI / (nx * 0.9) * 3, 2 * PI / (ny * 0.9) * 1, 2 * PI / (nz * 0.9) * 2)
    f = (np.cos(k[0] * x) * np.cos(k[1] * y) + 0.5 * np.sin(k[2] * z) + 2.0).ravel()
    exp = (-(k[0] ** 2 + k[1] ** 2) * np.cos(k[0] * x) * np.cos(k[1] * y) - k[2] ** 2 * 0.5 * np.sin(k[2] * z)).ravel()
    lap = np.zeros_like(f)
    po.lib().tdwo_laplace_real(C.byref(cfg), po.ptr(f), po.ptr(lap))
    assert np.abs(lap - exp).max() < 1e-12 * 10


def _abm_error(scheme, nsteps, T=1.0, lam=2.3):
    """PECE solution of y' = -i lam y on [0, T] with exact starting history;
    error of y(T) against exp(-i lam T)."""
    h = {po.AB3AM4: 3, po.AB4AM5: 4, po.AB5AM5: 5}[scheme]
    dt = T / nsteps
    cfg = po.make_cfg(1, 2, 1, 1, scheme=scheme, dt=dt)
    y0 = np.array([1.0 + 0.5j, -0.3 + 0.2j, 0.7j, 0.25], dtype=np.complex128)   # 2 components x 2 points
    ex = lambda t: np.exp(-1j * lam * t) * y0
    y = ex(0.0).copy()
    f = [lam * ex(-j * dt) for j in range(h)] + [np.zeros(4, dtype=np.complex128) for _ in range(5 - h)]
    L = po.lib()
    for _ in range(nsteps):
        L.tdwo_abm_step1(C.byref(cfg), 1, po.ptr(y), *[po.ptr(a) for a in f])
        hy = lam * y
        L.tdwo_abm_step4(C.byref(cfg), 1, po.ptr(hy), po.ptr(y), po.ptr(f[h - 1]))
        # rotate: the slot that held the partial corrector takes f(y_{k+1})
        f = [lam * y] + f[:h - 1] + f[h:]
    return np.abs(y - ex(T)).max()


@pytest.mark.parametrize("scheme,order", [(po.AB3AM4, 4), (po.AB4AM5, 5), (po.AB5AM5, 5)])
def test_abm_order_of_accuracy(scheme, order):
    """Halving dt divides the global error by 2^order; a wrong coefficient or
    history index makes the scheme inconsistent (order <= 1)."""
    e1 = _abm_error(scheme, 40)
    e2 = _abm_error(scheme, 80)
    rate = math.log2(e1 / e2)
    assert abs(rate - order) < 0.3, (e1, e2, rate)
    assert e2 < 1e-6


@pytest.fixture(scope="module")
def uniform6():
    """Self-consistent uniform SLDA state on a 6^3 lattice, every plane-wave
    state below the cutoff."""
    cfg = po.make_cfg(3, 6, 6, 6, dx=1.0, functional=po.SLDA, regularization=po.SPHERICAL, mf_correction=1,
                      scheme=po.AB5AM5, dt=0.01)
    n0 = 1.0 / (3 * PI ** 2)
    kc = 0.999999 * PI
    u, rc = po.uniform_solve(cfg, n0 / 2, n0 / 2, kc, 1e-9, eps=1e-12, scmix=0.5, muchange=0.5, maxiter=20000)
    assert rc == 0
    cfg.mu_a, cfg.mu_b, cfg.ec = u.mu_a, u.mu_b, u.ec
    wf, En, kt, cnt = po.uniform_states(cfg, u)
    fb = np.array([po.fbeta(e, u.beta) for e in En])
    N = 216
    pots = np.zeros(12 * N)
    pots[:N], pots[N:2 * N] = u.V_a, u.V_b
    pots[2 * N:4 * N:2] = u.delta
    pots[4 * N:6 * N] = 1.0
    return dict(cfg=cfg, u=u, wf=wf, En=En, fb=fb, pots=pots, N=N)


def test_uniform_states_reproduce_solver_densities(uniform6):
    """Densities of the plane-wave states equal the uniform solver's n_a, n_b,
    nu and tau at every lattice point (pca_uniform.h:380-407)."""
    s = uniform6
    P = po.Problem(s["cfg"], s["wf"], s["fb"], pots=s["pots"])
    P.initial_measure()
    N, u, d = s["N"], s["u"], P.dens
    assert np.allclose(d[2 * N:3 * N], u.n0_a, rtol=1e-10, atol=0)
    assert np.allclose(d[7 * N:8 * N], u.n0_b, rtol=1e-10, atol=0)
    assert np.allclose(d[0:2 * N:2], u.nu, rtol=1e-9, atol=0)
    assert np.abs(d[1:2 * N:2]).max() < 1e-12 * abs(u.nu)
    assert np.allclose(d[3 * N:4 * N], u.tau_a, rtol=1e-9, atol=0)
    assert np.allclose(d[8 * N:9 * N], u.tau_b, rtol=1e-9, atol=0)
    # currents of a set closed under k -> -k vanish
    for sl in (slice(4 * N, 7 * N), slice(9 * N, 12 * N)):
        assert np.abs(d[sl]).max() < 1e-12


def test_uniform_state_is_stationary(uniform6):
    """With self-consistent fields every plane-wave state is an eigenstate,
    (H - <H>) psi = 0: potentials recomputed from the densities equal the
    solver's, and a few PECE steps change neither densities nor energy."""
    s = uniform6
    N = s["N"]
    P = po.Problem(s["cfg"], s["wf"], s["fb"], pots=s["pots"].copy())
    P.initial_measure()
    pots = P.pots.copy()
    po.lib().tdwo_potentials(C.byref(s["cfg"]), 0, po.ptr(P.dens), po.ptr(pots))
    u = s["u"]
    assert np.allclose(pots[:N], u.V_a, rtol=1e-8, atol=1e-12)
    assert np.allclose(pots[2 * N:4 * N:2], u.delta, rtol=1e-8, atol=0)
    P.pots[:] = pots
    d0 = P.dens.copy()
    e0 = P.energy(0)
    # <H> of each state equals its BdG energy E_n (the u/v eigenvalue)
    out = np.zeros_like(P.wf)
    lap = np.zeros_like(P.wf)
    qpe = np.zeros(P.n)
    po.lib().tdwo_laplace(C.byref(s["cfg"]), 2 * P.n, po.ptr(P.wf), po.ptr(lap))
    po.lib().tdwo_apply_h(C.byref(s["cfg"]), P.n, po.ptr(P.wf), po.ptr(out), po.ptr(lap), po.ptr(P.pots), None,
                          None, po.ptr(qpe))
    assert np.abs(qpe - s["En"]).max() < 1e-8
    assert np.abs(out).max() < 1e-8
    P.f = P.f  # history starts at zero, which is the exact history here
    for it in range(4):
        P.step(it, grad_last=(it == 3))
    assert np.abs(P.dens - d0).max() <= 1e-10 * np.abs(d0).max()
    e1 = P.energy(4)
    assert abs(sum(e1[:7]) - sum(e0[:7])) <= 1e-10 * abs(sum(e0[:7]))


def test_product_uniform_solver_matches_oracle():
    """The library's host-side uniform solver and plane-wave generator (used
    to set up runs) agree with the oracle's restatement."""
    from wslda_amd import td
    p = td.make_params(3, 6, 6, 6, dx=1.0, functional=td.SLDA, regularization=td.SPHERICAL_CUTOFF, mf_correction=1,
                       scheme=td.AB5AM5, dt=0.01)
    cfg = po.Cfg()
    C.memmove(C.addressof(cfg), C.addressof(p), C.sizeof(cfg))
    n0 = 1.0 / (3 * PI ** 2)
    kc = 0.999999 * PI
    ur, rc = po.uniform_solve(cfg, n0 / 2, n0 / 2, kc, 1e-3, eps=1e-12, scmix=0.5, muchange=0.5, maxiter=20000)
    assert rc == 0
    ug, rcg = td.uniform_solve(p, n0 / 2, n0 / 2, kc, 1e-3, eps=1e-12, scmix=0.5, muchange=0.5, maxiter=20000)
    assert rcg == 0
    for k in ("mu_a", "mu_b", "V_a", "V_b", "delta", "nu", "tau_a", "tau_b", "ekin", "epot", "epair", "ec"):
        a, b = getattr(ug, k), getattr(ur, k)
        assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), k
    assert ug.nwf == ur.nwf
    cfg.mu_a, cfg.mu_b, cfg.ec = ur.mu_a, ur.mu_b, ur.ec
    p.mu_a, p.mu_b, p.ec = ug.mu_a, ug.mu_b, ug.ec
    wr, Er, _, _ = po.uniform_states(cfg, ur)
    wg, Eg, _, _ = td.uniform_states(p, ug)
    assert wg.shape == wr.shape
    assert np.abs(Eg - Er).max() < 1e-12
    assert np.abs(wg - wr).max() < 1e-12 * np.abs(wr).max()


def test_product_bdg_uniform_solver_matches_oracle():
    """The BdG branch of the library's uniform solver (td-gauss-bdg inputs:
    12^3, DX 1.15, aBdG -1.1, T 0.05 eF) against the oracle's."""
    from wslda_amd import td
    from tests import tdgauss
    s = tdgauss.setup(3, functional=po.BDG)
    cfg = s["cfg"]
    p = tdgauss.to_params(cfg)
    g = tdgauss.golden(po.BDG)["inputs"]
    L3 = (g["NX"] * g["DX"]) ** 3
    na, nb = s["u"].n0_a, s["u"].n0_b
    kc = 0.999999 * PI / g["DX"]
    ug, rc = td.uniform_solve(p, na, nb, kc, g["temperature"], eps=g["init0eps"], scmix=g["init0scmix"],
                              muchange=g["init0muchange"], maxiter=g["init0maxiter"])
    ur, rcr = po.uniform_solve(cfg, na, nb, kc, g["temperature"], eps=g["init0eps"], scmix=g["init0scmix"],
                               muchange=g["init0muchange"], maxiter=g["init0maxiter"])
    assert rc == 0 and rcr == 0
    for k in ("mu_a", "mu_b", "V_a", "V_b", "delta", "nu", "tau_a", "tau_b", "ekin", "epair", "ec"):
        a, b = getattr(ug, k), getattr(ur, k)
        assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), k
    assert ug.nwf == ur.nwf and ug.epot == 0.0
