"""Spin-symmetric mode (SPINSYMMETRY_MODE) of the oracle, pinned by the
identities that fix it.

The reference keeps only the E > 0 half of the quasiparticle spectrum and
folds spin a into spin b (pca_dens.cu:55-118, ccpca_dens.cu:69-135,
cpca_dens.cu:68-130, pca_uniform.h:325-429, 727-790).  For a spin-symmetric
BdG Hamiltonian every state (u, v, E) has the partner (v*, -u*, -E), so the
half-spectrum densities must equal the ordinary (non-symmetric) densities of
the doubled set {(u, v, f), (v*, -u*, 1 - f)} -- for ANY u, v, f, which
makes random states a complete check of every field (rho, nu, tau, j), and
the balanced uniform solution in spin-symmetric mode must equal the ordinary
balanced solution.  A dropped u or v term, a missing factor 2 on nu or on the
entropy, or a sign slip in the currents breaks one of these.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as po

TOL = 1e-12


def rel(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def dens_of(cfg, wf, fb, kt, cnt, grad=True):
    n = wf.shape[0] // 2
    P = po.Problem(cfg, wf, fb, kt=kt, cnt=cnt)
    L = po.lib()
    if grad:
        L.tdwo_derivatives(C.byref(cfg), 2 * n, po.ptr(P.wf), po.ptr(P.ddx), po.ptr(P.ddy), po.ptr(P.ddz),
                           po.ptr(P.lap))
    d = np.zeros(12 * P.N)
    L.tdwo_densities(C.byref(cfg), n, po.ptr(P.wf), po.ptr(P.ddx) if grad else None,
                     po.ptr(P.ddy) if grad else None, po.ptr(P.ddz) if grad else None,
                     po.ptr(P.lap) if grad else None, po.ptr(P.fbeta),
                     None if kt is None else po.ptr(P.kt), None if cnt is None else po.ptr(P.cnt),
                     po.ptr(d), int(grad))
    if grad:
        L.tdwo_density_tau(C.byref(cfg), po.ptr(d))
    return d


def random_states(dim, shape, n, seed):
    nx, ny, nz = shape
    cfg0 = po.make_cfg(dim, nx, ny, nz, dx=0.9)
    N = po.npts(cfg0)
    rng = np.random.default_rng(seed)
    wf = (rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))) * 0.05
    X = np.arange(N) // (N // nx)
    wf *= np.exp(1j * 2 * np.pi * X / nx)[None, :]   # currents along x
    fb = rng.uniform(0.0, 1.0, n)
    kt = cnt = None
    if dim == 2:
        kt = rng.choice([0.0, 0.7, -1.3, math.pi / 0.9], n)
    elif dim == 1:
        kt = np.concatenate([rng.choice([0.0, 0.5, -0.8], n), rng.choice([0.0, 1.1], n)])
        cnt = rng.integers(1, 5, n).astype(np.int32)
    return wf, fb, kt, cnt


def doubled(wf, fb, kt, cnt):
    """{(u, v, f)} + {(v*, -u*, 1 - f)} with the same transverse momenta."""
    n = wf.shape[0] // 2
    u, v = wf[:n], wf[n:]
    W = np.concatenate([u, np.conj(v), v, -np.conj(u)])
    F = np.concatenate([fb, 1.0 - fb])
    K = None if kt is None else (np.concatenate([kt, kt]) if kt.size == n else
                                  np.concatenate([kt[:n], kt[:n], kt[n:], kt[n:]]))
    Cn = None if cnt is None else np.concatenate([cnt, cnt])
    return W, F, K, Cn


@pytest.mark.parametrize("dim,shape", [(1, (32, 4, 4)), (2, (12, 10, 6)), (3, (8, 10, 6))])
@pytest.mark.parametrize("grad", [True, False])
def test_half_spectrum_densities_equal_doubled_set(dim, shape, grad):
    wf, fb, kt, cnt = random_states(dim, shape, 9, 11 + dim)
    cfg_ss = po.make_cfg(dim, *shape, dx=0.9, spinsymmetry=1)
    cfg = po.make_cfg(dim, *shape, dx=0.9)
    N = po.npts(cfg)
    d_ss = dens_of(cfg_ss, wf, fb, kt, cnt, grad)
    d_full = dens_of(cfg, *doubled(wf, fb, kt, cnt), grad)
    fields = [(slice(0, 2 * N), "nu"), (slice(2 * N, 3 * N), "rho_a"), (slice(7 * N, 8 * N), "rho_b")]
    if grad:
        fields += [(slice(3 * N, 4 * N), "tau_a"), (slice(8 * N, 9 * N), "tau_b"),
                   (slice(4 * N, 7 * N), "j_a"), (slice(9 * N, 12 * N), "j_b")]
    for sl, name in fields:
        assert rel(d_ss[sl], d_full[sl]) < TOL, name
    # spin a is a copy of spin b
    assert np.array_equal(d_ss[2 * N:7 * N], d_ss[7 * N:12 * N])
    # the doubled set's currents are not trivially zero
    if grad:
        assert np.abs(d_full[4 * N:5 * N]).max() > 1e-6


@pytest.mark.parametrize("functional", [po.SLDA, po.ASLDA, po.BDG])
def test_uniform_solution_equals_balanced_full_spectrum(functional):
    NX, DX = 10, 1.2
    LXYZ = (NX * DX) ** 3
    n0 = math.ceil(LXYZ / (6 * math.pi ** 2)) / LXYZ
    kc = 0.999999 * math.pi / DX
    out = []
    for ss in (0, 1):
        cfg = po.make_cfg(3, NX, NX, NX, dx=DX, functional=functional, spinsymmetry=ss,
                          sclgth=-0.8 if functional == po.BDG else 0.0)
        u, rc = po.uniform_solve(cfg, n0, n0, kc, 0.05, eps=1e-11, mu_rule=0 if functional == po.ASLDA else 1)
        assert rc == 0
        out.append((cfg, u))
    (_, a), (_, b) = out
    for k in ("n0_a", "n0_b", "mu_a", "mu_b", "V_a", "V_b", "delta", "nu", "tau_a", "tau_b", "S",
              "ekin", "epot", "epair", "ec"):
        x, y = getattr(a, k), getattr(b, k)
        assert abs(x - y) <= 1e-10 * max(abs(x), 1e-300), (k, x, y)
    assert b.nwf * 2 == a.nwf
    assert b.mu_a == b.mu_b


def test_uniform_states_half_spectrum():
    """The E > 0 plane-wave states with the spin-symmetric densities give the
    full spectrum's densities."""
    NX, DX = 8, 1.3
    LXYZ = (NX * DX) ** 3
    n0 = math.ceil(LXYZ / (6 * math.pi ** 2)) / LXYZ
    res = []
    for ss in (0, 1):
        cfg = po.make_cfg(3, NX, NX, NX, dx=DX, spinsymmetry=ss)
        u, rc = po.uniform_solve(cfg, n0, n0, 0.999999 * math.pi / DX, 0.1, eps=1e-12)
        assert rc == 0
        wf, En, kt, cnt = po.uniform_states(cfg, u)
        if ss:
            assert (En > 0).all()
        fb = np.array([po.fbeta(E, u.beta) for E in En])
        res.append((wf.shape[0] // 2, dens_of(cfg, wf, fb, kt, cnt)))
    (n_full, d_full), (n_ss, d_ss) = res
    assert 2 * n_ss == n_full
    N = NX ** 3
    for sl, name in [(slice(0, 2 * N), "nu"), (slice(2 * N, 4 * N), "rho_a, tau_a"),
                     (slice(7 * N, 9 * N), "rho_b, tau_b")]:
        assert rel(d_ss[sl], d_full[sl]) < 1e-11, name


def test_planted_mistake_is_caught():
    """The doubled-set identity fails for the ordinary densities of the half
    spectrum (what a missing fold would give)."""
    wf, fb, kt, cnt = random_states(3, (8, 10, 6), 9, 5)
    cfg = po.make_cfg(3, 8, 10, 6, dx=0.9)
    d_half = dens_of(cfg, wf, fb, kt, cnt)
    d_full = dens_of(cfg, *doubled(wf, fb, kt, cnt))
    N = po.npts(cfg)
    assert rel(d_half[7 * N:8 * N], d_full[7 * N:8 * N]) > 1e-3
