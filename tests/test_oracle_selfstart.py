"""The oracle's Taylor self-start (pca.c:1086-1392) against the exact
propagator of a linear problem.

With the BdG functional at scattering length 0 the pairing field vanishes
and there is no mean field, so H = diag(-lap/2 + k_t^2/2 + V_a - mu_a,
lap/2 - k_t^2/2 - V_b + mu_b) is fixed in time and the exact evolution of
i d psi/dt = (H - <H>) psi is psi(t) = exp(-i (H - e0) t) psi(0), e0 =
<psi(0)|H|psi(0)> (conserved).  The self-start takes K midpoint steps whose
propagator is the order-E Taylor polynomial of exp(-i (H - <H>) dt), so the
state after it must equal exp(-i (H - e0) K dt) psi(0) up to the truncation
error, which falls as dt^(E+1): halving dt must cut the error by ~2^(E+1).
The history must hold the reference's entries: fkm1 = (H - e0)
psi((K-1) dt) and fkm2..fkmK = (H - e0) psi((K-2) dt) (the host shift of
pca.c:1166 copies slot j to j+1 for increasing j).  H is built here as a
dense matrix from the DFT matrix and diagonalised by scipy.linalg.expm, not
from the oracle's FFT.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg

from oracle import pyoracle as po

NX = 16


def setup(scheme, dt, seed=3):
    cfg = po.make_cfg(1, NX, 4, 4, dx=1.0, functional=po.BDG, regularization=po.SPHERICAL, scheme=scheme,
                      mu_a=0.3, mu_b=0.2, ec=10.0, dt=dt, sclgth=0.0,
                      vext=po.gauss_vext([(0.2, (1.0, 0.0, 0.0), (-0.08, 0.0, 0.0))]))
    n = 3
    rng = np.random.default_rng(seed)
    wf = rng.standard_normal((2 * n, NX)) + 1j * rng.standard_normal((2 * n, NX))
    # smooth states (few modes) so the truncation error is measured, not aliasing
    F = np.fft.fft(wf, axis=1)
    F[:, np.abs(np.fft.fftfreq(NX) * NX) > 4] = 0.0
    wf = np.fft.ifft(F, axis=1)
    kt = rng.uniform(-0.8, 0.8, 2 * n)
    cnt = np.ones(n, dtype=np.int32)
    fbeta = np.full(n, 0.5)
    pots = np.zeros(12 * NX)
    pots[4 * NX:6 * NX] = 1.0
    return cfg, wf, kt, cnt, fbeta, pots


def dense_h(cfg, kt, n):
    """Per state: the u and v blocks of H as dense matrices."""
    x = cfg.dx * (np.arange(NX) - NX // 2)
    va = np.zeros(NX)
    vb = np.zeros(NX)
    po.lib().tdwo_vext_eval(C.byref(cfg), 0, po.ptr(va), po.ptr(vb))
    k = 2 * np.pi * np.fft.fftfreq(NX, d=cfg.dx)
    Fm = np.fft.fft(np.eye(NX), axis=0)                  # DFT matrix
    lap = np.linalg.solve(Fm, np.diag(-k * k) @ Fm)      # F^-1 diag(-k^2) F
    out = []
    for a in range(n):
        k2t = kt[a] ** 2 + kt[n + a] ** 2
        hu = -0.5 * lap + np.diag(0.5 * k2t + va - cfg.mu_a)
        hv = 0.5 * lap - np.diag(0.5 * k2t + vb - cfg.mu_b)
        out.append((hu, hv))
    return out


def exact(cfg, wf0, kt, n, t):
    """psi(t) and (H - e0) psi(t) of every state, psi(0) normalised."""
    H = dense_h(cfg, kt, n)
    psi, F = np.zeros_like(wf0), np.zeros_like(wf0)
    for a, (hu, hv) in enumerate(H):
        u, v = wf0[a], wf0[n + a]
        nr = np.vdot(u, u).real + np.vdot(v, v).real
        e0 = (np.vdot(u, hu @ u).real + np.vdot(v, hv @ v).real) / nr
        ut = scipy.linalg.expm(-1j * (hu - e0 * np.eye(NX)) * t) @ u
        vt = scipy.linalg.expm(-1j * (hv - e0 * np.eye(NX)) * t) @ v
        psi[a], psi[n + a] = ut, vt
        F[a], F[n + a] = hu @ ut - e0 * ut, hv @ vt - e0 * vt
    return psi, F


def run(scheme, dt):
    cfg, wf, kt, cnt, fbeta, pots = setup(scheme, dt)
    P = po.Problem(cfg, wf, fbeta, kt=kt, cnt=cnt, pots=pots)
    P.initial_measure()          # normalises (the reference starts from normalised states)
    wf0 = P.wf.copy()
    assert P.selfstart() == 0
    return cfg, P, wf0, kt


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("scheme", [po.AB3AM4, po.AB4AM5, po.AB5AM5])
def test_selfstart_matches_exact_propagator(scheme):
    K = {po.AB3AM4: 3, po.AB4AM5: 4, po.AB5AM5: 5}[scheme]
    E = {po.AB3AM4: 4, po.AB4AM5: 5, po.AB5AM5: 5}[scheme]
    errs = []
    for dt in (0.1, 0.05):
        cfg, P, wf0, kt = run(scheme, dt)
        n = P.n
        # the uniform pairing field stays zero: the problem is linear
        assert np.abs(P.pots[2 * NX:4 * NX]).max() == 0.0
        # normalisation of the dense states as the oracle does it
        psiK, _ = exact(cfg, wf0, kt, n, K * dt)
        err = rel(P.wf, psiK)
        errs.append(err)
        _, F1 = exact(cfg, wf0, kt, n, (K - 1) * dt)
        _, F2 = exact(cfg, wf0, kt, n, (K - 2) * dt)
        hist = P.history()
        assert rel(hist[0], F1) < 50 * err + 1e-12
        for j in range(1, K):
            assert rel(hist[j], F2) < 50 * err + 1e-12, j
            assert np.array_equal(hist[j], hist[1])
    # truncation error of the order-E Taylor polynomial: ~dt^(E+1) per run
    ratio = errs[0] / errs[1]
    assert errs[0] < 1e-4
    assert ratio > 0.6 * 2 ** (E + 1), (errs, ratio)


def test_selfstart_of_stationary_start_is_identity():
    """Plane-wave eigenstates of a self-consistent uniform SLDA solution:
    (H - <H>) psi = 0, so every Taylor term vanishes, the state is unchanged
    and the history is zero."""
    from wslda_amd import problems
    s = problems.uniform_start(1, (32, 4, 4), n_states=24)
    p = s["params"]
    cfg = po.Cfg()
    C.memmove(C.addressof(cfg), C.addressof(p), C.sizeof(cfg))
    P = po.Problem(cfg, s["wf"], s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    P.initial_measure()
    w0 = P.wf.copy()
    P.selfstart()
    # a 24-state slice is not the whole self-consistent set: compare with
    # the full set instead when it is small enough
    s2 = problems.uniform_start(1, (32, 4, 4))
    P2 = po.Problem(cfg, s2["wf"], s2["fbeta"], kt=s2["kt"], cnt=s2["cnt"], pots=s2["pots"])
    P2.initial_measure()
    w2 = P2.wf.copy()
    P2.selfstart()
    assert rel(P2.wf, w2) < 1e-12
    for h in P2.history():
        assert np.abs(h).max() < 1e-12 * np.abs(w2).max() * 10
