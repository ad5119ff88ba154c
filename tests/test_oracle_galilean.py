"""Galilean boost of the uniform SLDA / ASLDA / SLDAE solutions on the
oracle: the closed forms of tests/galilean.py at 1e-12, then three PECE
steps of the boosted ASLDA state keep the densities, currents and energy."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import galilean as G

PI = math.pi
NX = 12


def solution(functional, sclgth=0.0):
    # kc = 0.75 pi keeps every boosted momentum k +- q inside the zone
    cfg = po.make_cfg(3, NX, NX, NX, dx=1.0, functional=functional, sclgth=sclgth, scheme=po.AB5AM5, dt=0.01)
    n0 = 1.0 / (3 * PI ** 2)
    u, rc = po.uniform_solve(cfg, n0 / 2, n0 / 2, 0.75 * PI, 1e-9, eps=1e-14, scmix=0.5, muchange=0.5,
                             maxiter=20000, mu_rule=0)
    assert rc == 0
    cfg.mu_a, cfg.mu_b, cfg.ec = u.mu_a, u.mu_b, u.ec
    wf, En, kt, cnt = po.uniform_states(cfg, u)
    fb = np.array([po.fbeta(E, u.beta) for E in En])
    return cfg, u, wf, fb


def measure(cfg, wf, fb, pots):
    """Densities (with gradients), potentials, H psi (no <H> subtraction)
    and energies of a state set, all on the oracle."""
    P = po.Problem(cfg, wf, fb, pots=pots)
    P.initial_measure()
    po.lib().tdwo_potentials(C.byref(cfg), 0, po.ptr(P.dens), po.ptr(P.pots))
    n = P.n
    d = [np.zeros_like(P.wf) for _ in range(4)]
    po.lib().tdwo_derivatives(C.byref(cfg), 2 * n, po.ptr(P.wf), *(po.ptr(x) for x in d))
    al = np.zeros_like(P.wf)
    po.lib().tdwo_alpha_wf(C.byref(cfg), n, po.ptr(P.wf), po.ptr(P.pots), po.ptr(al))
    po.lib().tdwo_laplace(C.byref(cfg), 2 * n, po.ptr(al), po.ptr(al))
    H = np.zeros_like(P.wf)
    po.lib().tdwo_apply_h_full(C.byref(cfg), n, po.ptr(P.wf), po.ptr(H), *(po.ptr(x) for x in d), po.ptr(al),
                               po.ptr(P.pots), None, po.ptr(np.zeros(n)), None)
    lu, lv = G.component_eigenvalues(P.wf, H, n)
    e = P.energy(0)
    return P, dict(dens=P.dens.copy(), pots=P.pots.copy(), lu=lu, lv=lv, E=e[:7].sum(), Npart=e[7] + e[8])


@pytest.mark.parametrize("functional,sclgth", [(po.SLDA, 0.0), (po.ASLDA, 0.0), (po.SLDAE, -1.1)])
@pytest.mark.parametrize("m", [(1, 0, 0), (1, -1, 1)])
def test_boost_closed_forms(functional, sclgth, m):
    cfg, u, wf, fb = solution(functional, sclgth)
    N = NX ** 3
    n = wf.shape[0] // 2
    phase, q = G.lattice_phase(NX, NX, NX, cfg.dx, m)
    kvec = G.state_momenta(wf, n, (NX, NX, NX), cfg.dx)
    assert np.abs(kvec + np.abs(q)).max() < PI and np.abs(kvec - np.abs(q)).max() < PI
    _, rest = measure(cfg, wf, fb, G.rest_pots(u, N))
    wb, pb, dmu = G.boost(u, wf, G.rest_pots(u, N), phase, q, N)
    cb = po.Cfg.from_buffer_copy(cfg)
    cb.mu_a += dmu
    cb.mu_b += dmu
    _, bst = measure(cb, wb, fb, pb)
    err = G.check_closed_forms(rest, bst, u, q, phase, kvec, N)
    bad = {k: v for k, v in err.items() if not v < 1e-12}
    assert not bad, (functional, m, err)


def test_boosted_aslda_steps_keep_flow():
    """Three PECE steps (gradients every half-step) of the boosted ASLDA
    state from its exact history: each component turns at its own eigenvalue,
    (H - <H>) psi(t) = ((lambda_u - e) u(t), (lambda_v - e) v(t)) with
    u(t) = u e^{-i (lambda_u - e) t}, so the history entries at t = -j dt are
    known in closed form.  rho, tau, j and |nu| stay put and the energy holds
    (the pairing phase turns at the rate (2 - alpha) q^2, which the densities
    do not see)."""
    cfg, u, wf, fb = solution(po.ASLDA)
    N = NX ** 3
    n = wf.shape[0] // 2
    phase, q = G.lattice_phase(NX, NX, NX, cfg.dx, (1, 0, 0))
    wb, pb, dmu = G.boost(u, wf, G.rest_pots(u, N), phase, q, N)
    cfg.mu_a += dmu
    cfg.mu_b += dmu
    P, m = measure(cfg, wb, fb, pb)
    d0, e0 = m["dens"], m["E"]
    U, V = P.wf[:n], P.wf[n:]
    wu, wv = np.sum(np.abs(U) ** 2, 1), np.sum(np.abs(V) ** 2, 1)
    e = (m["lu"] * wu + m["lv"] * wv) / (wu + wv)
    ou, ov = m["lu"] - e, m["lv"] - e
    hist = P.history()
    for j in range(5):
        t = -j * cfg.dt
        hist[j][:n] = ou[:, None] * U * np.exp(-1j * ou * t)[:, None]
        hist[j][n:] = ov[:, None] * V * np.exp(-1j * ov * t)[:, None]
    for it in range(3):
        P.step(it, grad_last=1)
    d1 = P.dens
    for k in (2, 3, 7, 8):      # rho, tau of both spins
        assert np.abs(d1[k * N:(k + 1) * N] - d0[k * N:(k + 1) * N]).max() < 1e-11 * np.abs(d0[k * N:(k + 1) * N]).max()
    for k0 in (4, 9):           # currents
        assert np.abs(d1[k0 * N:(k0 + 3) * N] - d0[k0 * N:(k0 + 3) * N]).max() < 1e-11 * np.abs(d0[k0 * N:(k0 + 3) * N]).max()
    nu0 = np.abs(d0[0:2 * N:2] + 1j * d0[1:2 * N:2])
    nu1 = np.abs(d1[0:2 * N:2] + 1j * d1[1:2 * N:2])
    assert np.abs(nu1 - nu0).max() < 1e-11 * nu0.max()
    assert abs(P.energy(3)[:7].sum() - e0) < 1e-11 * abs(e0)
