"""Quantum friction in the oracle (tdwo_qf_densities, tdwo_qf_add,
tdwo_qf_remove; pca_dens.cu:360-480, pca_kernels.cu:497-539), pinned by
identities the mathematics fixes:

- Im(u* lap u) = div Im(u* grad u), so for band-limited states (all products
  below the lattice Nyquist) U_a and U_b are the spectral divergences of the
  currents j_a and j_b of the density kernel, in 1D, 2D and 3D and in
  spin-symmetric mode (the transverse -k_t^2 term is real and drops out);
- for plane waves u = al e^{ik.r}, v = be e^{ik.r} with transverse k_t,
  D = sum (k^2 + k_t^2) al be* (1 - 2f) w / tr and U = 0;
- the friction terms leave Delta unchanged when the phase of D equals that of
  Delta, scale it by (1 + qfbeta) a quarter turn away, by (1 + i qfgamma)
  with qfgamma alone, shift V_s by -U_s qfalpha / n_F, and are undone exactly.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as po

LATT = [(1, (32, 6, 4), 0.8), (2, (16, 12, 6), 0.9), (3, (12, 8, 10), 0.7)]


def band_limited_states(dim, shape, n, rng, kmax_frac=0.2):
    nx, ny, nz = shape
    cfg0 = po.make_cfg(dim, nx, ny, nz, dx=1.0)
    N = po.npts(cfg0)
    dims = [nx, ny, nz][:dim]
    idx = np.indices(dims).reshape(dim, -1)
    wf = np.zeros((2 * n, N), dtype=complex)
    for r in range(2 * n):
        for _ in range(3):
            m = [int(rng.integers(-int(kmax_frac * L), int(kmax_frac * L) + 1)) for L in dims]
            ph = sum(2 * math.pi * m[d] * idx[d] / dims[d] for d in range(dim))
            wf[r] += (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(1j * ph)
    return wf * 0.1


def kt_cnt(dim, n, rng, dz):
    if dim == 1:
        return np.concatenate([rng.choice([0.0, 0.5, -0.8], n), rng.choice([0.0, 1.1], n)]), \
            rng.integers(1, 5, n).astype(np.int32)
    if dim == 2:
        return rng.choice([0.0, 0.7, -1.3, math.pi / dz], n), None
    return None, None


def derivs(cfg, wf):
    n = wf.shape[0] // 2
    P = po.Problem(cfg, wf, np.zeros(n))
    po.lib().tdwo_derivatives(C.byref(cfg), 2 * n, po.ptr(P.wf), po.ptr(P.ddx), po.ptr(P.ddy), po.ptr(P.ddz),
                              po.ptr(P.lap))
    return P


@pytest.mark.parametrize("dim,shape,dx", LATT)
@pytest.mark.parametrize("ss", [0, 1])
def test_U_is_divergence_of_currents(dim, shape, dx, ss):
    rng = np.random.default_rng(5 + dim + 10 * ss)
    n = 6
    cfg = po.make_cfg(dim, *shape, dx=dx, spinsymmetry=ss)
    N = po.npts(cfg)
    wf = band_limited_states(dim, shape, n, rng)
    fb = rng.uniform(0, 1, n)
    kt, cnt = kt_cnt(dim, n, rng, dx)
    P = derivs(cfg, wf)
    P.fbeta[:] = fb
    L = po.lib()
    dens = np.zeros(12 * N)
    ktp = None if kt is None else po.ptr(np.ascontiguousarray(kt))
    kta = None if kt is None else np.ascontiguousarray(kt)
    cnta = None if cnt is None else np.ascontiguousarray(cnt)
    L.tdwo_densities(C.byref(cfg), n, po.ptr(P.wf), po.ptr(P.ddx), po.ptr(P.ddy), po.ptr(P.ddz), po.ptr(P.lap),
                     po.ptr(P.fbeta), None if kta is None else po.ptr(kta), None if cnta is None else po.ptr(cnta),
                     po.ptr(dens), 1)
    qf = np.zeros(4 * N)
    L.tdwo_qf_densities(C.byref(cfg), n, po.ptr(P.wf), po.ptr(P.lap), po.ptr(P.fbeta),
                        None if kta is None else po.ptr(kta), None if cnta is None else po.ptr(cnta), po.ptr(qf))
    for s, j0, U in ((0, 4, qf[:N]), (1, 9, qf[N:2 * N])):
        j = [np.ascontiguousarray(dens[(j0 + d) * N:(j0 + d + 1) * N]) for d in range(3)]
        div = np.zeros(N)
        L.tdwo_div_real(C.byref(cfg), po.ptr(j[0]), po.ptr(j[1]), po.ptr(j[2]), po.ptr(div))
        assert np.abs(div).max() > 1e-3   # the currents are not trivial
        assert np.abs(U - div).max() < 1e-11 * np.abs(div).max(), s
    if ss:
        assert np.array_equal(qf[:N], qf[N:2 * N])


@pytest.mark.parametrize("dim,shape,dx", LATT)
def test_D_of_plane_waves(dim, shape, dx):
    rng = np.random.default_rng(21 + dim)
    n = 5
    cfg = po.make_cfg(dim, *shape, dx=dx)
    N = po.npts(cfg)
    dims = list(shape)[:dim]
    idx = np.indices(dims).reshape(dim, -1)
    al = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    be = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    wf = np.zeros((2 * n, N), dtype=complex)
    K2 = np.zeros(n)
    for a in range(n):
        m = [int(rng.integers(-(L // 2) + 1, L // 2)) for L in dims]
        k = [2 * math.pi * m[d] / (dims[d] * dx) for d in range(dim)]
        e = np.exp(1j * sum(k[d] * idx[d] * dx for d in range(dim)))
        wf[a], wf[n + a] = al[a] * e, be[a] * e
        K2[a] = sum(x * x for x in k)
    fb = rng.uniform(0, 1, n)
    kt, cnt = kt_cnt(dim, n, rng, dx)
    w = np.ones(n)
    if dim == 1:
        K2 += kt[:n] ** 2 + kt[n:] ** 2
        w = cnt.astype(float)
    elif dim == 2:
        K2 += kt ** 2
        w = np.where(np.abs(kt) < 1e-12, 1.0, 2.0)
    tr = {1: shape[1] * dx * shape[2] * dx, 2: shape[2] * dx, 3: 1.0}[dim]
    P = derivs(cfg, wf)
    P.fbeta[:] = fb
    qf = np.zeros(4 * N)
    kta = None if kt is None else np.ascontiguousarray(kt)
    cnta = None if cnt is None else np.ascontiguousarray(cnt)
    po.lib().tdwo_qf_densities(C.byref(cfg), n, po.ptr(P.wf), po.ptr(P.lap), po.ptr(P.fbeta),
                               None if kta is None else po.ptr(kta), None if cnta is None else po.ptr(cnta),
                               po.ptr(qf))
    want = np.sum(K2 * al * np.conj(be) * (1 - 2 * fb) * w) / tr
    D = qf[2 * N::2] + 1j * qf[2 * N + 1::2]
    assert np.abs(D - want).max() < 1e-12 * abs(want)
    assert np.abs(qf[:2 * N]).max() < 1e-12 * abs(want)


def test_friction_terms_special_cases():
    N = 8
    cfg = po.make_cfg(1, N, 1, 1, dx=1.0, kF=1.3, qf=(0.4, 0.3, 0.2))
    nF = 1.3 ** 3 / (3 * math.pi ** 2)
    rng = np.random.default_rng(1)
    pots = np.zeros(12 * N)
    pots[:2 * N] = rng.standard_normal(2 * N)
    delta = rng.uniform(0.5, 1.0, N) * np.exp(1j * rng.uniform(-3, 3, N))
    pots[2 * N:4 * N:2], pots[2 * N + 1:4 * N:2] = delta.real, delta.imag
    orig = pots.copy()
    U = rng.standard_normal(2 * N)
    L = po.lib()
    for case, phase_shift, sw, pc in (("equal phase", 0.0, 0.5, 0.0), ("quarter turn", -math.pi / 2, 0.5, 0.0),
                                      ("gamma", 0.0, 0.5, 1.7)):
        qf = np.zeros(4 * N)
        qf[:2 * N] = U
        Dq = 0.3 * np.exp(1j * (np.angle(delta) + phase_shift))
        qf[2 * N::2], qf[2 * N + 1::2] = Dq.real, Dq.imag
        p = orig.copy()
        L.tdwo_qf_add(C.byref(cfg), po.ptr(qf), po.ptr(p), sw, pc)
        nd = p[2 * N:4 * N:2] + 1j * p[2 * N + 1:4 * N:2]
        if case == "equal phase":
            want = delta
        elif case == "quarter turn":
            want = delta * (1 + 0.3 * sw)
        else:
            want = delta * (1 + 1j * 0.2 * sw * pc)
        assert np.abs(nd - want).max() < 1e-14, case
        assert np.abs(p[:N] - (orig[:N] - U[:N] * 0.4 * sw / nF)).max() < 1e-14
        assert np.abs(p[N:2 * N] - (orig[N:2 * N] - U[N:] * 0.4 * sw / nF)).max() < 1e-14
        # originals saved over the friction densities, restored exactly
        assert np.array_equal(qf, orig[:4 * N])
        L.tdwo_qf_remove(C.byref(cfg), po.ptr(qf), po.ptr(p))
        assert np.array_equal(p, orig)
