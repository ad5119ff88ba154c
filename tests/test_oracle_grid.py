"""Grid-field helpers of the oracle (real-field gradient, per-component
derivative, divergence, Fermi-Dirac high-frequency filter, state overlaps)
against closed forms on sums of lattice plane waves.

A field sum_m a_m cos(k_m . r + phi_m) with k_m on the lattice has the exact
spectral derivative -sum a_m k_m,d sin(.), except that a component at the
Nyquist index of axis d has zero d-derivative (pca_derivative.cu:429-446);
the filter multiplies each mode by FD(|k_m|^2 / 2) = 1/(exp((e - mu)/T) + 1)
(pca_derivative.cu:784-827).  Axes have different lengths and spacings so a
swapped axis or spacing fails.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as po

LATT = [(1, (24, 1, 1), (0.7, 1.0, 1.0)), (2, (12, 8, 1), (0.9, 1.3, 1.0)), (3, (8, 12, 6), (0.9, 1.1, 0.7))]


def grid(dim, shape, sp):
    nx, ny, nz = shape
    cfg = po.make_cfg(dim, nx, ny, nz, dx=sp[0], dy=sp[1], dz=sp[2])
    idx = np.indices((nx, ny, nz) if dim == 3 else (nx, ny) if dim == 2 else (nx,)).reshape(dim, -1)
    r = [idx[d] * sp[d] for d in range(dim)]
    L = [shape[d] * sp[d] for d in range(dim)]
    return cfg, idx, r, L


def modes(dim, shape, rng):
    """(amplitude, integer wave numbers, phase); one mode at the x Nyquist index."""
    out = []
    for _ in range(4):
        m = [int(rng.integers(-(shape[d] // 2) + 1, shape[d] // 2)) for d in range(dim)]
        out.append((rng.uniform(0.5, 1.5), m, rng.uniform(0, 2 * math.pi)))
    nyq = [shape[0] // 2] + [int(rng.integers(0, shape[d] // 2)) for d in range(1, dim)]
    out.append((0.7, nyq, 0.0))
    return out


def field(modes_, r, L, dim, deriv=None):
    f = 0.0
    for a, m, ph in modes_:
        k = [2 * math.pi * m[d] / L[d] for d in range(dim)]
        arg = sum(k[d] * r[d] for d in range(dim)) + ph
        if deriv is None:
            f = f + a * np.cos(arg)
        else:
            kd = 0.0 if abs(m[deriv]) * 2 == SHAPES[dim][deriv] else k[deriv]
            f = f - a * kd * np.sin(arg)
    return f


SHAPES = {d: s for d, s, _ in LATT}


@pytest.mark.parametrize("dim,shape,sp", LATT)
def test_gradient_and_divergence_closed_form(dim, shape, sp):
    cfg, idx, r, L = grid(dim, shape, sp)
    N = po.npts(cfg)
    rng = np.random.default_rng(3 + dim)
    M = [modes(dim, shape, rng) for _ in range(3)]
    f = [np.ascontiguousarray(field(M[d], r, L, dim)) for d in range(3) if d < dim]
    while len(f) < 3:
        f.append(np.zeros(N))
    lib = po.lib()
    g = [np.zeros(N) for _ in range(3)]
    lib.tdwo_gradient_real(C.byref(cfg), po.ptr(f[0]), po.ptr(g[0]), po.ptr(g[1]), po.ptr(g[2]))
    scale = max(abs(a) * 2 * math.pi * max(abs(x) for x in m) / min(L) for Md in M for a, m, _ in Md) + 1.0
    for d in range(dim):
        assert np.abs(g[d] - field(M[0], r, L, dim, deriv=d)).max() < 1e-12 * scale, d
    # the x-Nyquist mode contributes nothing to d/dx, but does to d/dy
    dv = [np.zeros(N) for _ in range(3)]
    lib.tdwo_dvec_real(C.byref(cfg), po.ptr(f[0]), po.ptr(f[1]), po.ptr(f[2]), po.ptr(dv[0]), po.ptr(dv[1]),
                       po.ptr(dv[2]))
    div = np.zeros(N)
    lib.tdwo_div_real(C.byref(cfg), po.ptr(f[0]), po.ptr(f[1]), po.ptr(f[2]), po.ptr(div))
    want = 0.0
    for d in range(dim):
        exact = field(M[d], r, L, dim, deriv=d)
        assert np.abs(dv[d] - exact).max() < 1e-11 * scale, d
        want = want + exact
    assert np.abs(div - want).max() < 1e-11 * scale


@pytest.mark.parametrize("dim,shape,sp", LATT)
def test_fd_filter_closed_form(dim, shape, sp):
    cfg, idx, r, L = grid(dim, shape, sp)
    N = po.npts(cfg)
    rng = np.random.default_rng(11 + dim)
    M = modes(dim, shape, rng)
    mu, T = 1.7, 0.4

    def fd(m):
        e = 0.5 * sum((2 * math.pi * m[d] / L[d]) ** 2 for d in range(dim))
        return 1.0 / (math.exp((e - mu) / T) + 1.0)

    f = np.ascontiguousarray(field(M, r, L, dim))
    want = sum(a * fd(m) * np.cos(sum(2 * math.pi * m[d] / L[d] * r[d] for d in range(dim)) + ph)
               for a, m, ph in M)
    out = np.zeros(2 * N)
    f2 = np.ascontiguousarray(np.stack([f, 2 * f]))
    po.lib().tdwo_hf_filter_d(C.byref(cfg), 2, po.ptr(f2), po.ptr(out), mu, T)
    assert np.abs(out[:N] - want).max() < 1e-13 * np.abs(f).max()
    assert np.abs(out[N:] - 2 * want).max() < 1e-13 * np.abs(f).max()
    # the filter weights differ mode by mode (the test is not trivial)
    w = [fd(m) for _, m, _ in M]
    assert max(w) - min(w) > 0.1
    # complex field: sum a_m exp(i k_m . r)
    z = sum(a * np.exp(1j * (sum(2 * math.pi * m[d] / L[d] * r[d] for d in range(dim)) + ph)) for a, m, ph in M)
    zw = sum(a * fd(m) * np.exp(1j * (sum(2 * math.pi * m[d] / L[d] * r[d] for d in range(dim)) + ph))
             for a, m, ph in M)
    z = np.ascontiguousarray(z)
    zo = np.zeros_like(z)
    po.lib().tdwo_hf_filter_c(C.byref(cfg), po.ptr(z), po.ptr(zo), mu, T)
    assert np.abs(zo - zw).max() < 1e-13 * np.abs(z).max()


def test_overlap():
    cfg = po.make_cfg(3, 8, 12, 6, dx=0.9, dy=1.1, dz=0.7)
    N = po.npts(cfg)
    rng = np.random.default_rng(2)
    n = 5
    w1 = rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))
    w2 = rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))
    re, im = np.zeros(n), np.zeros(n)
    po.lib().tdwo_overlap(C.byref(cfg), n, po.ptr(w1), po.ptr(w2), po.ptr(re), po.ptr(im))
    dv = 0.9 * 1.1 * 0.7
    want = np.array([np.vdot(w1[a], w2[a]) + np.vdot(w1[n + a], w2[n + a]) for a in range(n)]) * dv
    assert np.abs(re - want.real).max() < 1e-12 * np.abs(want).max()
    assert np.abs(im - want.imag).max() < 1e-12 * np.abs(want).max()
