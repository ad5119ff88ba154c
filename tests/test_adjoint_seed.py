"""The adjoint seed of the plane-current observable (wslda_amd.adjoint
.PlaneCurrent), restated in numpy on a small periodic lattice: for random
states and a random direction, the directional derivative of
O = sum_plane (j_a + j_b)_z dA by central differences equals
Re sum conj(seed) . dpsi.  A wrong sign or a dropped d(g psi) term fails."""
import numpy as np


def dz(a, shape, dz_):
    """Spectral first derivative along z (the last axis), Nyquist mode zero."""
    nx, ny, nz = shape
    A = a.reshape(a.shape[0], nx, ny, nz)
    k = 2 * np.pi * np.fft.fftfreq(nz, d=dz_)
    if nz % 2 == 0:
        k[nz // 2] = 0.0
    return np.fft.ifft(1j * k * np.fft.fft(A, axis=-1), axis=-1).reshape(a.shape)


def current(psi, f, g, shape, d):
    n = psi.shape[0] // 2
    dpsi = dz(psi, shape, d)
    ja = (f[:, None] * (np.conj(psi[:n]) * dpsi[:n]).imag).sum(0)
    jb = -((1 - f)[:, None] * (np.conj(psi[n:]) * dpsi[n:]).imag).sum(0)
    return ((ja + jb) * g).sum()


def seed(psi, f, g, shape, d, drop=False, flip=False):
    n = psi.shape[0] // 2
    dpsi = dz(psi, shape, d)
    dgpsi = dz(psi * g[None, :], shape, d)
    s = np.empty_like(psi)
    extra = 0 if drop else 1
    sg = -1 if flip else 1
    s[:n] = -1j * sg * f[:, None] * (g[None, :] * dpsi[:n] + extra * dgpsi[:n])
    s[n:] = 1j * (1 - f)[:, None] * (g[None, :] * dpsi[n:] + extra * dgpsi[n:])
    return s


def test_plane_current_seed():
    rng = np.random.default_rng(7)
    shape = (4, 3, 10)
    N = int(np.prod(shape))
    n = 5
    d = 0.8
    psi = rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))
    dpsi = rng.standard_normal((2 * n, N)) + 1j * rng.standard_normal((2 * n, N))
    f = rng.uniform(0, 1, n)
    g = np.zeros(shape)
    g[:, :, 4] = 0.7
    g = g.ravel()
    h = 1e-6
    fd = (current(psi + h * dpsi, f, g, shape, d) - current(psi - h * dpsi, f, g, shape, d)) / (2 * h)
    ad = (np.conj(seed(psi, f, g, shape, d)) * dpsi).real.sum()
    assert abs(fd - ad) <= 1e-8 * abs(ad), (fd, ad)
    # planted mistakes fail it
    for kw in ({"drop": True}, {"flip": True}):
        bad = (np.conj(seed(psi, f, g, shape, d, **kw)) * dpsi).real.sum()
        assert abs(fd - bad) > 1e-3 * abs(fd), kw
