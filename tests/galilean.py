"""Galilean boost of a uniform solution: the closed forms that pin the
current-dependent terms of the ASLDA and SLDAE functionals.

Every state of the uniform solution is boosted by a lattice momentum q:
u -> e^{i q.r} u, v -> e^{-i q.r} v, Delta -> e^{2i q.r} Delta.  The
currents become j_s = n_s q and tau_s -> tau_s + q^2 n_s.  The functional's
coupling is regularised with mu - V (tdwslda_functionals.h:249-264), and V
gains the flow term (alpha_s - 1) q^2 / 2 from the j^2 terms
(:187-206), so the chemical potentials are raised by the same amount; then
mu - V, and with it the pairing field, is unchanged.  With that, a
Galilean-invariant functional must give, exactly on the lattice (all k +- q
inside the Brillouin zone, so the spectral derivatives of the plane waves are
exact):
  Delta(q) = Delta(0) e^{2i q.r},   V_s(q) = V_s(0) + (alpha_s - 1) q^2/2,
  A_s = (1 - alpha_s) q,             tau_s(q) = tau_s(0) + q^2 n_s,
  (H psi)_u = (E_k + k.q + (1 - alpha/2) q^2) u,
  (H psi)_v = (E_k + k.q - (1 - alpha/2) q^2) v   (E_k the rest energy:
     the Galilean shifts k.q +- q^2/2, less the chemical-potential shift),
  E(q) - E(0) = N q^2 / 2.
A sign error in A, in the A.grad term of H or in the j^2 terms of V breaks
these (checked by planting each in a scratch copy of the oracle)."""
import math

import numpy as np

PI = math.pi


def lattice_phase(nx, ny, nz, dx, m):
    """e^{i q.r} on the lattice for q = 2 pi m / L, and q."""
    q = np.array(m, dtype=float) * 2 * PI / (np.array([nx, ny, nz]) * dx)
    X, Y, Z = (np.arange(k) * dx for k in (nx, ny, nz))
    gx, gy, gz = np.meshgrid(X, Y, Z, indexing="ij")
    return np.exp(1j * (q[0] * gx + q[1] * gy + q[2] * gz)).ravel(), q


def state_momenta(wf, n, shape, dx):
    """Lattice momentum of each plane-wave state (its larger component)."""
    nx, ny, nz = shape
    ks = [2 * PI * np.fft.fftfreq(m, d=dx) for m in shape]
    out = np.zeros((n, 3))
    for a in range(n):
        c = wf[a] if np.abs(wf[a]).max() >= np.abs(wf[n + a]).max() else wf[n + a]
        F = np.fft.fftn(c.reshape(nx, ny, nz))
        i = np.unravel_index(np.abs(F).argmax(), F.shape)
        out[a] = [ks[0][i[0]], ks[1][i[1]], ks[2][i[2]]]
    return out


def boost(u, wf, pots, phase, q, N):
    """(wf, pots, dmu) of the boosted solution; pots start at the boosted
    solution's fields (V shifted by the flow term, Delta with the phase)."""
    n = wf.shape[0] // 2
    q2 = float(q @ q)
    dmu = (u.alph_a - 1.0) * q2 / 2.0
    w = wf.copy()
    w[:n] *= phase
    w[n:] *= np.conj(phase)
    p = pots.copy()
    p[:N] += dmu
    p[N:2 * N] += dmu
    d = u.delta * phase ** 2
    p[2 * N:4 * N:2] = d.real
    p[2 * N + 1:4 * N:2] = d.imag
    return w, p, dmu


def rest_pots(u, N):
    p = np.zeros(12 * N)
    p[:N], p[N:2 * N] = u.V_a, u.V_b
    p[2 * N:4 * N:2] = u.delta
    p[4 * N:5 * N], p[5 * N:6 * N] = u.alph_a, u.alph_b
    return p


def component_eigenvalues(wf, H, n):
    """lambda_u, lambda_v of every state: <u|(H psi)_u> / <u|u>, same for v."""
    U, V = wf[:n], wf[n:]
    lu = (np.sum(np.conj(U) * H[:n], 1) / np.sum(np.abs(U) ** 2, 1)).real
    lv = (np.sum(np.conj(V) * H[n:], 1) / np.sum(np.abs(V) ** 2, 1)).real
    return lu, lv


def check_closed_forms(rest, boosted, u, q, phase, kvec, N, tol=1e-12):
    """rest / boosted: dicts with dens, pots, lu, lv, E (sum of the first 7
    energy slots), Npart.  Returns the worst relative deviation per check."""
    al = u.alph_a
    q2 = float(q @ q)
    out = {}
    D0 = rest["pots"][2 * N:4 * N:2] + 1j * rest["pots"][2 * N + 1:4 * N:2]
    D1 = boosted["pots"][2 * N:4 * N:2] + 1j * boosted["pots"][2 * N + 1:4 * N:2]
    out["delta"] = np.abs(D1 - D0 * phase ** 2).max() / np.abs(D0).max()
    for s, (iv, ia, ij, ir, it) in enumerate(((0, 4, 4, 2, 3), (1, 5, 9, 7, 8))):
        alpha = boosted["pots"][ia * N:(ia + 1) * N]
        V0, V1 = rest["pots"][iv * N:(iv + 1) * N], boosted["pots"][iv * N:(iv + 1) * N]
        out["V_%d" % s] = np.abs(V1 - V0 - (alpha - 1.0) * q2 / 2).max() / np.abs(V0).max()
        n_s = boosted["dens"][ir * N:(ir + 1) * N]
        jscale = n_s.max() * np.sqrt(q2)
        j = boosted["dens"][ij * N:(ij + 3) * N].reshape(3, N)
        out["j_%d" % s] = max(np.abs(j[d] - n_s * q[d]).max() for d in range(3)) / jscale
        A = boosted["pots"][(6 + 3 * s) * N:(9 + 3 * s) * N].reshape(3, N)
        out["A_%d" % s] = max(np.abs(A[d] - (1.0 - alpha) * q[d]).max() for d in range(3)) / max(
            abs(1.0 - al) * np.sqrt(q2), np.sqrt(q2) * 1e-3)
        t0, t1 = rest["dens"][it * N:(it + 1) * N], boosted["dens"][it * N:(it + 1) * N]
        out["tau_%d" % s] = np.abs(t1 - t0 - q2 * n_s).max() / np.abs(t1).max()
    kq = kvec @ q
    c = (1.0 - al / 2.0) * q2
    scale = max(np.abs(rest["lu"]).max(), np.abs(rest["lv"]).max())
    out["lambda_u"] = np.abs(boosted["lu"] - rest["lu"] - kq - c).max() / scale
    out["lambda_v"] = np.abs(boosted["lv"] - rest["lv"] - kq + c).max() / scale
    out["energy"] = abs(boosted["E"] - rest["E"] - rest["Npart"] * q2 / 2) / abs(rest["E"])
    return out
