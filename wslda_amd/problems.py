"""Synthetic td workloads of the benchmark configurations.

Every workload is a uniform SLDA (or prescribed BdG) state expanded in plane
waves exactly as pca_uniform.h's create_uniform_wf does it (order of states,
u_k, v_k, branch signs, 1/sqrt(V) normalisation), then optionally perturbed
(noise, envelopes, imprinted phases) and fed to the td step.  Large state
sets are generated directly on the GPU.

Units: kF = 1, so eF = 1/2; energies and times below are quoted in eF.
"""
import math

import numpy as np

from . import td

PI = math.pi


def kaxis(n, L):
    i = np.arange(n)
    return np.where(i < n // 2, 2 * PI / L * i, 2 * PI / L * (i - n))


def bdg_branch(eta_a, eta_b, delta, branch):
    """Solution `branch` (0: upper, 1: lower) of the uniform 2x2 BdG problem
    (pca_uniform.h:689-696): energy E, u_k, v_k (v_k signed)."""
    sq = np.sqrt((eta_b - eta_a) ** 2 + 4.0 * (eta_a * eta_b + delta * delta))
    if branch == 0:
        E = -0.5 * (eta_b - eta_a) + 0.5 * sq
        v2 = delta * delta / ((0.5 * (eta_a + eta_b) + 0.5 * sq) ** 2 + delta * delta)
    else:
        E = -0.5 * (eta_b - eta_a) - 0.5 * sq
        v2 = delta * delta / ((0.5 * (eta_a + eta_b) - 0.5 * sq) ** 2 + delta * delta)
    u = np.sqrt(1.0 - v2)
    v = np.sqrt(v2)
    v = np.where(eta_b + E < 0.0, -v, v)
    return E, u, v


def plane_wave_list(params, mu_a, mu_b, V_a, V_b, delta, ec, alph=1.0):
    """(kx, ky, kz, E, u, v) of every state with -ec < E < ec in the order of
    create_uniform_wf (3D, all k) -- loops ix, iy, iz, then the two branches."""
    assert params.dim == 3
    kx = kaxis(params.nx, params.nx * params.dx)
    ky = kaxis(params.ny, params.ny * params.dy)
    kz = kaxis(params.nz, params.nz * params.dz)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = (KX ** 2 + KY ** 2 + KZ ** 2).ravel()
    ea = alph * k2 / 2.0 + V_a - mu_a
    eb = alph * k2 / 2.0 + V_b - mu_b
    rows = []
    for br in (0, 1):
        E, u, v = bdg_branch(ea, eb, delta, br)
        rows.append((E, u, v))
    # interleave the branches per k point, keep |E| < ec
    E = np.stack([rows[0][0], rows[1][0]], 1).ravel()
    U = np.stack([rows[0][1], rows[1][1]], 1).ravel()
    Vv = np.stack([rows[0][2], rows[1][2]], 1).ravel()
    K = np.repeat(np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], 1), 2, axis=0)
    sel = (E > -ec) & (E < ec)
    return K[sel], E[sel], U[sel], Vv[sel]


def states_on_gpu(torch, params, K, U, V, device="cuda:0", chunk=2048, envelope=None, phase=None):
    """psi_n(r) = exp(i k_n.r) (u_n, v_n) / sqrt(V_box) on the lattice, 3D,
    optionally times an envelope(r) and exp(i*phase(r)) on the u component."""
    n = K.shape[0]
    nx, ny, nz = params.nx, params.ny, params.nz
    N = nx * ny * nz
    vol = (nx * params.dx) * (ny * params.dy) * (nz * params.dz)
    X = torch.arange(nx, dtype=torch.float64, device=device) * params.dx
    Y = torch.arange(ny, dtype=torch.float64, device=device) * params.dy
    Z = torch.arange(nz, dtype=torch.float64, device=device) * params.dz
    gx, gy, gz = torch.meshgrid(X, Y, Z, indexing="ij")
    gx, gy, gz = gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)
    wf = torch.empty((2 * n, N), dtype=torch.complex128, device=device)
    Kt = torch.as_tensor(K, device=device)
    Ut = torch.as_tensor(U, device=device)
    Vt = torch.as_tensor(V, device=device)
    s = 1.0 / math.sqrt(vol)
    for a0 in range(0, n, chunk):
        a1 = min(n, a0 + chunk)
        ph = Kt[a0:a1, 0:1] * gx + Kt[a0:a1, 1:2] * gy + Kt[a0:a1, 2:3] * gz
        e = torch.polar(torch.ones_like(ph), ph) * s
        if envelope is not None:
            e = e * envelope
        eu = e if phase is None else e * phase
        wf[a0:a1] = eu * Ut[a0:a1, None]
        wf[n + a0:n + a1] = e * Vt[a0:a1, None]
    return wf


def _base_params(dim, nx, ny, nz, dx, scheme=td.AB5AM5, precision=td.FP64, vext=None, dt_eF=0.005,
                 functional=td.SLDA, sclgth=0.0, dperp=None):
    eF = 0.5
    dp = dx if dperp is None else dperp
    return td.make_params(dim, nx, ny, nz, dx=dx, dy=dx if dim >= 2 else dp, dz=dx if dim == 3 else dp,
                          functional=functional, regularization=td.SPHERICAL_CUTOFF,
                          mf_correction=1, scheme=scheme, dt=dt_eF / eF, vext=vext, precision=precision,
                          sclgth=sclgth)


def uniform_slda(params, T=1.0e-9, kc=None, eps=1.0e-14):
    """Self-consistent uniform SLDA (or ASLDA, by params.functional) state at
    kF = 1 (n_a = n_b = n/2).  kc: momentum cutoff of the regularised
    coupling (default 0.999999 pi / max(DX, DY, DZ), pca_utils.c:171); the
    solution's energy cutoff is ec = alpha_+ kc^2 / 2 - mu_+
    (pca_uniform.h:256-262).  eps: convergence tolerance on n_a, n_b and
    Delta.  The td step recomputes the mean fields from the states' own
    densities, so a start is stationary only to the tolerance the solution
    was converged to: 1e-9 leaves a one-off energy offset of ~3e-10 after
    the first step, 1e-14 one of ~3e-15."""
    n0 = 1.0 / (3.0 * PI ** 2)
    if kc is None:
        kc = 0.999999 * PI / max(params.dx, params.dy, params.dz)
    u, rc = td.uniform_solve(params, n0 / 2, n0 / 2, kc, T, eps=eps, scmix=0.5, muchange=0.5, maxiter=10000)
    if rc != 0:
        raise td.TDWError("uniform SLDA solution did not converge")
    return u


# Momentum cutoffs at which the uniform solution's own state set (every state
# with -ec < E < ec, as create_uniform_wf keeps them, pca_uniform.h:737) has
# the size the workloads name.  The set grows by whole k-shells: on the 24^3
# lattice (dx = 1) the shell counts step 16212 -> 16308 -> 16500, so 16384
# is not the size of any such set and the 24^3 workload takes the 16308
# states of the shells |m|^2 <= 157 (kc in [3.2825, 3.2925]; the middle of the
# interval keeps every state away from the cutoff).  The 1D workload's 128
# states form such a set on a 4x4 transverse box with DY = DZ = 2.5 for kc in
# [1.400, 1.410].
KC_24CUBED_16308 = 3.2875
KC_1D_128 = 1.405


def config2(torch, precision=td.FP64, n_states=None, seed=1, noise=1.0e-6, device="cuda:0", scheme=td.AB5AM5,
            shape=(24, 24, 24), dx=1.0, functional=td.SLDA, sclgth=0.0, kc=None, first=None):
    """Plain 3D uniform SLDA: 24^3 periodic lattice, the self-consistent
    uniform solution's own state set -- every plane-wave BdG state with
    -ec < E < ec, in create_uniform_wf's order (pca_uniform.h:636-850) --
    plus seeded random-phase noise of amplitude `noise`.  Without noise the
    start is stationary: (H - <H>) psi = 0 for every state.

    kc: momentum cutoff; the default on the 24^3 lattice is
    KC_24CUBED_16308 (16308 states, Delta = 0.485 eF), elsewhere the
    reference's default.  n_states, if given, must equal the size of the
    set (an error otherwise).  first: keep only the first `first` states of
    the set (a slice for parity tests; not a stationary start).
    functional=ASLDA: the same with the ASLDA solution (effective mass
    alpha = 1.094 at n_a = n_b); functional=SLDAE: the SLDAE solution at
    scattering length sclgth (k_F = 1 units)."""
    nx, ny, nz = shape
    p = _base_params(3, nx, ny, nz, dx, scheme=scheme, precision=precision, functional=functional, sclgth=sclgth)
    if kc is None and shape == (24, 24, 24) and dx == 1.0 and functional == td.SLDA:
        kc = KC_24CUBED_16308
    u = uniform_slda(p, kc=kc)
    K, E, U, V = plane_wave_list(p, u.mu_a, u.mu_b, u.V_a, u.V_b, u.delta, u.ec, alph=u.alph_a)
    if len(E) != u.nwf:
        raise td.TDWError("plane-wave set (%d) differs from the solver's state count (%d)" % (len(E), u.nwf))
    if n_states is not None and n_states != len(E):
        raise td.TDWError("the uniform solution at kc = %.6f has %d states below its cutoff, not %d"
                          % (u.kc, len(E), n_states))
    if first is not None:
        K, E, U, V = K[:first], E[:first], U[:first], V[:first]
    wf = states_on_gpu(torch, p, K, U, V, device=device)
    if noise > 0:
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        vol = (nx * dx) * (ny * dx) * (nz * dx)
        th = torch.rand(wf.shape, generator=g, dtype=torch.float64, device=device) * (2 * PI)
        wf += torch.polar(torch.full_like(th, noise / math.sqrt(vol)), th)
        del th
    p.mu_a, p.mu_b, p.ec = u.mu_a, u.mu_b, u.ec
    N = nx * ny * nz
    pots = np.zeros(12 * N)
    pots[:N] = u.V_a
    pots[N:2 * N] = u.V_b
    pots[2 * N:4 * N:2] = u.delta
    pots[4 * N:5 * N] = u.alph_a
    pots[5 * N:6 * N] = u.alph_b
    fb = np.array([td.fbeta(e, u.beta) for e in E])
    eF = 0.5
    return dict(params=p, wf=wf, fbeta=fb, kt=None, cnt=None, pots=pots, uniform=u, eF=eF, E=E,
                name="3D uniform %s %dx%dx%d, %d states" % ({td.ASLDA: "ASLDA", td.SLDAE: "SLDAE"}.get(
                    functional, "SLDA"), nx, ny, nz, len(E)))


def select_states(wf, En, kt, cnt, n_states, dim):
    """The n_states lowest-|E| states of a plane-wave set (kept in their
    original order), with their transverse momenta / degeneracies."""
    nall = En.size
    order = np.sort(np.argsort(np.abs(En), kind="stable")[:n_states])
    wfs = np.concatenate([wf[order], wf[nall + order]])
    kts = None
    if kt is not None:
        kts = kt[order] if dim == 2 else np.concatenate([kt[order], kt[nall + order]])
    cnts = None if cnt is None else cnt[order]
    return wfs, En[order], kts, cnts


def uniform_start(dim, shape, dx=1.0, n_states=None, scheme=td.AB5AM5, dt_eF=0.005, T=1.0e-9, dt_perp=None,
                  kc=None):
    """Self-consistent uniform SLDA state at kF = 1 on a 1D/2D/3D lattice
    (transverse box from shape for dim < 3, spacing dt_perp, default dx) and
    its plane-wave states (all, or the n_states of lowest |E|), host arrays:
    params, wf [2n, N], pots (12N), fbeta, kt, cnt."""
    nx, ny, nz = shape
    p = _base_params(dim, nx, ny, nz, dx, scheme=scheme, dt_eF=dt_eF, dperp=dt_perp)
    u = uniform_slda(p, T=T, kc=kc)
    p.mu_a, p.mu_b, p.ec = u.mu_a, u.mu_b, u.ec
    wf, En, kt, cnt = td.uniform_states(p, u)
    if n_states is not None and n_states < En.size:
        wf, En, kt, cnt = select_states(wf, En, kt, cnt, n_states, dim)
    N = wf.shape[1]
    pots = np.zeros(12 * N)
    pots[:N] = u.V_a
    pots[N:2 * N] = u.V_b
    pots[2 * N:4 * N:2] = u.delta
    pots[4 * N:6 * N] = 1.0
    fb = np.array([td.fbeta(e, u.beta) for e in En])
    return dict(params=p, wf=wf, pots=pots, fbeta=fb, kt=kt, cnt=cnt, uniform=u, En=En)


def config4(n_states=4096, seed=4, shape=(64, 64, 8), dx=1.0):
    """2D SLDA on a 64x64 lattice (transverse box of 8 points), the 4096
    lowest-|E| plane-wave states of the uniform solution, two Gaussian beams
    with parameters (amplitude, x, y, width) drawn uniformly from the
    physical range (seeded), and the particle number inside a disk of radius
    10 dx as the observable."""
    s = uniform_start(2, shape, dx=dx, n_states=n_states)
    p = s["params"]
    eF = 0.5
    rng = np.random.default_rng(seed)
    L = shape[0] * dx
    theta = []
    for _ in range(2):
        theta += [rng.uniform(0.05, 0.2) * eF, rng.uniform(-0.25, 0.25) * L, rng.uniform(-0.25, 0.25) * L,
                  rng.uniform(2.0, 5.0) * dx]
    s["theta"] = np.array(theta)
    s["center"] = (0.0, 0.0)
    s["radius"] = 10.0 * dx
    return s


# Cutoff at which the 32^3 uniform solution's own state set (dx = 1) has
# 24586 states, the shell count nearest 24576 (kc in [2.815, 2.821]).
KC_32CUBED_24586 = 2.818


def config3(torch, precision=td.SPLIT, n_states=None, seed=3, device="cuda:0", shape=(32, 32, 32), dx=1.0,
            omega_eF=0.1, envelope_frac=0.3, noise=1.0e-6, first=None, kc=None):
    """3D SLDA in a harmonic trap with an imprinted vortex line: 32^3
    lattice, trap 0.5 omega^2 r^2 with omega = 0.1 eF switched on at t = 0,
    the uniform solution's own state set (every plane-wave BdG state with
    -ec < E < ec: 24586 states at kc = KC_32CUBED_24586, the shell count
    nearest 24576), the phase exp(i phi) of a vortex line along z imprinted
    on the u components (so Delta ~ exp(i phi)), and seeded random-phase
    noise whose amplitude follows a Gaussian envelope exp(-r^2 / (2 R^2))
    (R = envelope_frac * L).  The start is the uniform stationary state
    quenched by the trap and the vortex, so the run stays a physical
    evolution for thousands of steps.  n_states, if given, must equal the
    size of the set; first keeps only the first states (a slice for parity
    tests)."""
    nx, ny, nz = shape
    eF = 0.5
    vext = td.gauss_vext([], omega=omega_eF * eF)
    p = _base_params(3, nx, ny, nz, dx, precision=precision, vext=vext)
    if kc is None and shape == (32, 32, 32) and dx == 1.0:
        kc = KC_32CUBED_24586
    u = uniform_slda(p, kc=kc)
    K, E, U, V = plane_wave_list(p, u.mu_a, u.mu_b, u.V_a, u.V_b, u.delta, u.ec)
    if len(E) != u.nwf:
        raise td.TDWError("plane-wave set (%d) differs from the solver's state count (%d)" % (len(E), u.nwf))
    if n_states is not None and n_states != len(E):
        raise td.TDWError("the uniform solution at kc = %.6f has %d states below its cutoff, not %d"
                          % (u.kc, len(E), n_states))
    if first is not None:
        K, E, U, V = K[:first], E[:first], U[:first], V[:first]
    X = (torch.arange(nx, dtype=torch.float64, device=device) - nx // 2) * dx
    Y = (torch.arange(ny, dtype=torch.float64, device=device) - ny // 2) * dx
    Z = (torch.arange(nz, dtype=torch.float64, device=device) - nz // 2) * dx
    gx, gy, gz = torch.meshgrid(X, Y, Z, indexing="ij")
    R = envelope_frac * nx * dx
    env = torch.exp(-(gx ** 2 + gy ** 2 + gz ** 2) / (2 * R * R)).reshape(-1)
    # phase of a vortex line through the lattice centre, offset by half a
    # cell so that no lattice point sits on the core
    phi = torch.atan2(gy + 0.5 * dx, gx + 0.5 * dx).reshape(-1)
    phase = torch.polar(torch.ones_like(phi), phi)
    wf = states_on_gpu(torch, p, K, U, V, device=device, phase=phase)
    if noise > 0:
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        vol = (nx * dx) * (ny * dx) * (nz * dx)
        chunk = 2048
        for a0 in range(0, wf.shape[0], chunk):
            rows = wf[a0:a0 + chunk]
            th = torch.rand(rows.shape, generator=g, dtype=torch.float64, device=device) * (2 * PI)
            rows += torch.polar(torch.full_like(th, noise / math.sqrt(vol)), th) * env
            del th
    p.mu_a, p.mu_b, p.ec = u.mu_a, u.mu_b, u.ec
    N = nx * ny * nz
    pots = np.zeros(12 * N)
    pots[:N] = u.V_a
    pots[N:2 * N] = u.V_b
    pots[2 * N:4 * N:2] = u.delta
    pots[4 * N:6 * N] = 1.0
    fb = np.array([td.fbeta(e, u.beta) for e in E])
    return dict(params=p, wf=wf, fbeta=fb, kt=None, cnt=None, pots=pots, uniform=u, eF=eF, E=E,
                name="3D trapped SLDA with a vortex line %dx%dx%d, %d states" % (nx, ny, nz, len(E)))


# Cutoff at which the uniform solution's own state set on the 32 x 32 x 96
# tube (dx = 1) has 12062 states, the shell count nearest 12000
# (kc in [1.566, 1.569]).
KC_TUBE_12062 = 1.5675


def tube_phase(nz, dz, width):
    """Phase profile along the tube axis with a step of pi at the barrier
    (z = 0) and the step back at the box edge z = +-L/2, each a tanh of the
    given width, so that the profile is periodic: -pi/2 on the left half,
    +pi/2 on the right half."""
    L = nz * dz
    z = (np.arange(nz) - nz // 2) * dz
    return 0.5 * math.pi * (np.tanh(z / width) - np.tanh((z - L / 2) / width) - np.tanh((z + L / 2) / width))


def config5(torch, precision=td.SPLIT, n_states=None, seed=5, device="cuda:0", shape=(32, 32, 96), dx=1.0,
            barrier_eF=0.3, barrier_width_dx=3.0, phase_width_dx=2.0, noise=1.0e-6, first=None, kc=None):
    """3D elongated-tube Josephson junction: 32 x 32 x 96 lattice (tube axis
    z), a Gaussian barrier V_ext = V0 exp(-z^2 / (2 w^2)) across the tube with
    V0 = 0.3 eF and w = 3 dx, on from t = 0, and a phase step of pi imprinted
    on the u components across it (tube_phase), on the uniform solution's own
    state set (12062 states at kc = KC_TUBE_12062, the shell count nearest
    12000) with seeded random-phase noise.  n_states, if given, must equal
    the size of the set; first keeps the first states of the set (one GPU's
    share of the states, or a slice for parity tests)."""
    nx, ny, nz = shape
    eF = 0.5
    w = barrier_width_dx * dx
    vext = td.gauss_vext([(barrier_eF * eF, (0.0, 0.0, 0.0), (0.0, 0.0, -0.5 / (w * w)))])
    p = _base_params(3, nx, ny, nz, dx, precision=precision, vext=vext)
    if kc is None and shape == (32, 32, 96) and dx == 1.0:
        kc = KC_TUBE_12062
    u = uniform_slda(p, kc=kc)
    K, E, U, V = plane_wave_list(p, u.mu_a, u.mu_b, u.V_a, u.V_b, u.delta, u.ec)
    if len(E) != u.nwf:
        raise td.TDWError("plane-wave set (%d) differs from the solver's state count (%d)" % (len(E), u.nwf))
    if n_states is not None and n_states != len(E):
        raise td.TDWError("the uniform solution at kc = %.6f has %d states below its cutoff, not %d"
                          % (u.kc, len(E), n_states))
    if first is not None:
        K, E, U, V = K[:first], E[:first], U[:first], V[:first]
    phz = torch.as_tensor(tube_phase(nz, dx, phase_width_dx * dx), device=device)
    phase = torch.polar(torch.ones(nx * ny * nz, dtype=torch.float64, device=device),
                        phz.repeat(nx * ny))
    wf = states_on_gpu(torch, p, K, U, V, device=device, phase=phase)
    if noise > 0:
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        vol = (nx * dx) * (ny * dx) * (nz * dx)
        chunk = 2048
        for a0 in range(0, wf.shape[0], chunk):
            rows = wf[a0:a0 + chunk]
            th = torch.rand(rows.shape, generator=g, dtype=torch.float64, device=device) * (2 * PI)
            rows += torch.polar(torch.full_like(th, noise / math.sqrt(vol)), th)
            del th
    p.mu_a, p.mu_b, p.ec = u.mu_a, u.mu_b, u.ec
    N = nx * ny * nz
    pots = np.zeros(12 * N)
    pots[:N] = u.V_a
    pots[N:2 * N] = u.V_b
    pots[2 * N:4 * N:2] = u.delta
    pots[4 * N:6 * N] = 1.0
    fb = np.array([td.fbeta(e, u.beta) for e in E])
    return dict(params=p, wf=wf, fbeta=fb, kt=None, cnt=None, pots=pots, uniform=u, eF=eF, E=E,
                name="3D tube Josephson junction %dx%dx%d, %d states" % (nx, ny, nz, len(E)))


def config1(n_states=128, shape=(64, 4, 4), dx=1.0, dperp=2.5, kc=KC_1D_128, amp_eF=0.1, width_dx=4.0):
    """1D uniform SLDA at unitarity: 64-point periodic lattice with dx = 1
    (a 4x4 transverse box of spacing 2.5, which carries the kF = 1 density),
    the uniform solution's own state set -- all 128 plane-wave BdG states
    with -ec < E < ec at kc = KC_1D_128 (Delta = 0.47 eF) -- and a Gaussian
    V_ext of amplitude 0.1 eF and width 4 dx on from t = 0 (host arrays)."""
    s = uniform_start(1, shape, dx=dx, dt_perp=dperp, kc=kc)
    if s["En"].size != n_states:
        raise td.TDWError("the 1D uniform solution has %d states below its cutoff, not %d"
                          % (s["En"].size, n_states))
    p = s["params"]
    eF = 0.5
    w = -0.5 / (width_dx * dx) ** 2
    p.vext = td.gauss_vext([(amp_eF * eF, (0.0, 0.0, 0.0), (w, 0.0, 0.0))])
    return s
