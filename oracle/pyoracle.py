"""ctypes view of the CPU FP64 oracle (oracle/tdw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (wslda_amd/) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtdw_oracle.so")

SLDA, ASLDA, BDG, SLDAE = 111, 112, 113, 115
SPHERICAL, CUBIC = 88, 89
AB3AM4, AB4AM5, AB5AM5 = 34, 45, 55
EKIN, EPOT, EPAIR, ECURRENT, EPOTEXT, EPAIREXT, EVELEXT, NPARTA, NPARTB, LZA, LZB = range(11)
MAXG = 4


class Vext(C.Structure):
    _fields_ = [("ngauss", C.c_int),
                ("amp", C.c_double * MAXG), ("cx", C.c_double * MAXG),
                ("cy", C.c_double * MAXG), ("cz", C.c_double * MAXG),
                ("wx", C.c_double * MAXG), ("wy", C.c_double * MAXG),
                ("wz", C.c_double * MAXG),
                ("omega", C.c_double), ("switch_on", C.c_int),
                ("t_start", C.c_double), ("t_stop", C.c_double), ("t_switch", C.c_double),
                ("flip_b", C.c_int)]


class Cfg(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("functional", C.c_int), ("regularization", C.c_int),
                ("mf_correction", C.c_int), ("scheme", C.c_int),
                ("mu_a", C.c_double), ("mu_b", C.c_double), ("ec", C.c_double),
                ("dt", C.c_double), ("t0", C.c_double), ("sclgth", C.c_double),
                ("p_nmin", C.c_double), ("p_nmax", C.c_double), ("denseps", C.c_double),
                ("vext", Vext), ("spinsymmetry", C.c_int),
                ("kF", C.c_double), ("qfalpha", C.c_double), ("qfbeta", C.c_double), ("qfgamma", C.c_double),
                ("force_a1", C.c_int), ("precision", C.c_int),
                ("hkf_mode", C.c_int), ("hkf_mu", C.c_double), ("hkf_T", C.c_double)]


class Uniform(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("n0_a", "n0_b", "mu_a", "mu_b", "V_a", "V_b", "alph_a", "alph_b", "kc", "ec",
                 "nu", "delta", "tau_a", "tau_b", "beta", "S", "ekin", "epot", "epair")] + \
               [("nwf", C.c_int), ("iters", C.c_int)]


class State(C.Structure):
    _fields_ = [("n", C.c_int),
                ("wf", C.c_void_p), ("lap", C.c_void_p), ("ddx", C.c_void_p),
                ("ddy", C.c_void_p), ("ddz", C.c_void_p),
                ("f", C.c_void_p * 5),
                ("fbeta", C.c_void_p), ("kt", C.c_void_p), ("cnt", C.c_void_p),
                ("dens", C.c_void_p), ("pots", C.c_void_p),
                ("qpe_last", C.c_double * 1), ("qpe", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        cp = C.POINTER(Cfg)
        sig = {
            "tdwo_npts": (C.c_size_t, [cp]),
            "tdwo_fft": (None, [cp, P, I]),
            "tdwo_laplace": (None, [cp, I, P, P]),
            "tdwo_derivatives": (None, [cp, I, P, P, P, P, P]),
            "tdwo_laplace_real": (None, [cp, P, P]),
            "tdwo_gradient_real": (None, [cp, P, P, P, P]),
            "tdwo_dvec_real": (None, [cp, P, P, P, P, P, P]),
            "tdwo_div_real": (None, [cp, P, P, P, P]),
            "tdwo_hf_filter_c": (None, [cp, P, P, D, D]),
            "tdwo_hf_filter_d": (None, [cp, I, P, P, D, D]),
            "tdwo_overlap": (None, [cp, I, P, P, P, P]),
            "tdwo_qf_densities": (None, [cp, I, P, P, P, P, P, P]),
            "tdwo_qf_add": (None, [cp, P, P, D, D]),
            "tdwo_qf_remove": (None, [cp, P, P]),
            "tdwo_abm_step1": (None, [cp, I, P, P, P, P, P, P]),
            "tdwo_abm_step4": (None, [cp, I, P, P, P]),
            "tdwo_norms": (None, [cp, I, P, P]),
            "tdwo_normalize": (None, [cp, I, P]),
            "tdwo_densities": (None, [cp, I, P, P, P, P, P, P, P, P, P, I]),
            "tdwo_density_tau": (None, [cp, P]),
            "tdwo_vext_eval": (None, [cp, I, P, P]),
            "tdwo_potentials": (None, [cp, I, P, P]),
            "tdwo_apply_h": (None, [cp, I, P, P, P, P, P, P, P]),
            "tdwo_apply_h_full": (None, [cp, I, P, P, P, P, P, P, P, P, P, P, P]),
            "tdwo_alpha_wf": (None, [cp, I, P, P, P]),
            "tdwo_energy": (None, [cp, I, P, P, P]),
            "tdwo_step": (I, [cp, C.POINTER(State), I, I]),
            "tdwo_initial_measure": (None, [cp, C.POINTER(State)]),
            "tdwo_taylor": (None, [cp, I, D, I, P, P, P]),
            "tdwo_selfstart": (I, [cp, C.POINTER(State)]),
            "tdwo_uniform_solve": (I, [cp, D, D, D, D, D, D, D, D, D, I, I, C.POINTER(Uniform)]),
            "tdwo_uniform_count": (I, [cp, C.POINTER(Uniform)]),
            "tdwo_uniform_wf": (I, [cp, C.POINTER(Uniform), I, I, P, P, P, P]),
            "tdwo_fbeta": (D, [D, D]),
            "tdwo_set_threads": (None, [I]),
            "tdwo_get_threads": (I, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_cfg(dim, nx, ny, nz, dx=1.0, dy=None, dz=None, functional=SLDA,
             regularization=SPHERICAL, mf_correction=0, scheme=AB4AM5,
             mu_a=0.0, mu_b=0.0, ec=0.0, dt=0.01, t0=0.0, sclgth=0.0,
             p_nmin=1.0e-7, p_nmax=1.0e-5, denseps=1.0e-8, vext=None, spinsymmetry=0, kF=1.0,
             qf=(0.0, 0.0, 0.0), force_a1=0, hkf=(0, 0.0, 0.0)):
    c = Cfg()
    c.dim, c.nx, c.ny, c.nz = dim, nx, ny, nz
    c.dx = dx
    c.dy = dx if dy is None else dy
    c.dz = dx if dz is None else dz
    c.functional, c.regularization, c.mf_correction, c.scheme = functional, regularization, mf_correction, scheme
    c.mu_a, c.mu_b, c.ec, c.dt, c.t0, c.sclgth = mu_a, mu_b, ec, dt, t0, sclgth
    c.p_nmin, c.p_nmax, c.denseps = p_nmin, p_nmax, denseps
    if vext is not None:
        c.vext = vext
    c.spinsymmetry = spinsymmetry
    c.kF = kF
    c.qfalpha, c.qfbeta, c.qfgamma = qf
    c.force_a1 = force_a1
    c.hkf_mode, c.hkf_mu, c.hkf_T = hkf
    return c


def gauss_vext(gaussians, omega=0.0, switch=None, flip_b=0):
    """gaussians: list of (amp, (cx,cy,cz), (wx,wy,wz)) with exp(w*(x-c)^2)."""
    v = Vext()
    v.ngauss = len(gaussians)
    for k, (a, cc, ww) in enumerate(gaussians):
        v.amp[k] = a
        v.cx[k], v.cy[k], v.cz[k] = cc
        v.wx[k], v.wy[k], v.wz[k] = ww
    v.omega = omega
    if switch is not None:
        v.switch_on = 1
        v.t_start, v.t_stop, v.t_switch = switch
    v.flip_b = flip_b
    return v


def npts(cfg):
    return int(lib().tdwo_npts(C.byref(cfg)))


class Problem:
    """All arrays of one td run, laid out as the reference keeps them."""

    def __init__(self, cfg, wf, fbeta, kt=None, cnt=None, pots=None):
        self.cfg = cfg
        N = npts(cfg)
        self.N = N
        n = wf.shape[0] // 2 if wf.ndim == 2 else wf.size // (2 * N)
        self.n = n
        # own copies: the oracle updates its arrays in place, and callers
        # start GPU runs from the same inputs
        self.wf = np.array(wf, dtype=np.complex128, copy=True).reshape(2 * n, N)
        self.lap = np.zeros_like(self.wf)
        self.ddx = np.zeros_like(self.wf)
        self.ddy = np.zeros_like(self.wf)
        self.ddz = np.zeros_like(self.wf)
        self.f = [np.zeros_like(self.wf) for _ in range(5)]
        self.fbeta = np.array(fbeta, dtype=np.float64, copy=True)
        self.kt = None if kt is None else np.array(kt, dtype=np.float64, copy=True)
        self.cnt = None if cnt is None else np.array(cnt, dtype=np.int32, copy=True)
        self.dens = np.zeros(12 * N)
        if pots is None:
            pots = np.zeros(12 * N)
            pots[4 * N:6 * N] = 1.0
        self.pots = np.array(pots, dtype=np.float64, copy=True)
        self.qpe = np.zeros(n)
        self.st = State()
        self.st.n = n
        self._sync()

    def _sync(self):
        s = self.st
        s.wf, s.lap = self.wf.ctypes.data, self.lap.ctypes.data
        s.ddx, s.ddy, s.ddz = self.ddx.ctypes.data, self.ddy.ctypes.data, self.ddz.ctypes.data
        for k in range(5):
            s.f[k] = self.f[k].ctypes.data
        s.fbeta = self.fbeta.ctypes.data
        s.kt = None if self.kt is None else self.kt.ctypes.data
        s.cnt = None if self.cnt is None else self.cnt.ctypes.data
        s.dens, s.pots, s.qpe = self.dens.ctypes.data, self.pots.ctypes.data, self.qpe.ctypes.data

    def history(self):
        """History arrays in the current rotation order (fkm1..fkm5)."""
        addr = {a.ctypes.data: a for a in self.f}
        return [addr[self.st.f[k]] for k in range(5)]

    def initial_measure(self):
        lib().tdwo_initial_measure(C.byref(self.cfg), C.byref(self.st))

    def step(self, it, grad_last=0):
        return lib().tdwo_step(C.byref(self.cfg), C.byref(self.st), it, grad_last)

    def selfstart(self):
        """Taylor self-start of pca.c:1086-1392 (history filled, state
        advanced by 3 / 4 / 5 steps, densities and potentials of the new
        state)."""
        return lib().tdwo_selfstart(C.byref(self.cfg), C.byref(self.st))

    def energy(self, it):
        e = np.zeros(11)
        lib().tdwo_energy(C.byref(self.cfg), it, ptr(self.dens), ptr(self.pots), ptr(e))
        return e


def uniform_solve(cfg, n0_a, n0_b, kc, T, eps=1e-9, scmix=0.5, muchange=0.5, maxiter=10000,
                  Tstart=None, dT=0.01, mu_rule=0):
    u = Uniform()
    rc = lib().tdwo_uniform_solve(C.byref(cfg), n0_a, n0_b, kc, T if Tstart is None else Tstart, T, dT,
                                  eps, scmix, muchange, maxiter, mu_rule, C.byref(u))
    return u, rc


def uniform_states(cfg, u):
    L = lib()
    n = L.tdwo_uniform_count(C.byref(cfg), C.byref(u))
    N = npts(cfg)
    wf = np.zeros((2 * n, N), dtype=np.complex128)
    En = np.zeros(n)
    kt = np.zeros(2 * n)
    cnt = np.zeros(n, dtype=np.int32)
    L.tdwo_uniform_wf(C.byref(cfg), C.byref(u), 0, n, ptr(wf), ptr(En), ptr(kt), ptr(cnt))
    if cfg.dim == 3:
        kt, cnt = None, None
    elif cfg.dim == 2:
        kt, cnt = kt[:n].copy(), None
    return wf, En, kt, cnt


def fbeta(E, beta):
    return lib().tdwo_fbeta(float(E), float(beta))
