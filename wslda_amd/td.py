"""Python view of libwslda_td.so (the MI355X td-W-SLDA step).

The library is plain C ABI (include/wslda_td.h); this module binds it with
ctypes and uses torch tensors only as device memory.  Names and argument
meaning follow the reference's kernel interface (hpc-engine/pca_kernels.h,
pca_derivative.h, pca_dens.h), so a run reads like the loop of pca.c.

Nothing here falls back to the CPU: if the library or a HIP device is
missing, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libwslda_td.so")

SLDA, ASLDA, BDG, SLDAE = 111, 112, 113, 115
SPHERICAL_CUTOFF, CUBIC_CUTOFF = 88, 89
AB3AM4, AB4AM5, AB5AM5 = 34, 45, 55
FP64, SPLIT = 0, 1
TDW_EINVAL, TDW_ENOCTX, TDW_ESIZE, TDW_ENOMEM = -1001, -1002, -1003, -1004
TDW_ESCRATCH = -1007
NOALLOC_SCRATCH_DOUBLES = 1 << 20
EKIN, EPOT, EPAIR, ECURRENT, EPOTEXT, EPAIREXT, EVELEXT, NPARTA, NPARTB, LZA, LZB = range(11)
NITEMS = 11
MAX_GAUSS = 4
MAX_USER_PARAMS = 32


class TDWError(RuntimeError):
    pass


class Vext(C.Structure):
    _fields_ = [("ngauss", C.c_int),
                ("amp", C.c_double * MAX_GAUSS), ("cx", C.c_double * MAX_GAUSS),
                ("cy", C.c_double * MAX_GAUSS), ("cz", C.c_double * MAX_GAUSS),
                ("wx", C.c_double * MAX_GAUSS), ("wy", C.c_double * MAX_GAUSS),
                ("wz", C.c_double * MAX_GAUSS),
                ("omega", C.c_double), ("switch_on", C.c_int),
                ("t_start", C.c_double), ("t_stop", C.c_double), ("t_switch", C.c_double),
                ("flip_b", C.c_int)]


class Params(C.Structure):
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


class KMode(C.Structure):
    """tdw_kmode: a transverse mode (ky, kz) and its degeneracy."""
    _fields_ = [("ky", C.c_double), ("kz", C.c_double), ("weight", C.c_int)]


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
                ("dens", C.c_void_p), ("pots", C.c_void_p), ("qpe", C.c_void_p)]


class CheckpointScalars(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("mu_a", "mu_b", "ec", "kF", "eF", "Effg", "beta")]


class StInfo(C.Structure):
    """tdw_st_info: the run constants of a static-solver (st-wslda-3d) output."""
    _fields_ = [("nwf", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + \
               [(k, C.c_double) for k in ("dx", "dy", "dz", "kF", "mu_a", "mu_b", "ec", "beta")] + \
               [("new_format", C.c_int)]


class TapeStep(C.Structure):
    _fields_ = [("n", C.c_int), ("scal", C.c_void_p), ("dens_p", C.c_void_p), ("pots_p", C.c_void_p),
                ("dens_c", C.c_void_p), ("pots_c", C.c_void_p)]


class AdjState(C.Structure):
    _fields_ = [("n", C.c_int), ("ly", C.c_void_p), ("lf", C.c_void_p * 5), ("lpots", C.c_void_p),
                ("w", C.c_void_p * 5), ("lgrid", C.c_void_p), ("theta_grad", C.c_void_p)]


class WDataParams(C.Structure):
    _fields_ = [("outprefix", C.c_char_p), ("writevar", C.c_char_p), ("writeprec", C.c_char),
                ("spinsymmetry", C.c_int), ("overwrite", C.c_int), ("inittype", C.c_int),
                ("t0", C.c_double), ("dt", C.c_double), ("codename", C.c_char_p)]


def make_params(dim, nx, ny, nz, dx=1.0, dy=None, dz=None, functional=SLDA,
                regularization=SPHERICAL_CUTOFF, mf_correction=0, scheme=AB4AM5,
                mu_a=0.0, mu_b=0.0, ec=0.0, dt=0.01, t0=0.0, sclgth=0.0,
                p_nmin=1.0e-7, p_nmax=1.0e-5, denseps=1.0e-8, vext=None, precision=FP64, spinsymmetry=0,
                kF=1.0, qf=(0.0, 0.0, 0.0)):
    p = Params()
    p.dim, p.nx, p.ny, p.nz = dim, nx, ny, nz
    p.dx = dx
    p.dy = dx if dy is None else dy
    p.dz = dx if dz is None else dz
    p.functional, p.regularization, p.mf_correction, p.scheme = functional, regularization, mf_correction, scheme
    p.mu_a, p.mu_b, p.ec, p.dt, p.t0, p.sclgth = mu_a, mu_b, ec, dt, t0, sclgth
    p.p_nmin, p.p_nmax, p.denseps = p_nmin, p_nmax, denseps
    if vext is not None:
        p.vext = vext
    p.spinsymmetry = spinsymmetry
    p.kF = kF
    p.qfalpha, p.qfbeta, p.qfgamma = qf
    p.precision = precision
    return p


def gauss_vext(gaussians, omega=0.0, switch=None, flip_b=0):
    """gaussians: list of (amp, (cx, cy, cz), (wx, wy, wz)): amp*exp(w*(x-c)^2)."""
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


_libs = {}


def _sigs():
    P, I, D, Z = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    return {
        "tdw_create": (I, [C.POINTER(Params), I, I, C.POINTER(C.c_void_p)]),
        "tdw_destroy": (I, [P]),
        "tdw_sync": (I, [P]),
        "tdw_npts": (Z, [P]),
        "tdw_set_params": (I, [P, C.POINTER(Params)]),
        "tdw_set_user_params": (I, [P, P]),
        "tdw_get_user_params": (I, [P, P]),
        "tdw_set_states": (I, [P, P, P, P]),
        "tdw_abm_step1": (I, [P, I, P, P, P, P, P, P]),
        "tdw_abm_step4": (I, [P, I, P, P, P]),
        "tdw_normalize": (I, [P, I, P, P]),
        "tdw_normalize_noalloc": (I, [P, I, P, P]),
        "tdw_laplace": (I, [P, I, P, P]),
        "tdw_derivatives": (I, [P, I, P, P, P, P, P]),
        "tdw_laplace_real": (I, [P, P, P]),
        "tdw_densities": (I, [P, I, P, P, P, P, P, P, P, I]),
        "tdw_density_tau": (I, [P, P]),
        "tdw_potentials": (I, [P, I, P, P]),
        "tdw_apply_h": (I, [P, I, I, P, P, P, P, P, P]),
        "tdw_apply_h_full": (I, [P, I, I, P, P, P, P, P, P, P, P, P, P]),
        "tdw_alpha_wf": (I, [P, I, P, P, P]),
        "tdw_overlap": (I, [P, I, P, P, P, P]),
        "tdw_energy": (I, [P, I, P, P, P]),
        "tdw_vext_eval": (I, [P, I, P, P]),
        "tdw_step": (I, [P, C.POINTER(State), I, I]),
        "tdw_step_fused": (I, [P, C.POINTER(State), I, I]),
        "tdw_finalize": (I, [P, C.POINTER(State)]),
        "tdw_fused_supported": (I, [P]),
        "tdw_initial_measure": (I, [P, C.POINTER(State)]),
        "tdw_selfstart": (I, [P, C.POINTER(State), P, P]),
        "tdw_adj_forward_step": (I, [P, C.POINTER(State), I, P, P, C.POINTER(TapeStep)]),
        "tdw_adj_forward_step_to": (I, [P, C.POINTER(State), I, P, P, C.POINTER(TapeStep)]),
        "tdw_adj_history_entry": (I, [P, I, C.POINTER(TapeStep), P, P, P]),
        "tdw_checkpoint_nwf": (I, [C.c_char_p, C.POINTER(C.c_int)]),
        "tdw_checkpoint_save": (I, [P, C.c_char_p, C.POINTER(State), D, P, P, C.POINTER(CheckpointScalars)]),
        "tdw_checkpoint_load": (I, [P, C.c_char_p, C.POINTER(State), C.POINTER(C.c_double), P, P,
                                    C.POINTER(CheckpointScalars)]),
        "tdw_wdata_create": (I, [P, C.POINTER(WDataParams), C.POINTER(C.c_void_p)]),
        "tdw_wdata_setconst": (I, [P, C.c_char_p, D, C.c_char_p]),
        "tdw_wdata_begin": (I, [P]),
        "tdw_wdata_write": (I, [P, I, P, P]),
        "tdw_wdata_cycles": (I, [P]),
        "tdw_wdata_destroy": (I, [P]),
        "tdw_adj_backward_step": (I, [P, C.POINTER(AdjState), I, P, C.POINTER(C.c_void_p), P, P,
                                      C.POINTER(TapeStep), P, P]),
        "tdw_uniform_solve": (I, [C.POINTER(Params), D, D, D, D, D, D, D, D, D, I, I, C.POINTER(Uniform)]),
        "tdw_uniform_count": (I, [C.POINTER(Params), C.POINTER(Uniform)]),
        "tdw_uniform_states": (I, [C.POINTER(Params), C.POINTER(Uniform), I, I, P, P, P, P]),
        "tdw_fbeta": (D, [D, D]),
        "tdw_stream": (P, [P]),
        "tdw_timing": (I, [P, I]),
        "tdw_timing_report": (I, [P, I, P, P, P]),
        "tdw_bind": (I, [P]),
        "tdw_bound": (P, []),
        "tdw_project_info": (I, [C.POINTER(Params), C.POINTER(C.c_uint)]),
        # drop-in layer (reference names)
        "set_gpu": (I, [I]),
        "gpu_malloc": (I, [Z, C.POINTER(C.c_void_p)]),
        "gpu_free": (I, [P]),
        "host_malloc_pl": (I, [Z, C.POINTER(C.c_void_p)]),
        "host_free_pl": (I, [P]),
        "memcopy_host2gpu": (I, [P, P, Z]),
        "memcopy_gpu2host": (I, [P, P, Z]),
        "memcopy_gpu2gpu": (I, [P, P, Z]),
        "memcopy_const": (I, [D, D, D, D, D, D]),
        "memcopy_const_params": (I, [P]),
        "memcopy_const_sclgth": (I, [D]),
        "create_cufftPlans": (I, [I, I, C.POINTER(C.c_size_t)]),
        "set_workspace_for_cufftPlan": (I, [P]),
        "compute_derivatives": (I, [I, P, P, P, P, P, I]),
        "compute_laplace": (I, [I, P, P, I]),
        "compute_laplace_real_f": (I, [P, P, I]),
        "calculate_densities": (I, [I, P, P, P, P, P, P, P, P, I, I]),
        "calculate_densities_1d": (I, [I, P, P, P, P, P, P, P, P, I, I]),
        "calculate_densities_2d": (I, [I, P, P, P, P, P, P, P, P, I, I]),
        "density_caculate_tau": (I, [P, I]),
        "compute_potentials": (I, [I, P, P, D, I]),
        "compute_energy": (I, [I, P, P, P, I]),
        "apply_hamiltonian": (I, [I, I, P, P, P, P, P, P, P, P, P, D, P, D, I]),
        "apply_hamiltonian_1d": (I, [I, I, P, P, P, P, P, P, P, P, D, P, D, I]),
        "apply_hamiltonian_2d": (I, [I, I, P, P, P, P, P, P, P, P, P, D, P, D, I]),
        "multiply_wf_by_alpha": (I, [I, P, P, P, I]),
        "normalize_wf": (I, [I, P, I]),
        "abm_step1": (I, [I, P, P, P, P, P, P, I, I]),
        "abm_step4": (I, [I, P, P, P, I, I]),
        "taylor_expansion_contribution": (I, [I, D, I, P, P, P, I]),
        "get_v_ext": (I, [I, I, I, P]),
        "get_delta_ext": (I, [I, I, P, P]),
        "get_velocity_ext": (I, [I, I, I, P]),
        "symmetrize_densities_device": (I, [P]),
        "calculate_quantum_friction_densities": (I, [I, P, P, P, P, P, I]),
        "calculate_quantum_friction_densities_1d": (I, [I, P, P, P, P, P, P, P, I]),
        "calculate_quantum_friction_densities_2d": (I, [I, P, P, P, P, P, P, I]),
        "compute_gradient_real_f": (I, [P, P, P, P, I]),
        "compute_derivative_real_vector_f": (I, [P, P, P, P, P, P, I]),
        "compute_divergence_real_vector_f": (I, [P, P, P, P, I]),
        "high_frequency_filter_d": (I, [P, P, D, D, I]),
        "high_frequency_filter_c": (I, [P, P, D, D, I]),
        "high_frequency_filter_massive_d": (I, [I, P, P, D, D, I]),
        "compute_ovelap": (I, [I, P, P, P, P, P, I]),
        "local_reductionR": (I, [P, I, P, I, I]),
        "memcopy_extra_data": (I, [C.c_size_t, P]),
        "mm_get_size_of_kernels_workspace": (C.c_size_t, [C.c_size_t]),
        "mm_get_size_of_derivatives_workspace": (C.c_size_t, [C.c_size_t]),
        "mm_get_size_of_total_workspace": (C.c_size_t, [C.c_size_t]),
        "mm_get_pointer_to_kernels_workspace": (P, [C.c_size_t]),
        "mm_get_pointer_to_derivatives_workspace": (P, [C.c_size_t]),
        "mm_get_pointer_to_total_workspace": (P, [C.c_size_t]),
        "mm_get_pointer_to_fftONE_workspace": (P, [C.c_size_t]),
        "tdw_st3d_info": (I, [C.c_char_p, I, C.POINTER(StInfo), P]),
        "tdw_st3d_load": (I, [C.POINTER(Params), C.c_char_p, I, P, C.POINTER(StInfo), I, I, I, P, P, P,
                              C.POINTER(CheckpointScalars)]),
        "tdw_st1d_kmodes": (I, [C.POINTER(Params), P, I, P]),
        "scan_stwslda1d_info_files": (I, [C.c_char_p, I, I, P, P, P]),
        "read_stwslda1d_wf": (I, [C.c_char_p, I, I, P, P, I, I, P, P, P, P]),
        "check_comptibility": (I, [C.c_char_p, I, C.c_char_p]),
        "read_checkpoint_info_pca": (I, [C.c_char_p, P, P, P, P, P, P, P, P, P, P, P]),
        "scan_s3dpca_info_files": (I, [C.c_char_p, I, P, P]),
        "read_s3dpca_wf": (I, [C.c_char_p, I, P, I, I, P, P]),
    }


EXPORTS = sorted(_sigs().keys())


def lib(path=None):
    """Load the HIP library (default: the engine's own libwslda_td.so; a
    project build's library by its path); raises if it was not built."""
    path = os.path.abspath(path or LIBPATH)
    L = _libs.get(path)
    if L is None:
        if not os.path.exists(path):
            raise TDWError("%s not built: run `python -m wslda_amd.build` (a project: "
                           "`python -m wslda_amd.build --project DIR --codedim D`)" % path)
        L = C.CDLL(path)
        for name, (res, args) in _sigs().items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[path] = L
    return L


def project_info(path):
    """(Params, flags) a project library was compiled with: lattice,
    functional, schemes and stabilisation constants of its predefines.h, and
    the TDW_PROJ_* switches; None for the plain library."""
    p = Params()
    fl = C.c_uint(0)
    if lib(path).tdw_project_info(C.byref(p), C.byref(fl)) != 1:
        return None
    return p, fl.value


def check(rc, what=""):
    if rc != 0:
        raise TDWError("%s failed with code %d" % (what, rc))


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


class TDContext:
    """One td run's library context (lattice, functional, constants)."""

    def __init__(self, params, nwf, device=0, libpath=None):
        import torch
        if not torch.cuda.is_available():
            raise TDWError("no HIP device: the td step runs only on the GPU")
        self.L = lib(libpath)
        self.torch = torch
        self.params = params
        self.nwf = nwf
        self.device = device
        h = C.c_void_p()
        check(self.L.tdw_create(C.byref(params), nwf, device, C.byref(h)), "tdw_create")
        self.h = h
        self.N = int(self.L.tdw_npts(h))

    def close(self):
        if getattr(self, "h", None):
            self.L.tdw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.L.tdw_sync(self.h), "tdw_sync")

    def set_params(self, params):
        self.params = params
        check(self.L.tdw_set_params(self.h, C.byref(params)), "tdw_set_params")

    def set_states(self, fbeta, kt=None, cnt=None):
        fb = np.ascontiguousarray(fbeta, dtype=np.float64)
        kt_ = None if kt is None else np.ascontiguousarray(kt, dtype=np.float64)
        cnt_ = None if cnt is None else np.ascontiguousarray(cnt, dtype=np.int32)
        check(self.L.tdw_set_states(self.h, fb.ctypes.data_as(C.c_void_p),
                                   None if kt_ is None else kt_.ctypes.data_as(C.c_void_p),
                                   None if cnt_ is None else cnt_.ctypes.data_as(C.c_void_p)),
              "tdw_set_states")

    # -- building blocks (device tensors) ---------------------------------
    def _pre(self):
        # torch queues on its own stream: order it before the library's
        self.torch.cuda.synchronize(self.device)

    def abm_step1(self, y, f):
        self._pre()
        fs = [_ptr(x) for x in f] + [None] * (5 - len(f))
        check(self.L.tdw_abm_step1(self.h, self.nwf_of(y), _ptr(y), *fs), "abm_step1")

    def abm_step4(self, hy, y, flast):
        self._pre()
        check(self.L.tdw_abm_step4(self.h, self.nwf_of(y), _ptr(hy), _ptr(y), _ptr(flast)), "abm_step4")

    def normalize_wf(self, wf, norms=None):
        self._pre()
        check(self.L.tdw_normalize(self.h, self.nwf_of(wf), _ptr(wf), _ptr(norms)), "normalize_wf")

    def compute_laplace(self, inp, out):
        self._pre()
        m = inp.numel() // self.N
        check(self.L.tdw_laplace(self.h, m, _ptr(inp), _ptr(out)), "compute_laplace")

    def compute_derivatives(self, inp, ddx, ddy, ddz, lap):
        self._pre()
        m = inp.numel() // self.N
        check(self.L.tdw_derivatives(self.h, m, _ptr(inp), _ptr(ddx), _ptr(ddy), _ptr(ddz), _ptr(lap)),
              "compute_derivatives")

    def compute_laplace_real(self, f, out):
        self._pre()
        check(self.L.tdw_laplace_real(self.h, _ptr(f), _ptr(out)), "compute_laplace_real_f")

    def calculate_densities(self, wf, ddx, ddy, ddz, lap, dens, gradients, weights=None):
        self._pre()
        check(self.L.tdw_densities(self.h, self.nwf_of(wf), _ptr(wf), _ptr(ddx), _ptr(ddy), _ptr(ddz),
                                  _ptr(lap), _ptr(weights), _ptr(dens), int(gradients)),
              "calculate_densities")

    def density_caculate_tau(self, dens):
        self._pre()
        check(self.L.tdw_density_tau(self.h, _ptr(dens)), "density_caculate_tau")

    def compute_potentials(self, it, dens, pots):
        self._pre()
        check(self.L.tdw_potentials(self.h, it, _ptr(dens), _ptr(pots)), "compute_potentials")

    def apply_hamiltonian(self, it, wf_in, wf_out, lap, pots, useqpe=None, qpe=None):
        self._pre()
        check(self.L.tdw_apply_h(self.h, it, self.nwf_of(wf_in), _ptr(wf_in), _ptr(wf_out), _ptr(lap),
                                _ptr(pots), _ptr(useqpe), _ptr(qpe)), "apply_hamiltonian")

    def apply_hamiltonian_full(self, it, wf_in, wf_out, ddx, ddy, ddz, lap, alap, pots, useqpe=None, qpe=None):
        """H application with the gradients and lap(alpha psi) (alap None:
        formed by the library); the form ASLDA needs."""
        self._pre()
        check(self.L.tdw_apply_h_full(self.h, it, self.nwf_of(wf_in), _ptr(wf_in), _ptr(wf_out), _ptr(ddx),
                                     _ptr(ddy), _ptr(ddz), _ptr(lap), _ptr(alap), _ptr(pots), _ptr(useqpe),
                                     _ptr(qpe)), "apply_hamiltonian_full")

    def multiply_wf_by_alpha(self, wf_in, wf_out, pots):
        self._pre()
        check(self.L.tdw_alpha_wf(self.h, self.nwf_of(wf_in), _ptr(wf_in), _ptr(pots), _ptr(wf_out)),
              "multiply_wf_by_alpha")

    def compute_energy(self, it, dens, pots):
        self._pre()
        e = np.zeros(NITEMS)
        check(self.L.tdw_energy(self.h, it, _ptr(dens), _ptr(pots), e.ctypes.data_as(C.c_void_p)),
              "compute_energy")
        return e

    def stream_ptr(self):
        return int(self.L.tdw_stream(self.h))

    def timing(self, enable):
        check(self.L.tdw_timing(self.h, int(enable)), "tdw_timing")

    def timing_report(self):
        """{kernel name: (total ms, launches)} since the last report."""
        mx = 64
        names = C.create_string_buffer(32 * mx)
        ms = np.zeros(mx)
        cnt = np.zeros(mx, dtype=np.int32)
        k = self.L.tdw_timing_report(self.h, mx, names, ms.ctypes.data_as(C.c_void_p),
                                    cnt.ctypes.data_as(C.c_void_p))
        if k < 0:
            raise TDWError("tdw_timing_report failed with code %d" % k)
        raw = names.raw
        out = {}
        for i in range(k):
            nm = raw[32 * i:32 * i + 32].split(b"\0")[0].decode()
            out[nm] = (float(ms[i]), int(cnt[i]))
        return out

    def vext(self, it):
        self._pre()
        va = self.torch.empty(self.N, dtype=self.torch.float64, device="cuda:%d" % self.device)
        vb = self.torch.empty_like(va)
        check(self.L.tdw_vext_eval(self.h, it, _ptr(va), _ptr(vb)), "tdw_vext_eval")
        self.sync()
        return va, vb

    def nwf_of(self, wf):
        return wf.numel() // (2 * self.N)


def uniform_solve(params, n0_a, n0_b, kc, T, eps=1e-9, scmix=0.5, muchange=0.5, maxiter=10000,
                  Tstart=None, dT=0.01, mu_rule=0):
    """solve_uniform_problem of pca_uniform.h (SLDA, ASLDA; the BdG twin for
    BDG); returns (Uniform, rc)."""
    u = Uniform()
    rc = lib().tdw_uniform_solve(C.byref(params), n0_a, n0_b, kc, T if Tstart is None else Tstart, T, dT,
                                 eps, scmix, muchange, maxiter, mu_rule, C.byref(u))
    if rc < 0:
        raise TDWError("tdw_uniform_solve failed with code %d" % rc)
    return u, rc


def st3d_info(prefix, iogroups):
    """Run constants and states per file of a static-solver 3D output
    (tdw_st3d_info); returns (StInfo, nwf_per_file)."""
    info = StInfo()
    per = np.zeros(iogroups, dtype=np.int32)
    rc = lib().tdw_st3d_info(prefix.encode(), iogroups, C.byref(info), per.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise TDWError("tdw_st3d_info failed with code %d" % rc)
    return info, per


def st3d_load(params, prefix, iogroups, info, per, lidx=0, uidx=None, spinsymmetry=0, potentials=True):
    """States [lidx, uidx) of a static-solver 3D output (tdw_st3d_load):
    returns wf [2n, N] (complex128, u rows then v rows), their occupations
    f(E), the 12*N potentials (or None) and the run scalars."""
    uidx = info.nwf if uidx is None else uidx
    n = uidx - lidx
    N = params.nx * params.ny * params.nz
    wf = np.zeros((2 * n, N), dtype=np.complex128)
    fb = np.zeros(n)
    pots = np.zeros(12 * N) if potentials else None
    sc = CheckpointScalars()
    rc = lib().tdw_st3d_load(C.byref(params), prefix.encode(), iogroups, per.ctypes.data_as(C.c_void_p),
                             C.byref(info), lidx, uidx, spinsymmetry, wf.ctypes.data_as(C.c_void_p),
                             fb.ctypes.data_as(C.c_void_p),
                             pots.ctypes.data_as(C.c_void_p) if pots is not None else None, C.byref(sc))
    if rc != 0:
        raise TDWError("tdw_st3d_load failed with code %d" % rc)
    return wf, fb, pots, sc


def uniform_states(params, u):
    """All plane-wave BdG states of a uniform solution, as pca_uniform.h's
    create_uniform_wf orders them: (wf [2n, N] complex128, E_n, kt, cnt)."""
    n = lib().tdw_uniform_count(C.byref(params), C.byref(u))
    N = params.nx if params.dim == 1 else params.nx * params.ny if params.dim == 2 else params.nx * params.ny * params.nz
    wf = np.zeros((2 * n, N), dtype=np.complex128)
    En = np.zeros(n)
    kt = np.zeros(2 * n)
    cnt = np.zeros(n, dtype=np.int32)
    check(lib().tdw_uniform_states(C.byref(params), C.byref(u), 0, n, wf.ctypes.data_as(C.c_void_p),
                                   En.ctypes.data_as(C.c_void_p), kt.ctypes.data_as(C.c_void_p),
                                   cnt.ctypes.data_as(C.c_void_p)), "tdw_uniform_states")
    if params.dim == 3:
        return wf, En, None, None
    if params.dim == 2:
        return wf, En, kt[:n].copy(), None
    return wf, En, kt, cnt


def fbeta(E, beta):
    return lib().tdw_fbeta(float(E), float(beta))


class TDRun:
    """All device arrays of one td run, in the reference's layout, and the
    time loop of pca.c (initial measurement, PECE steps, energies)."""

    def __init__(self, ctx, wf, pots=None, fused=False):
        torch = ctx.torch
        self.ctx = ctx
        dev = "cuda:%d" % ctx.device
        N, n = ctx.N, ctx.nwf
        self.n = n
        if isinstance(wf, np.ndarray):
            self.wf = torch.as_tensor(np.ascontiguousarray(wf).reshape(2 * n, N), dtype=torch.complex128).to(dev)
        else:
            self.wf = wf.reshape(2 * n, N).to(dev).contiguous()
        z = lambda: torch.zeros((2 * n, N), dtype=torch.complex128, device=dev)
        self.lap = z()
        self.ddx = z()
        self.ddy = z() if ctx.params.dim >= 2 else None
        self.ddz = z() if ctx.params.dim >= 3 else None
        # split precision: the history entries hold FP32 values
        self.split = ctx.params.precision == SPLIT
        if self.split:
            self.f = [torch.zeros((2 * n, N), dtype=torch.complex64, device=dev) for _ in range(5)]
        else:
            self.f = [z() for _ in range(5)]
        self.dens = torch.zeros(12 * N, dtype=torch.float64, device=dev)
        if pots is None:
            pots = np.zeros(12 * N)
            pots[4 * N:6 * N] = 1.0
        self.pots = torch.as_tensor(np.ascontiguousarray(pots, dtype=np.float64)).to(dev)
        self.qpe = torch.zeros(n, dtype=torch.float64, device=dev)
        self.fused = fused
        self.st = State()
        self.st.n = n
        self.st.wf, self.st.lap = self.wf.data_ptr(), self.lap.data_ptr()
        self.st.ddx = self.ddx.data_ptr()
        self.st.ddy = self.ddy.data_ptr() if self.ddy is not None else None
        self.st.ddz = self.ddz.data_ptr() if self.ddz is not None else None
        for k in range(5):
            self.st.f[k] = self.f[k].data_ptr()
        self.st.dens, self.st.pots, self.st.qpe = self.dens.data_ptr(), self.pots.data_ptr(), self.qpe.data_ptr()
        torch.cuda.synchronize(ctx.device)

    def history(self):
        by = {t.data_ptr(): t for t in self.f}
        return [by[self.st.f[k]] for k in range(5)]

    def initial_measure(self):
        check(self.ctx.L.tdw_initial_measure(self.ctx.h, C.byref(self.st)), "tdw_initial_measure")

    def selfstart(self):
        """Taylor self-start of the history (selfstart 1, pca.c:1086-1392):
        the state advances by 3 / 4 / 5 steps, the history is filled and the
        densities and potentials belong to the new state.  AB4AM5 / AB3AM4
        need one / two state-sized scratch arrays, allocated here for the
        call."""
        torch = self.ctx.torch
        k = {AB3AM4: 3, AB4AM5: 4, AB5AM5: 5}[self.ctx.params.scheme]
        scr = [torch.empty_like(self.wf) for _ in range(max(0, 5 - k))]
        ptrs = [t.data_ptr() for t in scr] + [None, None]
        check(self.ctx.L.tdw_selfstart(self.ctx.h, C.byref(self.st), ptrs[0], ptrs[1]), "tdw_selfstart")
        self.ctx.sync()
        del scr

    def step(self, it, grad_last=False):
        fn = self.ctx.L.tdw_step_fused if self.fused else self.ctx.L.tdw_step
        check(fn(self.ctx.h, C.byref(self.st), it, int(grad_last)), "tdw_step")

    def finalize(self):
        """Apply a fused step's pending normalisation / <H> subtraction."""
        check(self.ctx.L.tdw_finalize(self.ctx.h, C.byref(self.st)), "tdw_finalize")

    def energy(self, it):
        return self.ctx.compute_energy(it, self.dens, self.pots)

    def save_checkpoint(self, path, t0, fbeta, kt=None, **scalars):
        """Writes the run as a td checkpoint file (tdwslda_checkpoint.c
        layout); scalars: mu_a, mu_b, ec, kF, eF, Effg, beta."""
        sc = CheckpointScalars(**scalars)
        fb = np.ascontiguousarray(fbeta, dtype=np.float64)
        ktv = None if kt is None else np.ascontiguousarray(kt, dtype=np.float64)
        check(self.ctx.L.tdw_checkpoint_save(self.ctx.h, path.encode(), C.byref(self.st), float(t0),
                                        fb.ctypes.data_as(C.c_void_p),
                                        None if ktv is None else ktv.ctypes.data_as(C.c_void_p), C.byref(sc)),
              "tdw_checkpoint_save")

    def load_checkpoint(self, path):
        """Reads a td checkpoint file into the run; returns (t0, fbeta, kt,
        scalars dict)."""
        n = self.n
        t0 = C.c_double()
        fb = np.zeros(n)
        kt = np.zeros(2 * n)
        sc = CheckpointScalars()
        check(self.ctx.L.tdw_checkpoint_load(self.ctx.h, path.encode(), C.byref(self.st), C.byref(t0),
                                        fb.ctypes.data_as(C.c_void_p), kt.ctypes.data_as(C.c_void_p), C.byref(sc)),
              "tdw_checkpoint_load")
        dim = self.ctx.params.dim
        ktr = None if dim == 3 else (kt[:n].copy() if dim == 2 else kt)
        return t0.value, fb, ktr, {k: getattr(sc, k) for k, _ in CheckpointScalars._fields_}


class WData:
    """W-data measurement output of a run (tdw_wdata_*: <outprefix>.wtxt and
    <outprefix>_<var>.wdat, one datablock per measurement), as the td codes
    write it (wslda_writevars.c).  consts: {name: value} written to the
    metadata before the first cycle (pca.c:1026-1030 sets kF, eF, mu_a, mu_b)."""

    def __init__(self, ctx, outprefix, writevar="default", writeprec="d", spinsymmetry=0, overwrite=1,
                 inittype=0, t0=0.0, dt=1.0, codename=None, consts=None):
        self.ctx = ctx
        self.L = ctx.L
        wp = WDataParams()
        wp.outprefix = str(outprefix).encode()
        wp.writevar = writevar.encode()
        wp.writeprec = writeprec.encode()[:1]
        wp.spinsymmetry, wp.overwrite, wp.inittype = int(spinsymmetry), int(overwrite), int(inittype)
        wp.t0, wp.dt = float(t0), float(dt)
        wp.codename = codename.encode() if codename else None
        h = C.c_void_p()
        check(ctx.L.tdw_wdata_create(ctx.h, C.byref(wp), C.byref(h)), "tdw_wdata_create")
        self.h = h
        for k, v in (consts or {}).items():
            check(self.L.tdw_wdata_setconst(self.h, k.encode(), float(v), None), "tdw_wdata_setconst")
        check(self.L.tdw_wdata_begin(self.h), "tdw_wdata_begin")

    def write(self, it, dens, pots):
        check(self.L.tdw_wdata_write(self.h, int(it), _ptr(dens), _ptr(pots)), "tdw_wdata_write")

    @property
    def cycles(self):
        return self.L.tdw_wdata_cycles(self.h)

    def close(self):
        if self.h:
            self.L.tdw_wdata_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
