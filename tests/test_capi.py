"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from wslda_amd import td

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


PROJECT_HEADER = os.path.join(ROOT, "include", "wslda_td_project.h")
PROJECT_1D = os.path.join(ROOT, "tests", "projects", "td_gauss_1d")


def declared_functions(headers=None):
    names = set()
    if headers is None:
        headers = [h for h in glob.glob(os.path.join(ROOT, "include", "*.h")) if h != PROJECT_HEADER]
    for h in headers:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|size_t|void|double|tdw_ctx|tdw_project_inputs)\s*\**\s*([a-zA-Z_]\w*)\s*\(",
                             src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(td.LIBPATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert declared_functions() == set(td.EXPORTS)


def test_params_layout_matches_oracle_cfg():
    from oracle import pyoracle as po
    # the library's tdw_params starts with the oracle's cfg layout field by field
    for (a, _), (b, _) in zip(po.Cfg._fields_, td.Params._fields_):
        assert a == b
        assert getattr(po.Cfg, a).offset == getattr(td.Params, b).offset


def test_project_library_exports():
    """A project build (the td-gauss problem compiled for the 1D code) exports
    the engine's entry points and the problem's own functions, reports its
    compiled settings, and its process_params is the problem's (host call,
    no GPU): amplitude and times in eF, widths as -1/(2 s^2) in 1/kF."""
    import numpy as np
    path = os.path.join(PROJECT_1D, "libwslda_td.so")
    lib = ctypes.CDLL(path)
    names = declared_functions() | declared_functions([PROJECT_HEADER])
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    info = td.project_info(path)
    assert info is not None
    p, flags = info
    assert (p.dim, p.nx, p.ny, p.nz, p.functional) == (1, 12, 12, 12, td.SLDA)
    assert abs(p.dx - 1.15) < 1e-15 and p.p_nmin == 1e-7 and p.p_nmax == 1e-5 and p.denseps == 1e-8
    assert p.scheme == td.AB4AM5 and p.regularization == td.SPHERICAL_CUTOFF and p.mf_correction == 0
    assert flags == 0x01   # ENABLE_V_EXT only
    assert td.project_info(td.LIBPATH) is None
    prm = np.zeros(32)
    prm[:8] = [0.5, 2.0, 0.0, 0.0, 5.0, 55.0, 15.0, 1.0]
    kF = ctypes.c_double(1.3)
    mu = np.array([0.4, 0.45])
    lib.process_params.restype = None
    lib.process_params(prm.ctypes.data_as(ctypes.c_void_p), ctypes.byref(kF), mu.ctypes.data_as(ctypes.c_void_p),
                       ctypes.c_size_t(0), None)
    eF = 0.5 * 1.3 ** 2
    assert abs(prm[0] - 0.5 * eF) < 1e-15
    assert abs(prm[1] - (-0.5 / (2.0 / 1.3) ** 2)) < 1e-15 and prm[2] == 0.0 and prm[3] == 0.0
    assert np.allclose(prm[4:7], np.array([5.0, 55.0, 15.0]) / eF, rtol=1e-15)
