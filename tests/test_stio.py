"""Static-solver initial states (inittype 3 of the td driver, pca.c:628-725):
the info / scan / wave-function readers of pca_io.h:340-417, 1519-1633 on
files written in the static solver's layout, both file-name generations,
host only.  Parity with files a real st-wslda-3d run wrote is unpinned: the
reference holds none."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from tests import stfiles
from wslda_amd import td

SHAPE = (6, 4, 5)


def params(shape=SHAPE, dx=0.8):
    return td.make_params(3, *shape, dx=dx)


@pytest.mark.parametrize("new", [True, False])
def test_st3d_roundtrip(tmp_path, new):
    per = [4, 5, 6]
    ref = stfiles.write_st3d(str(tmp_path), SHAPE, 0.8, per, new=new, seed=3)
    info, per_read = td.st3d_info(str(tmp_path), 3)
    assert info.nwf == 15 and list(per_read) == per and info.new_format == int(new)
    assert (info.nx, info.ny, info.nz) == SHAPE and info.dx == 0.8 and info.beta == ref["beta"]
    assert (info.mu_a, info.mu_b, info.ec, info.kF) == (ref["mu"][0], ref["mu"][1], ref["ec"], ref["kF"])
    p = params()
    for spinsym in (0, 1):
        wf, fb, pots, sc = td.st3d_load(p, str(tmp_path), 3, info, per_read, spinsymmetry=spinsym)
        assert np.array_equal(wf[:15], ref["u"]) and np.array_equal(wf[15:], ref["v"])
        assert np.array_equal(pots, ref["pots"])
        f = np.array([td.fbeta(e, ref["beta"]) for e in ref["en"]])
        assert np.array_equal(fb, f)
        Na = float(np.sum(np.abs(ref["u"]) ** 2 * f[:, None]))
        Nb = float(np.sum(np.abs(ref["v"]) ** 2 * (1 - f)[:, None]))
        eF = 0.5 * ref["kF"] ** 2
        assert sc.eF == eF and sc.beta == ref["beta"] and sc.mu_a == ref["mu"][0] and sc.ec == ref["ec"]
        assert math.isclose(sc.Effg, 0.6 * (Na + Nb if spinsym else Na) * eF, rel_tol=1e-12)


def test_st3d_state_ranges_across_files(tmp_path):
    per = [4, 5, 6]
    ref = stfiles.write_st3d(str(tmp_path), SHAPE, 0.8, per, seed=5)
    info, per_read = td.st3d_info(str(tmp_path), 3)
    for lo, hi in ((0, 4), (3, 11), (9, 10), (4, 15), (13, 15)):
        wf, fb, _, _ = td.st3d_load(params(), str(tmp_path), 3, info, per_read, lo, hi, potentials=False)
        n = hi - lo
        assert np.array_equal(wf[:n], ref["u"][lo:hi]) and np.array_equal(wf[n:], ref["v"][lo:hi]), (lo, hi)


def test_st3d_error_codes(tmp_path):
    lib = td.lib()
    d = str(tmp_path)
    info = td.StInfo()
    per = np.zeros(3, dtype=np.int32)
    pp = per.ctypes.data_as(C.c_void_p)
    # no info file at all: WSLDA_ERR_INITSTATE_FILE_NOT_FOUND
    assert lib.tdw_st3d_info(d.encode(), 3, C.byref(info), pp) == 10033
    # a 1D / 2D solution offered to the 3D code
    for tag, code in ((1, 10030), (2, 10031)):
        sub = os.path.join(d, "dim%d" % tag)
        os.makedirs(sub)
        stfiles.write_st3d(sub, SHAPE, 0.8, [2, 2, 2], dim_tag=tag)
        assert lib.tdw_st3d_info(sub.encode(), 3, C.byref(info), pp) == code
    # per-file counts that do not add up to the header's nwf
    sub = os.path.join(d, "count")
    os.makedirs(sub)
    stfiles.write_st3d(sub, SHAPE, 0.8, [2, 2, 2])
    assert lib.tdw_st3d_info(sub.encode(), 2, C.byref(info), pp) == 10006
    # a missing per-file info file
    assert lib.tdw_st3d_info(sub.encode(), 4, C.byref(info), np.zeros(4, np.int32).ctypes.data_as(C.c_void_p)) == 10007
    # missing / truncated wave-function files
    sub = os.path.join(d, "wf")
    os.makedirs(sub)
    stfiles.write_st3d(sub, SHAPE, 0.8, [2, 3])
    info, per2 = td.st3d_info(sub, 2)
    os.remove(os.path.join(sub, "wf.0001_vn.wdat"))
    with pytest.raises(td.TDWError, match="code -2"):
        td.st3d_load(params(), sub, 2, info, per2)
    wf, _, _, _ = td.st3d_load(params(), sub, 2, info, per2, 0, 2, potentials=False)   # file 0 alone is whole
    with open(os.path.join(sub, "wf.0000_un.wdat"), "r+b") as f:
        f.truncate(16 * 6 * 4 * 5 + 8)
    with pytest.raises(td.TDWError, match="code -21"):
        td.st3d_load(params(), sub, 2, info, per2, 0, 2, potentials=False)
    # the files' lattice must be the run's
    with pytest.raises(td.TDWError, match="code -1001"):
        td.st3d_load(params(dx=0.9), sub, 2, info, per2)
