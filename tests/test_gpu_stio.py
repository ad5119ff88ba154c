"""The driver's read_s3dpca_wf (pca_io.h:1562) through the drop-in layer:
N comes from the bound context; the states it reads equal the files'."""
import ctypes as C
import sys

import numpy as np
import pytest

from tests import stfiles

pytestmark = pytest.mark.gpu


def test_dropin_read_s3dpca_wf(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from wslda_amd import td
    print("case: drop-in read_s3dpca_wf, 6x4x6 lattice, 12 states in 3 files", flush=True)
    sys.stdout.flush()
    shape = (6, 4, 6)
    per = [3, 4, 5]
    ref = stfiles.write_st3d(str(tmp_path), shape, 0.8, per, seed=9)
    p = td.make_params(3, *shape, dx=0.8)
    ctx = td.TDContext(p, 12, 0)
    L = td.lib()
    td.check(L.tdw_bind(ctx.h), "tdw_bind")
    N = ctx.N
    perr = np.array(per, dtype=np.int32)
    lo, hi = 2, 9
    wf = np.zeros((2 * (hi - lo), N), dtype=np.complex128)
    en = np.zeros(hi - lo)
    rc = L.read_s3dpca_wf(str(tmp_path).encode(), 3, perr.ctypes.data_as(C.c_void_p), lo, hi,
                          wf.ctypes.data_as(C.c_void_p), en.ctypes.data_as(C.c_void_p))
    assert rc == 0
    assert np.array_equal(wf[:hi - lo], ref["u"][lo:hi]) and np.array_equal(wf[hi - lo:], ref["v"][lo:hi])
    assert np.array_equal(en, ref["en"][lo:hi])
    L.tdw_bind(None)
    ctx.close()
