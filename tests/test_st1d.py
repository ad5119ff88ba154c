"""st-wslda-1d starts (inittype 1): the transverse-mode table of
wslda_wavevectors.h:25-100 against a brute-force count, and
scan_stwslda1d_info_files (pca_io.h:845-912) on files written in the
static solver's layout, host only.  Parity with files a real st-wslda-1d run
wrote is unpinned: the reference holds none."""
import ctypes as C
import math

import numpy as np
import pytest

from tests import st1dfiles
from wslda_amd import td


def kmodes(p):
    L = td.lib()
    cnt = C.c_int()
    td.check(L.tdw_st1d_kmodes(C.byref(p), None, 0, C.byref(cnt)), "count")
    arr = (td.KMode * cnt.value)()
    td.check(L.tdw_st1d_kmodes(C.byref(p), C.cast(arr, C.c_void_p), cnt.value, C.byref(cnt)), "kmodes")
    return arr


def kgrid(n, L):
    return np.array([2 * math.pi / L * (i if i < n // 2 else i - n) for i in range(n)])


@pytest.mark.parametrize("ny,nz,cubic", [(6, 4, 0), (8, 8, 0), (6, 6, 1)])
def test_kmode_table(ny, nz, cubic):
    p = td.make_params(1, 16, ny, nz, dx=0.9, dy=1.1, dz=0.7,
                       regularization=td.CUBIC_CUTOFF if cubic else td.SPHERICAL_CUTOFF)
    arr = kmodes(p)
    ky, kz = kgrid(ny, ny * 1.1), kgrid(nz, nz * 0.7)
    pts = [(a, b) for iy, a in enumerate(ky) for iz, b in enumerate(kz)
           if cubic or not (iy == ny // 2 or iz == nz // 2)]
    assert sum(m.weight for m in arr) == len(pts) == (ny * nz if cubic else (ny - 1) * (nz - 1))
    for m in arr:
        k2 = m.ky ** 2 + m.kz ** 2
        assert m.weight == sum(1 for a, b in pts if abs(a * a + b * b - k2) < 1e-12)
    # every class appears once
    k2s = sorted(m.ky ** 2 + m.kz ** 2 for m in arr)
    assert all(b - a > 1e-12 for a, b in zip(k2s, k2s[1:]))


@pytest.mark.parametrize("new", [True, False])
def test_scan_info_files(tmp_path, new):
    p = td.make_params(1, 16, 6, 4, dx=0.9, dy=1.1, dz=0.7)
    arr = kmodes(p)
    modes = [(m.ky, m.kz, m.weight) for m in arr][:3]
    per = [3, 2, 4]
    st1dfiles.write_st1d(str(tmp_path), 16, modes, per, new=new, seed=1)
    L = td.lib()
    kv = (td.KMode * 3)(*[td.KMode(*m) for m in modes])
    nwf = C.c_int(sum(n * m[2] for n, m in zip(per, modes)))
    got = (C.c_int * 3)()
    assert L.scan_stwslda1d_info_files(str(tmp_path).encode(), 1, 3, C.cast(kv, C.c_void_p), C.byref(nwf),
                                       C.cast(got, C.c_void_p)) == 0
    assert list(got) == per and nwf.value == sum(per)
    nwf = C.c_int(sum(n * m[2] for n, m in zip(per, modes)))
    assert L.scan_stwslda1d_info_files(str(tmp_path).encode(), 3, 3, C.cast(kv, C.c_void_p), C.byref(nwf),
                                       C.cast(got, C.c_void_p)) == 0
    assert nwf.value == sum(n * m[2] for n, m in zip(per, modes))
    bad = C.c_int(nwf.value + 1)
    assert L.scan_stwslda1d_info_files(str(tmp_path).encode(), 1, 3, C.cast(kv, C.c_void_p), C.byref(bad),
                                       C.cast(got, C.c_void_p)) == 10026
    np.array([3, 2], dtype=np.int32).tofile(str(tmp_path / ("wf.0001.info" if new else "s1dpca.0001.info")))
    nwf = C.c_int(0)
    assert L.scan_stwslda1d_info_files(str(tmp_path).encode(), 1, 3, C.cast(kv, C.c_void_p), C.byref(nwf),
                                       C.cast(got, C.c_void_p)) == 10032
