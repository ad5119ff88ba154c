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


@pytest.mark.parametrize("codedim", [1, 2, 3])
def test_dropin_read_stwslda1d_wf(tmp_path, codedim):
    """read_stwslda1d_wf on the bound context: a 1D run takes each state
    once with its mode's degeneracy; a 3D run expands each state over its
    mode's (ky, kz) as u(x) e^{i(ky y + kz z)} / sqrt(LY LZ)
    (pca_io.h:1177-1372)."""
    import math
    from tests import st1dfiles
    from wslda_amd import td
    from tests.test_st1d import kmodes
    nx, ny, nz, dx, dy, dz = 12, 6, 4, 0.9, 1.1, 0.7
    p = td.make_params(codedim, nx, ny, nz, dx=dx, dy=dy, dz=dz)
    arr = kmodes(p)
    modes = [(m.ky, m.kz, m.weight) for m in arr][:3]
    per = [2, 3, 1]
    ref = st1dfiles.write_st1d(str(tmp_path), nx, modes, per, new=True, seed=4)
    kv = (td.KMode * 3)(*[td.KMode(*m) for m in modes])
    nwf = C.c_int(sum(n * m[2] for n, m in zip(per, modes)))
    got = (C.c_int * 3)()
    L = td.lib()
    LY, LZ = ny * dy, nz * dz
    ky_grid = [2 * math.pi / LY * (i if i < ny // 2 else i - ny) for i in range(ny)]
    kz_grid = [2 * math.pi / LZ * (i if i < nz // 2 else i - nz) for i in range(nz)]

    def members(mky, mkz):
        return [(a, b) for i1, a in enumerate(ky_grid) for i2, b in enumerate(kz_grid)
                if i1 != ny // 2 and i2 != nz // 2 and abs(a * a + b * b - mky ** 2 - mkz ** 2) < 1e-12]

    def members_2d(mky, mkz):
        # grouped by kz in order of first appearance, kz >= 0 kept
        groups = {}
        for a, b in members(mky, mkz):
            key = next((k for k in groups if abs(k - b) < 1e-12), b)
            groups.setdefault(key, []).append(a)
        return [(a, kz) for kz, ys in groups.items() if kz > -1e-12 for a in ys]

    # the scan of a 2D run reads the mode table of the bound context
    N = nx if codedim == 1 else nx * ny if codedim == 2 else nx * ny * nz
    nbig = sum(per) * 8
    ctx = td.TDContext(p, nbig, 0)
    td.check(L.tdw_bind(ctx.h), "tdw_bind")
    try:
        assert L.scan_stwslda1d_info_files(str(tmp_path).encode(), codedim, 3, C.cast(kv, C.c_void_p),
                                           C.byref(nwf), C.cast(got, C.c_void_p)) == 0
    finally:
        L.tdw_bind(None)
    n = nwf.value
    if codedim == 2:
        assert n == sum(c * len(members_2d(m[0], m[1])) for c, m in zip(per, modes))
    ctx.close()
    ctx = td.TDContext(p, n, 0)
    wf = np.zeros((2 * n, N), dtype=np.complex128)
    fb = np.zeros(n)
    kk = np.zeros(2 * n)
    cnt = np.zeros(n, dtype=np.int32)
    td.check(L.tdw_bind(ctx.h), "tdw_bind")
    try:
        rc = L.read_stwslda1d_wf(str(tmp_path).encode(), codedim, 3, C.cast(kv, C.c_void_p),
                                 C.cast(got, C.c_void_p), 0, n, wf.ctypes.data_as(C.c_void_p),
                                 fb.ctypes.data_as(C.c_void_p), kk.ctypes.data_as(C.c_void_p),
                                 cnt.ctypes.data_as(C.c_void_p))
    finally:
        L.tdw_bind(None)
    assert rc == 0
    U, V, E = np.concatenate(ref["u"]), np.concatenate(ref["v"]), np.concatenate(ref["en"])
    if codedim == 1:
        assert np.array_equal(wf[:n], U) and np.array_equal(wf[n:], V) and np.array_equal(fb, E)
        assert np.array_equal(kk[:n], ref["ky"]) and np.array_equal(kk[n:], ref["kz"])
        w = {(m[0], m[1]): m[2] for m in modes}
        assert list(cnt) == [w[(a, b)] for a, b in zip(ref["ky"], ref["kz"])]
    elif codedim == 2:
        row, s0 = 0, 0
        for (mky, mkz, wgt), cnt_m in zip(modes, per):
            for a, b in members_2d(mky, mkz):
                ph = np.exp(1j * a * np.arange(ny) * dy) / math.sqrt(LY)
                for k in range(cnt_m):
                    assert np.abs(wf[row] - (U[s0 + k][:, None] * ph[None, :]).ravel()).max() < 1e-14
                    assert np.abs(wf[n + row] - (V[s0 + k][:, None] * ph[None, :]).ravel()).max() < 1e-14
                    assert fb[row] == E[s0 + k] and kk[row] == b
                    row += 1
            s0 += cnt_m
        assert row == n
    else:
        iy, iz = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
        row, s0 = 0, 0
        for (mky, mkz, wgt), cnt_m in zip(modes, per):
            mem = members(mky, mkz)
            assert len(mem) == wgt
            for a, b in mem:
                ph = (np.exp(1j * (a * iy * dy + b * iz * dz)) / math.sqrt(LY * LZ)).ravel()
                for k in range(cnt_m):
                    want_u = (U[s0 + k][:, None] * ph[None, :]).ravel()
                    want_v = (V[s0 + k][:, None] * ph[None, :]).ravel()
                    assert np.abs(wf[row] - want_u).max() < 1e-14
                    assert np.abs(wf[n + row] - want_v).max() < 1e-14
                    assert fb[row] == E[s0 + k]
                    row += 1
            s0 += cnt_m
        assert row == n
    ctx.close()
