"""Writes st-wslda-1d output in the layout the td driver's inittype-1
readers expect (pca_io.h:845-1372): per transverse mode ikz the files
wf.%04d.info (int dim = 1, int nwf) or s1dpca.%04d.info, the u and v of its
states (NX complex each), their energies and their (ky, kz) pairs."""
import os

import numpy as np


def write_st1d(prefix, nx, modes, per, new=True, seed=0):
    """modes: list of (ky, kz, weight); per: states per mode.  Returns the
    written u, v (rows in file order), energies and (ky, kz) per state."""
    rng = np.random.default_rng(seed)
    out = {"u": [], "v": [], "en": [], "ky": [], "kz": []}
    for ikz, ((ky, kz, _w), n) in enumerate(zip(modes, per)):
        u = rng.standard_normal((n, nx)) + 1j * rng.standard_normal((n, nx))
        v = rng.standard_normal((n, nx)) + 1j * rng.standard_normal((n, nx))
        en = rng.uniform(-2, 2, n)
        kk = np.array([[ky, kz]] * n, dtype=np.float64)
        names = ({"info": "wf.%04d.info", "u": "wf.%04d_un.wdat", "v": "wf.%04d_vn.wdat", "k": "wf.%04d_kkyz.dat",
                  "e": "wf.%04d_en.dat"} if new else
                 {"info": "s1dpca.%04d.info", "u": "s1dpca.%04d.wfu", "v": "s1dpca.%04d.wfv",
                  "k": "s1dpca.%04d.kkyz", "e": "s1dpca.%04d.en"})
        f = lambda k: os.path.join(prefix, names[k] % ikz)
        np.array([1, n], dtype=np.int32).tofile(f("info"))
        u.astype(np.complex128).tofile(f("u"))
        v.astype(np.complex128).tofile(f("v"))
        kk.tofile(f("k"))
        en.tofile(f("e"))
        out["u"].append(u)
        out["v"].append(v)
        out["en"].append(en)
        out["ky"] += [ky] * n
        out["kz"] += [kz] * n
    return out
