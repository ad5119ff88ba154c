"""Static-solver (st-wslda-3d) output files written in the layout the td
driver reads (pca_io.h:340-417, 1519-1633): <dir>/wf.info (or s3dpca.info),
per-file info, u / v / energy files and the potentials, raw little-endian
binary."""
import os
import struct

import numpy as np


def write_st3d(d, shape, dx, per_file, new=True, seed=0, dim_tag=3, kF=1.0, mu=(0.3, 0.25), ec=4.0, beta=20.0):
    nx, ny, nz = shape
    N = nx * ny * nz
    nwf = int(sum(per_file))
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((nwf, N)) + 1j * rng.standard_normal((nwf, N))
    v = rng.standard_normal((nwf, N)) + 1j * rng.standard_normal((nwf, N))
    en = rng.uniform(-0.5, 0.5, nwf)
    pots = rng.standard_normal(12 * N)
    pre = "wf" if new else "s3dpca"
    head = struct.pack("<5i", dim_tag if new else 8, nwf, nx, ny, nz)
    head += struct.pack("<8d", dx, dx, dx, kF, mu[0], mu[1], ec, beta)
    with open(os.path.join(d, pre + ".info"), "wb") as f:
        f.write(head)
    a = 0
    for k, m in enumerate(per_file):
        with open(os.path.join(d, "%s.%04d.info" % (pre, k)), "wb") as f:
            f.write(struct.pack("<2i", 3 if new else 8, m))
        names = ("_un.wdat", "_vn.wdat", "_en.dat") if new else (".wfu", ".wfv", ".en")
        stem = os.path.join(d, "%s.%04d" % (pre, k))
        u[a:a + m].astype(np.complex128).tofile(stem + names[0])
        v[a:a + m].astype(np.complex128).tofile(stem + names[1])
        en[a:a + m].astype(np.float64).tofile(stem + names[2])
        a += m
    pots.astype(np.float64).tofile(os.path.join(d, pre + ".pud"))
    return dict(u=u, v=v, en=en, pots=pots, kF=kF, mu=mu, ec=ec, beta=beta, nwf=nwf)
