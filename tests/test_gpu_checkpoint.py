"""td checkpoint files: the library's writer against an independent reading
of the tdwslda_checkpoint.c layout, and save -> load -> continue against an
uninterrupted run (bit-identical)."""
import os
import struct

import numpy as np
import pytest

from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def read_checkpoint(path, dim, scheme, N):
    """The file as tdwslda_checkpoint.c:118-220 reads it for one rank."""
    b = open(path, "rb").read()
    t0, mu_a, mu_b, ec, kF = struct.unpack_from("<5d", b, 0)
    (nwf,) = struct.unpack_from("<i", b, 40)
    eF, Effg, beta = struct.unpack_from("<3d", b, 44)
    pad = nwf + 1            # nwfip + nwfp with one rank
    off = 68
    fb = np.frombuffer(b, np.float64, nwf, off)
    off += 8 * pad
    kz = ky = None
    if dim <= 2:
        kz = np.frombuffer(b, np.float64, nwf, off)
        off += 8 * pad
    if dim == 1:
        ky = np.frombuffer(b, np.float64, nwf, off)
        off += 8 * pad
    narr = 4 + (scheme >= 45) + (scheme >= 55)
    arrs = []
    for _ in range(narr):
        arrs.append(np.frombuffer(b, np.complex128, 2 * nwf * N, off).reshape(2 * nwf, N))
        off += 2 * pad * N * 16
    pots = np.frombuffer(b, np.float64, 12 * N, off)
    off += 12 * N * 8
    return dict(t0=t0, mu=(mu_a, mu_b), ec=ec, kF=kF, nwf=nwf, eF=eF, Effg=Effg, beta=beta, fbeta=fb, kz=kz,
                ky=ky, arrays=arrs, pots=pots, size=len(b), expected_size=off)


@pytest.mark.parametrize("dim,shape,nst,scheme,fused", [(3, (12, 12, 12), 96, 55, True), (2, (16, 16, 8), 64, 45, False),
                                                         (1, (32, 8, 8), 48, 34, False)])
def test_checkpoint_roundtrip(torch, tmp_path, dim, shape, nst, scheme, fused):
    say("case: checkpoint dim=%d scheme=%d fused=%s" % (dim, scheme, fused))
    from wslda_amd import problems, td
    s = problems.uniform_start(dim, shape, n_states=nst, scheme=scheme)
    p = s["params"]
    p.vext = td.gauss_vext([(0.05, (1.0, 0.0, 0.0), (-0.1, -0.1 if dim >= 2 else 0.0, -0.1 if dim == 3 else 0.0))])
    n = s["wf"].shape[0] // 2
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(s["fbeta"], s["kt"], s["cnt"])
    R = td.TDRun(ctx, s["wf"], s["pots"], fused=fused)
    R.initial_measure()
    for it in range(7):
        R.step(it, grad_last=(it == 3))
    path = str(tmp_path / "checkpoint_0")
    sc = dict(mu_a=p.mu_a, mu_b=p.mu_b, ec=p.ec, kF=1.0, eF=0.5, Effg=3.25, beta=s["uniform"].beta)
    R.save_checkpoint(path, 7 * p.dt, s["fbeta"], s["kt"], **sc)
    N = ctx.N
    f = read_checkpoint(path, dim, scheme, N)
    assert f["size"] == f["expected_size"] == os.path.getsize(path)
    assert f["nwf"] == n and f["t0"] == 7 * p.dt and f["mu"] == (p.mu_a, p.mu_b) and f["Effg"] == 3.25
    assert np.array_equal(f["fbeta"], s["fbeta"])
    if dim <= 2:
        assert np.array_equal(f["kz"], s["kt"][:n])
    if dim == 1:
        assert np.array_equal(f["ky"], s["kt"][n:])
    h = {34: 3, 45: 4, 55: 5}[scheme]
    assert np.array_equal(f["arrays"][0], R.wf.cpu().numpy())
    hist = R.history()
    for j in range(len(f["arrays"]) - 1):
        assert np.array_equal(f["arrays"][1 + j], hist[j].cpu().numpy()), j
    assert np.array_equal(f["pots"], R.pots.cpu().numpy())
    # continue the original, and a fresh run loaded from the file
    ctx2 = td.TDContext(p, n, 0)
    ctx2.set_states(s["fbeta"], s["kt"], s["cnt"])
    R2 = td.TDRun(ctx2, np.zeros_like(s["wf"]), None, fused=fused)
    t0, fb, kt, sc2 = R2.load_checkpoint(path)
    assert t0 == 7 * p.dt and np.array_equal(fb, s["fbeta"]) and sc2["Effg"] == 3.25
    for it in range(7, 12):
        R.step(it, grad_last=(it == 11))
        R2.step(it, grad_last=(it == 11))
    R.finalize()
    R2.finalize()
    assert torch.equal(R.wf, R2.wf)
    assert torch.equal(R.dens, R2.dens)
    for a, b in zip(R.history()[:h], R2.history()[:h]):
        assert torch.equal(a, b)
    ctx.close()
    ctx2.close()
