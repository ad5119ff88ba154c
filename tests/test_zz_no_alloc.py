"""tdw_normalize_noalloc (no device allocation; partial sums in the library's
fixed per-device scratch array) against tdw_normalize through the C ABI:
the same normalised states and norms, bit for bit, on two calls in a row,
the second on a smaller input; inputs beyond the scratch size are refused
with TDW_ESCRATCH before any work is queued."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def random_states(torch, n, N, seed):
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    re = torch.randn(2 * n, N, generator=g, dtype=torch.float64)
    im = torch.randn(2 * n, N, generator=g, dtype=torch.float64)
    return torch.complex(re, im).to("cuda:0")


def test_normalize_noalloc_matches_normalize(torch):
    from wslda_amd import td
    print("case: tdw_normalize_noalloc vs tdw_normalize, 24^3, 2 calls (300 then 37 states)", flush=True)
    p = td.make_params(3, 24, 24, 24, dx=1.0)
    nwf = 300
    ctx = td.TDContext(p, nwf, 0)
    L = ctx.L
    N = ctx.N
    for n, seed, with_norms in ((300, 11, True), (37, 12, False)):
        a = random_states(torch, n, N, seed)
        b = a.clone()
        na = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        nb = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        td.check(L.tdw_normalize(ctx.h, n, ptr(a), ptr(na)), "tdw_normalize")
        td.check(L.tdw_normalize_noalloc(ctx.h, n, ptr(b), ptr(nb) if with_norms else None),
                 "tdw_normalize_noalloc")
        ctx.sync()
        assert torch.equal(a, b)
        if with_norms:
            assert torch.equal(na, nb)
        # normalised: sum |u|^2 + |v|^2 dV = 1
        s = ((b.abs() ** 2).sum(dim=1)[:n] + (b.abs() ** 2).sum(dim=1)[n:]).cpu().numpy()
        assert np.abs(s - 1.0).max() < 1e-13
    # beyond the scratch: refused on the host, nothing launched
    big = td.make_params(3, 32, 32, 96, dx=1.0)
    nbig = td.NOALLOC_SCRATCH_DOUBLES // 24 + 1     # ceil(32*32*96 / 4096) = 24 chunks per state
    ctx2 = td.TDContext(big, nbig, 0)
    dummy = torch.zeros(1, dtype=torch.complex128, device="cuda:0")
    assert ctx2.L.tdw_normalize_noalloc(ctx2.h, nbig, ptr(dummy), ptr(dummy)) == td.TDW_ESCRATCH
    assert ctx2.L.tdw_normalize_noalloc(ctx2.h, 0, ptr(dummy), None) == td.TDW_EINVAL
    ctx2.close()
    ctx.close()
