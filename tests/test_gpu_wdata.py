"""W-data measurement output: the library's data set read back by an
independent reader written from the format description
(lib/wdata/doc/README.md: <prefix>.wtxt metadata, <prefix>_<var>.wdat
datablocks of nx*ny*nz (datadim 3), nx*ny (2) or nx (1) elements; real 8 B,
complex 16 B, vector 3 x 8 B; real4 / complex8 / vector4 in FP32), compared
with the run's own device densities and potentials (bit-exact), V_ext with
the oracle's, and the .wtxt text with the printf layout of
wdata_print_metadata (lib/wdata/c/wdata.c:495-584)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import tdgauss
from tests.test_gpu_parity import gpu_run, say

pytestmark = pytest.mark.gpu

SIZES = {"real": (8, 1, np.float64), "real8": (8, 1, np.float64), "real4": (4, 1, np.float32),
         "complex": (16, 1, np.complex128), "complex16": (16, 1, np.complex128), "complex8": (8, 1, np.complex64),
         "vector": (8, 3, np.float64), "vector8": (8, 3, np.float64), "vector4": (4, 3, np.float32)}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.skip("no HIP device")
    return t


def read_wdata(wtxt):
    """Metadata and every variable's cycles, as the format description says."""
    md, var, link, const = {}, {}, {}, {}
    for line in open(wtxt):
        line = line.split("#", 1)[0].split()
        if not line:
            continue
        tag = line[0]
        if tag == "var":
            var[line[1]] = dict(type=line[2], unit=line[3], format=line[4])
        elif tag == "link":
            link[line[1]] = line[2]
        elif tag == "const":
            const[line[1]] = float(line[2])
        else:
            md[tag.lower()] = line[1]
    nx, ny, nz, dd = int(md["nx"]), int(md["ny"]), int(md["nz"]), int(md["datadim"])
    bl = {1: nx, 2: nx * ny, 3: nx * ny * nz}[dd]
    cyc = int(md["cycles"])
    data = {}
    base = os.path.join(os.path.dirname(wtxt), os.path.basename(md["prefix"]))
    for name, v in var.items():
        sz, comps, dt = SIZES[v["type"]]
        fn = "%s_%s.%s" % (base, name, v["format"])
        if cyc == 0 and not os.path.exists(fn):    # files appear with the first cycle
            data[name] = None
            continue
        raw = open(fn, "rb").read()
        nel = bl * comps
        assert len(raw) == cyc * nel * sz, (name, len(raw))
        data[name] = np.frombuffer(raw, dt).reshape(cyc, comps, bl) if comps > 1 else \
            np.frombuffer(raw, dt).reshape(cyc, bl)
    return md, var, link, const, data


def expected_wtxt_body(nx, ny, nz, dx, dy, dz, dd, prefix, cycles, t0, dt, variables, links, consts):
    """The metadata text after the header line, in wdata_print_metadata's layout."""
    L = ["# W-DATA 0.3.0",
         "# More information about the WData format see: https://gitlab.fizyka.pw.edu.pl/wtools/wdata/-/wikis/home",
         "nx %24d   # lattice" % nx, "ny %24d   # lattice" % ny, "nz %24d   # lattice" % nz,
         "dx %24g   # spacing" % dx, "dy %24g   # spacing" % dy, "dz %24g   # spacing" % dz,
         "datadim %19d   # dimension of block size: 1=nx, 2=nx*ny, 3=nx*ny*nz" % dd,
         ("prefix %20s   # prefix for files belonging to this data set`" % prefix) if len(prefix) <= 20 else
         ("prefix %s   # prefix for files belonging to this data set`" % prefix),
         "cycles %20d   # number of cycles (measurements)" % cycles,
         "t0 %24g   # time value for the first cycle" % t0,
         "dt %24g   # time interval between cycles" % dt]
    if variables:
        L += ["", "# variables",
              "# tag                  name                    type                    unit                  format"]
        L += ["var%24s%24s%24s%24s" % v for v in variables]
    if links:
        L += ["", "# links", "# tag                  name                 link-to"]
        L += ["link%23s%24s" % l for l in links]
    if consts:
        L += ["", "# consts", "# tag                  name                   value                    unit"]
        L += ["const%22s%24g%24s" % c for c in consts]
    return "\n".join(L) + "\n\n"


@pytest.mark.parametrize("dim", [3, 2])
def test_wdata_dataset(torch, tmp_path, dim):
    say("case: wdata dim=%d" % dim)
    from wslda_amd import td
    s = tdgauss.setup(dim)
    cfg = s["cfg"]
    ctx, R = gpu_run(torch, s)
    N = ctx.N
    it0 = 2000                      # t*eF = 7: V_ext switching on
    prefix = str(tmp_path / "run")
    consts = dict(kF=s["kF"], eF=s["eF"], mu_a=cfg.mu_a, mu_b=cfg.mu_b)
    W = td.WData(ctx, prefix, writevar="all", writeprec="d", t0=it0 * cfg.dt, dt=3 * cfg.dt, consts=consts)
    Wf = td.WData(ctx, str(tmp_path / "runf"), writevar="rho delta j nu", writeprec="f", spinsymmetry=1,
                  t0=it0 * cfg.dt, dt=3 * cfg.dt)
    R.initial_measure()
    snaps = []
    it = it0
    for cyc in range(3):
        if cyc:
            for _ in range(3):
                R.step(it, grad_last=True)
                it += 1
        W.write(it, R.dens, R.pots)
        Wf.write(it, R.dens, R.pots)
        ctx.sync()
        snaps.append((it, R.dens.cpu().numpy().copy(), R.pots.cpu().numpy().copy()))
    assert W.cycles == 3 and Wf.cycles == 3
    W.close()
    Wf.close()

    md, var, link, const, data = read_wdata(prefix + ".wtxt")
    assert int(md["cycles"]) == 3 and int(md["datadim"]) == dim
    assert (int(md["nx"]), int(md["ny"]), int(md["nz"])) == (cfg.nx, cfg.ny, cfg.nz)
    names = ["rho_a", "rho_b", "delta", "j_a", "j_b", "nu", "tau_a", "tau_b", "V_a", "V_b", "V_ext_a", "V_ext_b",
             "delta_ext", "velocity_ext_a", "velocity_ext_b", "alpha_a", "alpha_b", "A_a", "A_b"]
    assert list(var) == names and not link
    assert const["kF"] == pytest.approx(s["kF"], rel=1e-5)
    for k, (itk, d, p) in enumerate(snaps):
        va, vb = np.zeros(N), np.zeros(N)
        po.lib().tdwo_vext_eval(C.byref(cfg), itk, po.ptr(va), po.ptr(vb))
        assert np.abs(va).max() > 0
        assert np.array_equal(data["rho_a"][k], d[2 * N:3 * N])
        assert np.array_equal(data["rho_b"][k], d[7 * N:8 * N])
        assert np.array_equal(data["tau_a"][k], d[3 * N:4 * N])
        assert np.array_equal(data["tau_b"][k], d[8 * N:9 * N])
        assert np.array_equal(data["j_a"][k].ravel(), d[4 * N:7 * N])
        assert np.array_equal(data["j_b"][k].ravel(), d[9 * N:12 * N])
        assert np.array_equal(data["nu"][k], d[:2 * N].view(np.complex128))
        assert np.array_equal(data["delta"][k], p[2 * N:4 * N].view(np.complex128))
        # the built-in problem has no Delta_ext (get_delta_ext gives zeros)
        assert not data["delta_ext"][k].any()
        assert np.array_equal(data["alpha_a"][k], p[4 * N:5 * N])
        assert np.array_equal(data["A_b"][k].ravel(), p[9 * N:12 * N])
        assert not data["velocity_ext_a"][k].any() and not data["velocity_ext_b"][k].any()
        # V_ext against the oracle, V written as V - V_ext
        assert np.abs(data["V_ext_a"][k] - va).max() <= 1e-14 * np.abs(va).max()
        assert np.abs(data["V_ext_b"][k] - vb).max() <= 1e-14 * max(np.abs(vb).max(), 1e-300)
        assert np.array_equal(data["V_a"][k], p[:N] - data["V_ext_a"][k])
        assert np.array_equal(data["V_b"][k], p[N:2 * N] - data["V_ext_b"][k])
    # .wtxt text after the generator line
    txt = open(prefix + ".wtxt").read()
    head, body = txt.split("\n\n", 1)
    assert head.startswith("# Generated by td-wslda-%dd, engine version " % dim)
    variables = [(nm, var[nm]["type"], "none", "wdat") for nm in names]
    cs = [(k, v, "none") for k, v in consts.items()]
    assert body == expected_wtxt_body(cfg.nx, cfg.ny, cfg.nz, cfg.dx, cfg.dy, cfg.dz, dim, prefix, 3,
                                      it0 * cfg.dt, 3 * cfg.dt, variables, [], cs)

    # FP32, spin-symmetric set: _b as links, blocks are the FP32 casts
    md, var, link, const, data = read_wdata(str(tmp_path / "runf.wtxt"))
    assert list(var) == ["rho_a", "delta", "j_a", "nu"]
    assert var["rho_a"]["type"] == "real4" and var["delta"]["type"] == "complex8" and var["j_a"]["type"] == "vector4"
    assert link == {"rho_b": "rho_a", "j_b": "j_a"}
    assert not os.path.exists(str(tmp_path / "runf_rho_b.wdat"))
    for k, (itk, d, p) in enumerate(snaps):
        assert np.array_equal(data["rho_a"][k], d[2 * N:3 * N].astype(np.float32))
        assert np.array_equal(data["j_a"][k].ravel(), d[4 * N:7 * N].astype(np.float32))
        assert np.array_equal(data["delta"][k].view(np.float32), p[2 * N:4 * N].astype(np.float32))
    ctx.close()


def test_wdata_overwrite_and_first_cycle(torch, tmp_path):
    say("case: wdata overwrite / inittype")
    from wslda_amd import td
    s = tdgauss.setup(1)
    cfg = s["cfg"]
    ctx, R = gpu_run(torch, s)
    N = ctx.N
    R.initial_measure()
    prefix = str(tmp_path / "st")
    # a run started from a static solution (inittype 1..3) writes V itself at it == 0
    W = td.WData(ctx, prefix, writevar="V", inittype=1)
    W.write(0, R.dens, R.pots)
    W.write(2000, R.dens, R.pots)
    W.close()
    ctx.sync()
    p = R.pots.cpu().numpy()
    _, _, _, _, data = read_wdata(prefix + ".wtxt")
    va, vb = np.zeros(N), np.zeros(N)
    po.lib().tdwo_vext_eval(C.byref(cfg), 2000, po.ptr(va), po.ptr(vb))
    assert np.abs(va).max() > 0
    assert np.array_equal(data["V_a"][0], p[:N])
    assert np.abs(data["V_a"][1] - (p[:N] - va)).max() <= 1e-14 * max(1.0, np.abs(va).max())
    # overwrite = 0 refuses existing files; overwrite = 1 starts the set again
    with pytest.raises(td.TDWError, match="-1005"):
        td.WData(ctx, prefix, writevar="V", overwrite=0)
    W2 = td.WData(ctx, prefix, writevar="V", overwrite=1)
    assert W2.cycles == 0
    W2.write(7, R.dens, R.pots)
    W2.close()
    md, _, _, _, data = read_wdata(prefix + ".wtxt")
    assert int(md["cycles"]) == 1 and data["V_a"].shape == (1, N)
    # default variable list and the reference's device helpers
    W3 = td.WData(ctx, str(tmp_path / "def"), writevar="")
    W3.close()
    _, var, _, _, _ = read_wdata(str(tmp_path / "def.wtxt"))
    assert list(var) == ["rho_a", "rho_b", "delta", "j_a", "j_b"]
    td.lib().tdw_bind(ctx.h)
    h = np.ones(3 * N)
    assert td.lib().get_velocity_ext(dim_of(cfg), 0, 5, h.ctypes.data_as(C.c_void_p)) == 0
    assert not h.any()
    dl = np.zeros(2 * N)
    dptr = C.c_void_p(R.pots.data_ptr() + 2 * N * 8)
    assert td.lib().get_delta_ext(1, 5, dptr, dl.ctypes.data_as(C.c_void_p)) == 0
    assert not dl.any()
    td.lib().tdw_bind(None)
    ctx.close()


def dim_of(cfg):
    return int(cfg.dim)
