"""Set-up of the td-gauss-slda testsuite problem (testsuite/td-gauss-slda):
12^3 lattice, DX=1.15, SLDA, spherical cutoff, AB4AM5, UNIFORM_TEST_MODE
(Na = ceil(L^3/(6 pi^2)), Nb = Na+1), uniform start at T=0.05 eF, a Gaussian
V_ext switched on between t*eF = 5 and 55 (problem-definition.h v_ext and
process_params), params7=1 flips its sign for spin b."""
import json
import math
import os

import numpy as np

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))


def golden(functional=po.SLDA, force_a1=0):
    name = {po.BDG: "td_gauss_bdg.json", po.ASLDA: "td_gauss_aslda.json",
            po.SLDAE: "td_gauss_sldae_a1.json" if force_a1 else "td_gauss_sldae.json"}.get(
        functional, "td_gauss_slda.json")
    return json.load(open(os.path.join(HERE, "golden", name)))


def mu_rule_of(functional):
    """The uniform solver's chemical-potential update of the code that wrote
    each golden: the SLDA and BdG logs come from the 2021.11.06 code (the
    rule now kept commented at pca_uniform.h:425-426), the ASLDA log from
    the 2026.09.19 code (the active rule, pca_uniform.h:427-428)."""
    return 0 if functional in (po.ASLDA, po.SLDAE) else 1


def run_tag(functional, dim):
    """The golden run of each suite: td-gauss-sldae has only the
    spin-symmetric ss1 runs."""
    return ("ss1-%dd" if functional == po.SLDAE else "ss0-%dd") % dim


def setup(dim, scheme=po.AB4AM5, functional=po.SLDA, force_a1=0, mu_rule=None):
    """td-gauss-slda (default), td-gauss-aslda (functional=ASLDA: the
    uniform alpha_a, alpha_b go into the effective-mass slots of pots) or
    td-gauss-bdg (functional=BDG, scattering length aBdG, BdG uniform
    solution) problem."""
    G = golden(functional, force_a1)
    g = G["inputs"]
    NX, DXv = g["NX"], g["DX"]
    LXYZ = (NX * DXv) ** 3
    Na = math.ceil(1.0 / (6.0 * math.pi ** 2) * LXYZ)
    ss = int(g.get("spinsymmetry", 0))
    Nb = Na if ss else Na + 1          # UNIFORM_TEST_MODE, pca.c:198-199
    if functional == po.SLDAE:
        scheme = po.AB4AM5 if force_a1 else po.AB3AM4   # Makefile.1.s / Makefile.1.s.34
    cfg = po.make_cfg(dim, NX, NX, NX, dx=DXv, scheme=scheme, functional=functional, spinsymmetry=ss,
                      sclgth=g["aBdG"] if functional == po.BDG else g.get("sclgth", 0.0), force_a1=force_a1)
    kc = 0.999999 * math.pi / DXv
    u, rc = po.uniform_solve(cfg, Na / LXYZ, Nb / LXYZ, kc, g["temperature"], eps=g["init0eps"],
                             scmix=g["init0scmix"], muchange=g["init0muchange"],
                             maxiter=g["init0maxiter"],
                             mu_rule=mu_rule_of(functional) if mu_rule is None else mu_rule)
    assert rc == 0
    n = u.n0_a + u.n0_b
    eF = (3 * math.pi ** 2 * n) ** (2 / 3) / 2
    kF = (3 * math.pi ** 2 * n) ** (1 / 3)
    Effg = 0.6 * n * eF * LXYZ
    wf, En, kt, cnt = po.uniform_states(cfg, u)
    fb = np.array([po.fbeta(E, u.beta) for E in En])
    p = G["runs"][run_tag(functional, dim)]["params"]
    A = p[0] * eF

    def wcoef(s):
        s = s / kF
        return -0.5 / s ** 2 if s > 0.1 else 0.0

    cfg.vext = po.gauss_vext([(-A, (0, 0, 0), (wcoef(p[1]), wcoef(p[2]), wcoef(p[3])))],
                             switch=(p[4] / eF, p[5] / eF, p[6] / eF), flip_b=int(p[7] > 0.5))
    cfg.mu_a, cfg.mu_b, cfg.ec = u.mu_a, u.mu_b, u.ec
    cfg.dt = g["dt_eF"] / eF
    N = po.npts(cfg)
    pots = np.zeros(12 * N)
    pots[:N] = u.V_a
    pots[N:2 * N] = u.V_b
    pots[2 * N:4 * N:2] = u.delta
    if functional in (po.ASLDA, po.SLDAE):
        pots[4 * N:5 * N] = u.alph_a
        pots[5 * N:6 * N] = u.alph_b
    else:
        pots[4 * N:6 * N] = 1.0
    return dict(cfg=cfg, u=u, wf=wf, fbeta=fb, kt=kt, cnt=cnt, pots=pots, eF=eF, kF=kF,
                Effg=Effg, timesteps=g["timesteps"])


def wlog_row(e, Effg):
    """Columns 3..13 of the td-gauss logger (Na, Nb, N, Etot, Ekin, ...)."""
    et = e[:7].sum()
    return [e[7], e[8], e[7] + e[8], et / Effg] + [e[k] / Effg for k in range(7)]


def digits_match(mine, ref, sig=10):
    """True if mine equals ref when both are printed with %.{sig}g (1 ulp slack
    in the last printed digit)."""
    if ref == 0.0:
        return abs(mine) < 1e-20
    ulp = 10 ** (math.floor(math.log10(abs(ref))) - (sig - 1))
    return abs(mine - ref) <= 1.01 * ulp


def to_params(cfg, precision=0):
    """Library parameters (wslda_amd.td.Params) equal to an oracle Cfg."""
    import ctypes
    from wslda_amd import td
    p = td.Params()
    ctypes.memmove(ctypes.addressof(p), ctypes.addressof(cfg), ctypes.sizeof(cfg))
    p.precision = precision
    return p
