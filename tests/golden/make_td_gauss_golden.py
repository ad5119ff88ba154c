"""Extract the td-gauss-slda, td-gauss-aslda, td-gauss-bdg and
td-gauss-sldae(-a1) golden trajectories into small JSON fixtures.

Reads the reference testsuite's own expected outputs (data, not code):
  testsuite/td-gauss-{slda,aslda,bdg}/ref/test.ss0-{1d,2d,3d}.wlog.ref
  testsuite/td-gauss-{slda,aslda,bdg}/ref/test.ss0-{1d,2d,3d}.cmp.ref
and the test inputs (testsuite/td-gauss-{slda,aslda,bdg}/input.txt.{1,2,3},
predefines.h: 12^3 lattice, DX=1.15, SLDA, ASLDA or BDG, spherical cutoff, AB4AM5
default, UNIFORM_TEST_MODE, params7=1 appended by test.desc for the ss0
runs; aBdG = -1.1 is the scattering length of the BDG runs).

Run in a tree where the reference is mounted; the JSON it writes are the
committed fixtures tests/golden/td_gauss_{slda,aslda,bdg}.json.
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def rows(path):
    out = []
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        t = line.split()
        # columns: id, t*eF, Na, Nb, N, Etot, Ekin, Epot, Epair, Ecurr, Eext, Edext, Evext, rt, date
        out.append([int(t[0])] + [float(x) for x in t[1:13]])
    return out


def header(path):
    h = {}
    for line in open(path):
        if line.startswith("#") and "=" in line:
            k, v = line[1:].split("=", 1)
            h[k.strip()] = v.strip()
    return h


def final_values(path):
    """test.*.cmp.ref: name, value, tolerance of the run's final state."""
    out = {}
    for line in open(path):
        t = line.split()
        if len(t) == 3:
            out[t[0]] = [float(t[1]), float(t[2])]
    return out


def params(path, flip):
    pr = [0.0] * 8
    for line in open(path):
        t = line.split()
        if t and t[0].startswith("params"):
            pr[int(t[0][6:])] = float(t[1])
    if flip:
        pr[7] = 1.0  # test.desc appends "params7 1" for the ss0 runs
    return pr


# td-gauss-sldae: the spin-symmetric ss1 runs (Makefile.1.s.34:
# SPINSYMMETRY_MODE, AB3AM4), N_a = N_b, sclgth = -1.1, no sign flip
for suite, functional, ss in (("slda", "SLDA", "ss0"), ("aslda", "ASLDA", "ss0"), ("bdg", "BDG", "ss0"),
                              ("sldae", "SLDAE", "ss1"), ("sldae-a1", "SLDAE", "ss1")):
    data = {"columns": ["id", "t_eF", "Na", "Nb", "N", "Etot/Effg", "Ekin/Effg", "Epot/Effg",
                        "Epair/Effg", "Ecurr/Effg", "Eext/Effg", "Edext/Effg", "Evext/Effg"],
            "inputs": {"NX": 12, "NY": 12, "NZ": 12, "DX": 1.15, "functional": functional,
                       "regularization": "SPHERICAL_CUTOFF", "scheme": "AB4AM5",
                       "temperature": 0.05, "init0Tstart": 0.05, "init0muchange": 0.5,
                       "init0scmix": 0.5, "init0eps": 1e-9, "init0maxiter": 10000,
                       "dt_eF": 0.0035, "timesteps": 286, "measurements": 100, "aBdG": -1.1,
                       },
            "runs": {}}
    if ss == "ss1":
        # td-gauss-sldae: Makefile.1.s.34 (AB3AM4); td-gauss-sldae-a1:
        # Makefile.1.s (default AB4AM5) with SLDA_FORCE_A1 in predefines.h
        data["inputs"].update({"scheme": "AB4AM5" if suite == "sldae-a1" else "AB3AM4", "spinsymmetry": 1,
                               "sclgth": -1.1, "force_a1": int(suite == "sldae-a1")})
    for d, tag in ((1, ss + "-1d"), (2, ss + "-2d"), (3, ss + "-3d")):
        p = os.path.join(REF, "testsuite/td-gauss-%s/ref/test.%s.wlog.ref" % (suite, tag))
        inp = os.path.join(REF, "testsuite/td-gauss-%s/input.txt.%d" % (suite, d))
        c = os.path.join(REF, "testsuite/td-gauss-%s/ref/test.%s.cmp.ref" % (suite, tag))
        data["runs"][tag] = {"header": header(p), "rows": rows(p), "params": params(inp, ss == "ss0"),
                             "final": final_values(c)}
    out = os.path.join(HERE, "td_gauss_%s.json" % suite.replace("-", "_"))
    json.dump(data, open(out, "w"), indent=0)
    print("wrote", out)
