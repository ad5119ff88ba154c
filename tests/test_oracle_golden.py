"""The oracle against the reference testsuite's own golden trajectories
(testsuite/td-gauss-{slda,bdg}/ref/test.ss0-*.wlog.ref, 10 significant
digits)."""
import pytest

from oracle import pyoracle as po
from tests import tdgauss

# columns of the golden rows: id, t_eF, Na, Nb, N, Etot, Ekin, Epot, Epair, Ecurr, Eext, ...
CHECKED = [(2, "Na"), (3, "Nb"), (4, "N"), (5, "Etot"), (6, "Ekin"), (7, "Epot"), (8, "Epair"), (10, "Eext")]


def run_and_compare(dim, tag, nmeas, functional=po.SLDA):
    s = tdgauss.setup(dim, functional=functional)
    P = po.Problem(s["cfg"], s["wf"], s["fbeta"], kt=s["kt"], cnt=s["cnt"], pots=s["pots"])
    P.initial_measure()
    ref = tdgauss.golden(functional)["runs"][tag]["rows"]
    it = 0
    bad = []
    for m in range(nmeas + 1):
        if m > 0:
            for k in range(s["timesteps"]):
                P.step(it, grad_last=(k == s["timesteps"] - 1))
                it += 1
        row = tdgauss.wlog_row(P.energy(it), s["Effg"])
        # row: Na, Nb, N, Etot, Ekin, Epot, Epair, Ecurr, Eext, Edext, Evext
        mine = {"Na": row[0], "Nb": row[1], "N": row[2], "Etot": row[3], "Ekin": row[4],
                "Epot": row[5], "Epair": row[6], "Eext": row[8]}
        for col, name in CHECKED:
            r = ref[m][col]
            if name == "Eext":
                # a small difference of large sums: compared on the scale of
                # the total energy (10 printed digits of Etot ~ 1e-10 absolute)
                ok = abs(mine[name] - r) < 2e-11
            else:
                ok = tdgauss.digits_match(mine[name], r)
            if not ok:
                bad.append((m, name, mine[name], r))
    assert not bad, bad[:10]


@pytest.mark.parametrize("functional", [po.SLDA, po.BDG])
def test_uniform_header_1d(functional):
    s = tdgauss.setup(1, functional=functional)
    h = tdgauss.golden(functional)["runs"]["ss0-1d"]["header"]
    assert abs(s["kF"] - float(h["kF"])) < 1e-5
    assert abs(s["eF"] - float(h["eF"])) < 1e-6
    assert abs(s["Effg"] - float(h["Effg"])) < 1e-4
    assert abs(s["u"].mu_a / s["eF"] - float(h["mu_a/eF"])) < 1e-6
    assert abs(s["u"].mu_b / s["eF"] - float(h["mu_b/eF"])) < 1e-6
    assert abs(s["u"].ec - float(h["E_cut"])) < 1e-4
    assert s["wf"].shape[0] // 2 == int(h["nwf"])


def test_td_gauss_1d_full_trajectory():
    """All 101 rows (28,600 ABM steps) of the 1D spin-imbalanced run."""
    run_and_compare(1, "ss0-1d", 100)


def test_td_gauss_2d_first_rows():
    """First 12 rows (through the switch-on of V_ext at t*eF=5..20)."""
    run_and_compare(2, "ss0-2d", 11)


@pytest.mark.slow
def test_td_gauss_3d_first_rows():
    run_and_compare(3, "ss0-3d", 2)


def test_td_gauss_bdg_1d_full_trajectory():
    """BdG functional: all 101 rows of the 1D spin-imbalanced run
    (testsuite/td-gauss-bdg, aBdG = -1.1)."""
    run_and_compare(1, "ss0-1d", 100, functional=po.BDG)


def test_td_gauss_bdg_2d_first_rows():
    run_and_compare(2, "ss0-2d", 11, functional=po.BDG)
