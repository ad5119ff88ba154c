"""bench.py's command line and byte model, on the CPU: more than one GPU is
refused with a message (this build runs one GPU), and the per-pass byte
units add up to the step totals the counters measure (FP64 20 S and split
13 S per state-step at AB5AM5; compulsory 13 S and 9 S)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def step_sum(units, scheme, split):
    # one fused step: predictor, two density passes, two yz passes, corrector
    return (units("f_pred", scheme, split) + 2 * units("densities", scheme, split)
            + 2 * units("f_hyz", scheme, split) + units("f_corr", scheme, split))


def test_design_units_add_up_to_the_measured_step():
    assert step_sum(bench.design_units, 55, False) == 20
    assert step_sum(bench.design_units, 55, True) == 13


def test_algorithmic_units_add_up_to_the_compulsory_step():
    assert step_sum(bench.algorithmic_units, 55, False) == bench.step_units(55, False) == 13
    assert step_sum(bench.algorithmic_units, 55, True) == bench.step_units(55, True) == 9


def test_aslda_units_add_up():
    """The ASLDA step at AB5AM5: the folded step (one predictor and one
    corrector pass, two density and two H pass pairs) moves 27 S per
    state-step, its split form 16 S, the modular one 35 S."""
    u = bench.modular_units
    pair = 2 * (u("dens_gx") + u("dens_gyz") + u("ah_x") + u("ah_yz"))
    assert u("a_pred") + u("a_corr") + pair == 27
    assert u("abm_step1") + u("abm_step4q") + 2 * (u("norms") + u("scale")) + u("sub_qpe") + pair == 35
    # split precision: FP32 history, copies, H and T
    us = lambda k: bench.modular_units(k, True)
    assert us("a_pred") + us("a_corr") + 2 * (us("dens_gx") + us("dens_gyz") + us("ah_x") + us("ah_yz")) == 16
