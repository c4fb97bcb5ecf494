"""Control-file keywords -> engine configuration (Src/RINPUT.c:48-563)."""
import math

import pytest

from meinv.control import ControlFileError, parse_text, to_config

BASE = """
lines =  6173.3433,  2.5
data_path = ./d.fits
wavelength_path = ./w.fits
"""


def test_defaults_match_reference_table():
    cfg = to_config(parse_text(BASE))
    assert (cfg.lambda0, cfg.geff) == (6173.3433, 2.5)
    assert (cfg.num_iter, cfg.num_run) == (60, 4)
    assert (cfg.convg_criteria, cfg.chisq_criteria, cfg.damp) == (1e-3, 1.0, 1e-3)
    assert list(cfg.weights) == [1.0] * 4
    assert (cfg.lambda_limits[0], cfg.lambda_limits[1]) == (1e-5, 1e3)
    assert [tuple(cfg.bounds[i]) for i in (0, 3, 4, 5, 6, 7, 8)] == \
        [(5, 6000), (-20, 20), (10, 80), (0.4, 0.8), (2, 90), (0, 1e6), (0.1, 0.85)]
    assert cfg.bounds[1][1] == pytest.approx(math.pi)
    # Btheta_Regularization = YES, 0.5, 1e-4 -> value in rad
    assert cfg.regul[1] == 1 and cfg.regul_value[1] == pytest.approx(0.5 * math.pi)
    assert cfg.regul_weight[1] == 1e-4 and sum(cfg.regul) == 1
    assert (cfg.vcoeffi, cfg.lcoeffi, cfg.icriteria, cfg.svd_threshold) == (100, 250, 100, 1e-6)
    assert (cfg.nfigures, cfg.hmi_ref, cfg.output_fit) == (6, 0, 0)
    assert list(cfg.step) == [1000., 2., 2., 2., 20., 0.3, 30., 1e5, 0.3]


def test_comments_overrides_and_units():
    text = BASE + """
num_run = 6      # comment after a value
weights = 1, 4, 4, 3
! Bmod_fixed = YES
*Btheta_fixed = YES
Bounds_Btheta = 0.1, 0.9
Bounds_Dopp = 100, 400
Btheta_Regularization = NO
Dopp_Regularization = YES, 25, 1e-3
HMI_REF = YES
output_fit = yes
Icriteria = 0.1
"""
    cf = parse_text(text)
    cfg = to_config(cf)
    assert cfg.num_run == 6
    assert list(cfg.weights) == [1, 4, 4, 3]
    assert sum(cfg.fixed) == 0          # both lines are comments
    assert cfg.bounds[1][0] == pytest.approx(0.1 * math.pi) and cfg.bounds[1][1] == pytest.approx(0.9 * math.pi)
    assert tuple(cfg.bounds[4]) == (100, 400)
    assert cfg.regul[1] == 0 and cfg.regul[4] == 1 and cfg.regul_value[4] == 25
    assert cfg.hmi_ref == 1 and cfg.output_fit == 1 and cfg.icriteria == 0.1


def test_errors_the_reference_treats_as_fatal():
    with pytest.raises(ControlFileError):
        parse_text("num_run = 3\n")
    with pytest.raises(ControlFileError):
        to_config(parse_text(BASE + "lambda_accept = 1.2\n"))
    with pytest.raises(ControlFileError):
        to_config(parse_text(BASE + "weights = 1, 2\n"))
    cf = parse_text(BASE + "not_a_key = 3\nnum_run = 2\nnum_run = 3\n")
    assert any("unknown" in w for w in cf.warnings) and any("redefined" in w for w in cf.warnings)
    assert to_config(cf).num_run == 3
