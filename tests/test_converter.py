"""Sampler-to-physical parameter conversion of the likelihood caller
(behaviour of pygpuAAK/convert.py, with whole-name key matching)."""
import numpy as np
import pytest

from emri_kludge_suite_amd.convert import PARAMS, Converter

KEYS = ["iota", "s", "p", "e", "ln_M", "ln_mu", "gamma", "psi", "alph", "cos_theta_S", "phi_S", "cos_theta_K",
        "phi_K", "D"]


def test_conversions_and_order():
    c = Converter("emri", KEYS)
    x = np.array([0.5, 0.3, 9.0, 0.2, np.log(1e6), np.log(10.0), 1.0, 2.0, 3.0, 0.4, 4.0, -0.2, 5.0, 1.5])[:, None]
    phys = c.physical_order(c.convert(x.copy()))
    assert phys.shape == (14, 1)
    np.testing.assert_allclose(phys[:, 0], [0.5, 0.3, 9.0, 0.2, 1e6, 10.0, 1.0, 2.0, 3.0, np.arccos(0.4), 4.0,
                                            np.arccos(-0.2), 5.0, 1.5])


def test_periodic_angles_wrap_in_place():
    c = Converter("emri", KEYS)
    x = np.zeros((14, 3))
    x[6] = [-1.0, 7.0, 2.0]     # gamma
    x[10] = [-0.5, 13.0, 1.0]   # phi_S
    x[8] = [7.0, -7.0, 0.1]     # alph
    x[0] = [9.0, 9.0, 9.0]      # iota: not periodic here
    c.recycle(x)
    for row in (6, 8, 10):
        assert np.all((x[row] >= 0) & (x[row] < 2 * np.pi))
    np.testing.assert_allclose(x[6], [2 * np.pi - 1.0, 7.0 - 2 * np.pi, 2.0])
    assert np.all(x[0] == 9.0)


def test_key_order_is_free_and_aliases_resolve():
    keys = ["D", "alpha", "a"] + [k for k in PARAMS if k not in ("D", "alph", "s")]
    c = Converter("emri", keys)
    x = np.arange(14, dtype=float)[:, None]
    phys = c.physical_order(x.copy())
    assert phys[PARAMS.index("D"), 0] == 0 and phys[PARAMS.index("alph"), 0] == 1 and phys[PARAMS.index("s"), 0] == 2


def test_unknown_or_missing_keys_are_errors():
    with pytest.raises(ValueError):
        Converter("emri", KEYS[:-1] + ["distance_pc"])
    c = Converter("emri", KEYS[:-1])
    with pytest.raises(ValueError):
        c.physical_order(np.zeros((13, 1)))
    with pytest.raises(ValueError):
        Converter("mbh", KEYS)
