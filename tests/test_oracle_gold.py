"""The CPU oracle against the reference's own regression gold
(test/test_input_output/gold, written by test/run_tests.sh:15-23 with
transport_test.cpp / adsorption_test.cpp settings, compared at max-abs 1e-10
by scripts/compare_files.py:16-19)."""
import numpy as np
import pytest

import _oracle as O

# test/transport_test.cpp:9-16 and test/adsorption_test.cpp:10-17
SETTINGS = {"transport": dict(dx=1.0e-9, Fbody=1.0e-8, T=300.0, P=2.0e6),
            "adsorption": dict(dx=12.0e-9, Fbody=1.0e-8, T=300.0, P=2.0e6)}


@pytest.mark.parametrize("kind", ["transport", "adsorption"])
@pytest.mark.parametrize("fast", [False, True])
def test_oracle_reproduces_gold(kind, fast, fixture_geometry, gold):
    pore, kn, lp = fixture_geometry
    o = O.Oracle(pore, kn, lp, fast=fast, **SETTINGS[kind])
    assert o.ncell == 3198
    steps = o.run(kind, 100, 1e-10)
    assert steps == 100
    ux, uy, rho = o.complete_domain()
    for name, a in (("ux", ux), ("uy", uy), ("rho", rho)):
        g = gold[(kind, name)]
        err = np.abs(a - g).max()
        assert err <= 1e-10, (kind, name, err)           # the reference's own criterion
        # gold holds 15 significant digits: the oracle sits at print precision
        assert err <= 1e-14 * max(1.0, np.abs(g).max()), (kind, name, err)


def test_transport_residual_trace(fixture_geometry):
    """Residual trace of the transport regression (step 0 and step 99)."""
    pore, kn, lp = fixture_geometry
    o = O.Oracle(pore, kn, lp, **SETTINGS["transport"])
    o.run("transport", 100, 1e-10)
    h = o.diff_history()
    assert len(h) == 100
    assert abs(h[0] - 0.466825156246162) < 1e-14
    assert abs(h[99] - 0.00151596085691759) < 1e-15


def test_adsorption_first_residual_clamped(fixture_geometry):
    """calculateDensityDifference clamps a zero first residual to 1e-5 (coreLB.hpp:1280)."""
    pore, kn, lp = fixture_geometry
    o = O.Oracle(pore, kn, lp, **SETTINGS["adsorption"])
    o.run("adsorption", 3, 1e-10)
    h = o.diff_history()
    assert h[0] == 1e-5 and h[1] > 1e-5


def test_break_on_tolerance(fixture_geometry):
    """run() stops when diff < tol without advancing t (mantiisApp.hpp:302-303)."""
    pore, kn, lp = fixture_geometry
    o = O.Oracle(pore, kn, lp, **SETTINGS["transport"])
    n = o.run("transport", 100, 0.3)
    h = o.diff_history()
    assert n == len(h) and h[-1] < 0.3 and (h[:-1] >= 0.3).all()
