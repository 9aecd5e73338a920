"""The CPU restatement (oracle/lmg_oracle.c) against the reference Fortran's own outputs.

Every fixture was written by the reference compiled from its sources (tests/golden/make_golden.py).
The restatement follows the reference's expression order, so fields must agree bit for bit.
"""
import numpy as np
import pytest

from helpers import COEF, case, load

CASES = ["g1_c0s4", "g2_c4s8", "g3_c1s8", "g4_c0s4_f", "g5_c4s8_ext"]


@pytest.mark.parametrize("name", ["g1_c0s4", "g2_c4s8", "g3_c1s8"])
def test_geometry_coefficients_source_bitwise(name):
    p, o, coeffs, b, g = case(name)
    geo = o.geometry()
    for k in ("x", "y", "z", "dkp", "dkw"):
        assert np.array_equal(geo[k], g[k]), k
    for r, k in enumerate(("diu", "djv", "rdiv", "rdju", "didjp", "rdidjp", "rp")):
        assert np.array_equal(geo["met"][r], g["met_" + k]), k
    for k, a in zip(COEF, coeffs):
        assert np.array_equal(a, g[k]), k
    assert np.array_equal(b[: o.n1], g["b"][: o.n1])


def test_tables_match_reference():
    for name in CASES:
        p, o, coeffs, b, g = case(name)
        assert np.array_equal(o.nijkp1, g["nijkp1"])
        assert np.array_equal(o.off, g["mg_off"])
        assert np.array_equal(o.npts, g["npts_all"])


@pytest.mark.parametrize("name", CASES)
def test_solve_bitwise(name):
    p, o, coeffs, b, g = case(name)
    f = np.zeros(o.ntot)
    hist, res = o.solve(coeffs, f, b, int(g["ncycles"]), cycle=int(g["cycle"]))
    assert np.array_equal(f, g["f"])
    assert np.array_equal(res, g["res"])
    assert np.array_equal(b, g["b"])
    assert np.array_equal(hist, g["hist"])
    o.extrap_p(f)
    o.p_ref_subtract(f)
    assert np.array_equal(f[: o.n1], g["p_post"])


def test_c0_full_history_and_pressure():
    """c0 (one 64x64x32 block) at full size: 30 V(2,2) cycles, 4 levels, Dirichlet outflow."""
    p, o, coeffs, b, g = case("g6_c0")
    f = np.zeros(o.ntot)
    hist, res = o.solve(coeffs, f, b, 30)
    assert np.array_equal(hist, g["hist"])
    assert np.array_equal(f[g["p_idx"]], g["f_sample"])
    o.extrap_p(f)
    o.p_ref_subtract(f)
    pp = f[: o.n1]
    assert np.array_equal(pp[g["p_idx"]], g["p_sample"])
    assert pp.sum() == g["p_sum"]


def test_jumble_persists_across_calls():
    """One call with 6 cycles equals 6 calls with 1 cycle (sporlx's SAVEd counter)."""
    p, o, coeffs, b, g = case("g1_c0s4")
    f = np.zeros(o.ntot)
    for _ in range(6):
        o.solve(coeffs, f, b, 1)
    assert np.array_equal(f, g["f"])
    # per cycle on 3 levels: numpre on levels 1-2, numcits on level 3, numpost on levels 2-1
    calls = 2 * p.numpre + p.numcits + 2 * p.numpost
    assert o.jumble == 13 * calls * 6
