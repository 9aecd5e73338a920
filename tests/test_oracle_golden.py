"""The CPU restatement (oracle/lmg_oracle.c) against the reference Fortran's own outputs.

Every fixture was written by the reference compiled from its sources (tests/golden/make_golden.py).
The restatement follows the reference's expression order, so fields must agree bit for bit.
"""
import numpy as np
import pytest

from helpers import COEF, case, load

CASES = ["g1_c0s4", "g2_c4s8", "g3_c1s8", "g4_c0s4_f", "g5_c4s8_ext", "g11_c4s8_spor2",
         "g12_c1s8_spor2"]


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
    hist, res = o.solve(coeffs, f, b, int(g["ncycles"]), cycle=int(g["cycle"]),
                        meth_smooth=int(g.get("meth_smooth", 1)))
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


def test_sporlx2_leaves_jumble_alone():
    """sporlx2 (meth_smooth = 2) sweeps in natural order and never advances sporlx's SAVEd
    counter; the fixtures g11/g12 pin its fields (test_solve_bitwise)."""
    p, o, coeffs, b, g = case("g11_c4s8_spor2")
    o.jumble = 26
    f = np.zeros(o.ntot)
    o.solve(coeffs, f, b, 1, meth_smooth=2)
    assert o.jumble == 26


def _ordered_block_sums(o, field, level):
    """(sum, max, min) over each block's interior, summed one value after the other in the
    reference's loop order (k outer, j, i inner; global_sum.F:130-144)."""
    out = []
    for bl in range(o.nblk):
        n = o.nijkp1[level, bl]
        a = o.off[level, bl]
        blk = field[a: a + int(np.prod(n))].reshape(n[2], n[1], n[0])[1:-1, 1:-1, 1:-1]
        s = 0.0
        for v in blk.ravel().tolist():
            s = s + v
        out.append((s, float(blk.max()), float(blk.min())))
    return out


@pytest.mark.parametrize("name", ["g1_c0s4", "g2_c4s8", "g3_c1s8", "g5_c4s8_ext",
                                  "g12_c1s8_spor2"])
def test_global_sum_is_the_references(name):
    """The oracle's global_sum of the reference's final f, b and residu on every level equals
    the reference's own global_sum bit for bit (fixtures written by global_sum.F run on those
    fields); the reference's per-block global_sum_block results are the ordered interior sums,
    maxima and minima of each block, and the total is their sum in block order."""
    p, o, coeffs, b, g = case(name)
    for level in range(o.nlev):
        for q, key in enumerate(("f", "b", "res")):
            got = o.global_sum(level, g[key])
            assert np.array_equal(got, g["gsum"][level, q]), (level, key, got)
            blk = _ordered_block_sums(o, g[key], level)
            assert np.array_equal(np.array(blk), g["gsum_blk"][level, :, q]), (level, key)
            tot = 0.0
            for s, _, _ in blk:
                tot = tot + s
            assert tot == g["gsum"][level, q, 0]
            assert max(m for _, m, _ in blk) == g["gsum"][level, q, 1]
            assert min(m for _, _, m in blk) == g["gsum"][level, q, 2]
