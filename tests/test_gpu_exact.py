"""Reference-order FP64 GPU path against the reference Fortran's golden outputs, through the C ABI.

Fields (f, residu, b on every level) must be bit-identical; the norm history, whose sums run in a
different order on the GPU, must agree to 1e-12 relative.
"""
import numpy as np
import pytest

from helpers import COEF, case, level_slice

pytestmark = pytest.mark.gpu

CASES = ["g1_c0s4", "g2_c4s8", "g3_c1s8", "g4_c0s4_f", "g5_c4s8_ext", "g11_c4s8_spor2",
         "g12_c1s8_spor2"]
GSUM_CASES = ["g1_c0s4", "g2_c4s8", "g3_c1s8", "g5_c4s8_ext", "g12_c1s8_spor2"]


def solver_for(p):
    from fortranutils_amd import lmgx
    return lmgx.Solver(lmgx.Topology.from_problem(p))


def rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_exact_solve_matches_reference(name):
    p, o, coeffs, b, g = case(name)
    s = solver_for(p)
    s.set_operator(coeffs)
    f = np.zeros(o.ntot)
    res = np.zeros(o.ntot)
    bb = b.copy()
    meth = int(g.get("meth_smooth", 1))
    print(f"case {name} ntot={o.ntot} meth_smooth={meth}", flush=True)
    hist = s.lmg_solve(f, bb, res, max_cycles=int(g["ncycles"]), tolf=0.0, cycle=int(g["cycle"]),
                       numpre=p.numpre, numpost=p.numpost, numcits=p.numcits, omega=p.omega,
                       omega_c=p.omega_c, smoother=0 if meth == 1 else 2)
    assert hist.shape == g["hist"].shape
    assert np.array_equal(f, g["f"])
    assert np.array_equal(res, g["res"])
    assert np.array_equal(bb, g["b"])
    assert rel(hist, g["hist"]) < 1e-12
    # the SAVEd sweep-order counter advanced as the reference's does (sporlx only)
    fo = np.zeros(o.ntot)
    o.solve(coeffs, fo, b.copy(), int(g["ncycles"]), cycle=int(g["cycle"]), meth_smooth=meth)
    assert s.jumble == o.jumble
    assert (o.jumble == 0) == (meth == 2)


@pytest.mark.parametrize("name", ["g2_c4s8", "g5_c4s8_ext"])
def test_device_entry_point_and_post_steps(name):
    import torch
    p, o, coeffs, b, g = case(name)
    s = solver_for(p)
    s.set_operator([torch.from_numpy(a).cuda() for a in coeffs])
    f = torch.zeros(o.ntot, dtype=torch.float64, device="cuda")
    res = torch.zeros_like(f)
    bb = torch.from_numpy(b).cuda()
    hist = s.lmg_solve(f, bb, res, max_cycles=int(g["ncycles"]), tolf=0.0, numpre=p.numpre,
                       numpost=p.numpost, numcits=p.numcits)
    assert np.array_equal(f.cpu().numpy(), g["f"])
    assert rel(hist, g["hist"]) < 1e-12
    s.extrap_p(f)
    s.p_ref_subtract(f)
    assert np.array_equal(f.cpu().numpy()[: o.n1], g["p_post"])


def test_c0_full_30_cycles():
    """c0: one 64x64x32 block, 4 levels, V(2,2), 30 cycles; history to 1e-12, p bit-exact."""
    import torch
    p, o, coeffs, b, g = case("g6_c0")
    s = solver_for(p)
    s.set_operator([torch.from_numpy(a).cuda() for a in coeffs])
    f = torch.zeros(o.ntot, dtype=torch.float64, device="cuda")
    res = torch.zeros_like(f)
    bb = torch.from_numpy(b).cuda()
    hist = s.lmg_solve(f, bb, res, max_cycles=30, tolf=0.0, numpre=2, numpost=2)
    assert rel(hist, g["hist"]) < 1e-12
    fh = f.cpu().numpy()
    assert np.array_equal(fh[g["p_idx"]], g["f_sample"])
    s.extrap_p(f)
    s.p_ref_subtract(f)
    pp = f.cpu().numpy()[: o.n1]
    assert np.array_equal(pp[g["p_idx"]], g["p_sample"])


def test_single_passes_match_oracle():
    """norms, sporlx (3 sweeps + swaps), restriction + zeroing, prolongation + swap, one by one."""
    import torch
    p, o, coeffs, b, g = case("g3_c1s8")
    s = solver_for(p)
    s.set_operator(coeffs)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(o.ntot)
    dev = lambda a: torch.from_numpy(a.copy()).cuda()
    # norms on every level
    for lvl in range(o.nlev):
        r_o = np.zeros(o.ntot)
        n_o = o.norms(lvl, coeffs, x, b, r_o)
        r_d = dev(np.zeros(o.ntot))
        n_d = s.norms(lvl, dev(x), dev(b), r_d)
        sl = level_slice(o, lvl)
        assert np.array_equal(r_d.cpu().numpy()[sl], r_o[sl])
        assert rel(n_d, n_o) < 1e-13
    # one sporlx call per level, same jumble
    for lvl in range(o.nlev):
        xo = x.copy()
        s.jumble = o.jumble
        o.sporlx(lvl, coeffs, xo, b, 1.6)
        xd = dev(x)
        s.sporlx(lvl, xd, dev(b), 1.6)
        assert np.array_equal(xd.cpu().numpy(), xo)
        assert s.jumble == o.jumble
    # restriction (+ zero coarse f), prolongation (+ coarse halo copy + swap)
    for lc in range(1, o.nlev):
        bo, fo = b.copy(), x.copy()
        o.restrict(lc, x, bo)
        fo[level_slice(o, lc)] = 0.0
        bd, fd = dev(b), dev(x)
        s.restrict(lc, dev(x), bd, fd)
        assert np.array_equal(bd.cpu().numpy(), bo)
        assert np.array_equal(fd.cpu().numpy(), fo)
    for lf in range(o.nlev - 1):
        fo = x.copy()
        o.prolong(lf, fo)
        fd = dev(x)
        s.prolong(lf, fd)
        assert np.array_equal(fd.cpu().numpy(), fo)
    # one sporlx2 call per level (7 natural-order sweeps + swaps; jumble untouched)
    for lvl in range(o.nlev):
        xo = x.copy()
        o.sporlx2(lvl, coeffs, xo, b, 1.6)
        xd = dev(x)
        j0 = s.jumble
        s.sporlx2(lvl, xd, dev(b), 1.6)
        assert np.array_equal(xd.cpu().numpy(), xo)
        assert s.jumble == j0
    # swap and global sum
    for lvl in range(o.nlev):
        xo = x.copy()
        o.swap(lvl, xo)
        xd = dev(x)
        s.swap(lvl, xd)
        assert np.array_equal(xd.cpu().numpy(), xo)
        gs_o = o.global_sum(lvl, x)
        gs_d = s.global_sum(lvl, dev(x))
        assert abs(gs_d[0] - gs_o[0]) <= 1e-12 * np.abs(x).sum()
        assert gs_d[1] == gs_o[1] and gs_d[2] == gs_o[2]


@pytest.mark.parametrize("name", GSUM_CASES)
def test_global_sum_matches_reference(name):
    """lmgx_global_sum_device on the reference's final f, b and residu of every level against
    the reference's own global_sum (fixture): max and min bit for bit, the sum (a fixed
    two-stage tree on the GPU, serial block order in the reference) to 1e-13 of the sum of
    magnitudes; and the sum against a numpy sum over the interior mask."""
    import torch
    from helpers import interior_mask
    p, o, coeffs, b, g = case(name)
    s = solver_for(p)
    for lvl in range(o.nlev):
        sl = level_slice(o, lvl)
        msk = interior_mask(o, lvl)
        for q, key in enumerate(("f", "b", "res")):
            fld = g[key]
            got = s.global_sum(lvl, torch.from_numpy(fld.copy()).cuda())
            want = g["gsum"][lvl, q]
            mag = np.abs(fld[sl][msk]).sum()
            assert got[1] == want[1] and got[2] == want[2], (lvl, key)
            assert abs(got[0] - want[0]) <= 1e-13 * mag, (lvl, key, got[0], want[0])
            assert abs(got[0] - fld[sl][msk].sum()) <= 1e-13 * mag


@pytest.mark.parametrize("cfg,scale", [("c3", 8), ("c4", 8), ("c0", 4)])
def test_p_mean_subtract_matches_oracle(cfg, scale):
    """Mean-pressure gauge on the device against the oracle's: the same mean to the rounding
    of two summation orders, the same shifted field (halos included), zero interior mean after."""
    import torch
    from fortranutils_amd import problems as P
    from helpers import interior_mask
    import oracle as O
    p = P.config(cfg, scale)
    o = O.Oracle(p)
    s = solver_for(p)
    rng = np.random.default_rng(17)
    x = 5.0 + rng.standard_normal(o.n1)
    xo = x.copy()
    m_o = o.p_mean_subtract(xo)
    xd = torch.from_numpy(x.copy()).cuda()
    m_d = s.p_mean_subtract(xd)
    assert abs(m_d - m_o) <= 1e-14 * abs(m_o)
    got = xd.cpu().numpy()
    assert np.max(np.abs(got - xo)) <= 1e-14 * np.abs(x).max()
    assert np.array_equal(got, x - m_d)
    msk = interior_mask(o, 0)
    assert abs(got[msk].mean()) <= 1e-14 * np.abs(x).max()
