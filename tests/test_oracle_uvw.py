"""Velocity correction uvw_adv (reference code/uvw_adv.F, code/uvw_adv_block.F) in the oracle.

* Bit for bit against the reference Fortran's own output (fixtures g7/g8 written by
  tests/golden/make_golden.py: 2 and 4 blocks with interfaces, periodic z and walls, rho^{n+1}
  differing from rho).
* Closed form on a uniform grid: with p linear along the component's axis and a constant density,
  every corrected face gets u_int - 1.5 dt g / rho, and the first face of a wall side keeps u_int.
* What the correction is for: after a converged solve with the reference's operator (rho = 1),
  the source of the corrected velocities (sour_p) is the constant that makes the source
  compatible, so the discrete divergence of the corrected field is uniform to solver precision
  (one projection step of the LES time step, code/step.F:353-425).
"""
import numpy as np
import pytest

from helpers import load, fields
from fortranutils_amd import problems as P
import oracle as O

SEED_NOISE = 7


@pytest.mark.parametrize("name", ["g7_c4s8_uvw", "g8_c1s8_uvw"])
def test_uvw_adv_bitwise_reference(name):
    g = load(name)
    p = P.config(str(g["cfg"]), int(g["scale"]))
    o = O.Oracle(p)
    geo = o.geometry()
    fl = fields(p, o)
    rhonp1 = g["rhonp1"]
    # the fixture's rho^{n+1} is rho times seeded noise; check the inputs are the same
    rng = np.random.default_rng(SEED_NOISE)
    assert np.array_equal(rhonp1, fl["rho"] * (1.0 + 0.1 * rng.random(fl["rho"].size)))
    for iv, key, src in [(1, "u_post", "u"), (2, "v_post", "v"), (3, "w_post", "w")]:
        u = fl[src].copy()
        o.uvw_adv(iv, geo, u, fl[src], rhonp1, g["p_post"], p.dtime)
        assert np.array_equal(u, g[key]), key
        assert np.count_nonzero(u != fl[src]) > 0.5 * u.size


def _uniform(nx=8, ny=6, nz=4, lx=2.0, ly=1.5, lz=1.0):
    yv = np.linspace(0.0, ly, ny + 1)
    zv = np.linspace(0.0, lz, nz + 1)
    blocks = P._chain_x([(nx, ny, nz)], 0.0, lx, yv, zv, 1)
    return P.Problem("uniform", blocks, nummgl=2, rho_kind="mixing"), (lx / nx, ly / ny, lz / nz)


@pytest.mark.parametrize("ivar", [1, 2, 3])
def test_uvw_adv_closed_form(ivar):
    p, h = _uniform()
    o = O.Oracle(p)
    geo = o.geometry()
    b = p.blocks[0]
    n0, n1, n2 = b.nijkp1
    k, j, i = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    grad = 0.75
    coord = (i, j, k)[ivar - 1] * h[ivar - 1]          # cell-centre spacing h along the axis
    pfield = (grad * coord).ravel().astype(np.float64)
    rho = np.full(o.n1, 1.6)
    rng = np.random.default_rng(1)
    ui = rng.standard_normal(o.n1)
    u = ui.copy()
    dt = 0.01
    o.uvw_adv(ivar, geo, u, ui, rho, pfield, dt)
    U, UI = u.reshape(n2, n1, n0), ui.reshape(n2, n1, n0)
    # loop range: interior cells, the first face skipped on a wall (lower side not periodic /
    # interface); the problem is periodic in z only
    lo = [2, 2, 2]
    side = (0, 2, 4)[ivar - 1]
    if b.bc[side] not in (12, 13):
        lo[ivar - 1] = 3
    sl = (slice(lo[2] - 1, n2 - 1), slice(lo[1] - 1, n1 - 1), slice(lo[0] - 1, n0 - 1))
    want = UI[sl] - 1.5 * dt * grad / 1.6
    assert np.allclose(U[sl], want, rtol=0, atol=1e-13)
    # the first face on a wall keeps u_int
    if lo[ivar - 1] == 3:
        first = [slice(1, -1)] * 3
        first[3 - ivar] = slice(1, 2)
        assert np.array_equal(U[tuple(first)], UI[tuple(first)])


@pytest.mark.parametrize("cfg,scale", [("c4", 8), ("c1", 8)])
def test_projection_makes_divergence_uniform(cfg, scale):
    p = P.config(cfg, scale)
    o = O.Oracle(p)
    geo = o.geometry()
    fl = fields(p, o)
    for c in "uvw":
        o.swap(0, fl[c])     # u* carries swapped halos, as in LESter: one value per shared face
    one = np.ones(o.n1)
    a = o.coeff(geo)                                   # reference operator (rho = 1)
    s0 = o.sour(geo, fl["u"], fl["v"], fl["w"], one, one, p.dtime)
    # blocks joined by interfaces form one Neumann problem each (c1: two x-chains); each needs
    # a source of zero sum, so the mean is removed per chain
    comp = list(range(o.nblk))

    def root(x):
        while comp[x] != x:
            x = comp[x]
        return x
    for bl, blk in enumerate(p.blocks):
        for side in range(6):
            if blk.bc[side] == 13:
                comp[root(bl)] = root(blk.bl2bl[side] - 1)
    ids = np.full(o.n1, -1)
    for bl, off in enumerate(o.off[0]):
        n = o.nijkp1[0, bl]
        m = np.full((n[2], n[1], n[0]), -1)
        m[1:-1, 1:-1, 1:-1] = root(bl)
        ids[off: off + m.size] = m.ravel()
    mask = ids >= 0
    mean = np.zeros(o.n1)
    for c in set(ids[mask]):
        mean[ids == c] = s0[: o.n1][ids == c].mean()
    b = s0.copy()
    b[: o.n1] -= mean                                  # compatible right-hand side
    mean = mean[mask]
    f = np.zeros(o.ntot)
    hist, _ = o.solve(a, f, b, 60)
    assert hist[-1, 0, 0] < 1e-11 * hist[0, 0, 1]
    o.extrap_p(f)
    o.p_ref_subtract(f)
    new = {}
    for iv, c in [(1, "u"), (2, "v"), (3, "w")]:
        new[c] = fl[c].copy()
        o.uvw_adv(iv, geo, new[c], fl[c], one, f, p.dtime)
    s1 = o.sour(geo, new["u"], new["v"], new["w"], one, one, p.dtime)[: o.n1][mask]
    scale_s = np.abs(s0[: o.n1][mask]).max()
    assert np.abs(s1 - mean).max() < 1e-8 * scale_s
    # and without the correction the divergence is far from uniform
    assert np.abs(s0[: o.n1][mask] - mean).max() > 1e-2 * scale_s
