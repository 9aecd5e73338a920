"""The metric form of level 1's operator (MetF: the FP32 relaxation's couplings recomputed per
cell from the densities and the grid's face factors, LMGX_METRIC=1) restated in numpy, FP64, and
checked against the oracle's coefficients (themselves bitwise the reference's coeff_p_block with
the 1/rho_face scaling).  Every interior cell of level 1, boundary planes included:

* the face factors g are rdju/diu at (i, j) for the i-face between i-1 and i, rdiv/djv for the
  j-face between j-1 and j and didjp/rp for the k-faces, and a coupling is g dk 2 / (rho_a + rho_b)
  with dk = dkp(k) for i- and j-faces and 1/dkw(k) for the k-face between k-1 and k;
* a coupling across a side that is neither periodic nor an interface is zero, and on a
  face-Dirichlet side it is added to the diagonal instead; the diagonal is the sum of the
  couplings and that excess.

A wrong face factor (a neighbour's, a transposed metric), a wrong k factor or a wrong side rule
moves some coupling by far more than the 1e-13 allowed for the different rounding.
"""
import numpy as np
import pytest

from helpers import fields
from fortranutils_amd import problems as P
import oracle as O

M_DIU, M_DJV, M_RDIV, M_RDJU, M_DIDJP, M_RDIDJP, M_RP = range(7)
SWAP = (12, 13)        # PERIODIC, INTERFACE


def metric_form(o, geo, rho, p, mutate=None):
    """Per block: the seven level-1 coefficient arrays (natural layout, halos zero) of the metric
    form, as the kernels define it (mutate: a planted mistake, for the test of the check)."""
    g = o.g
    met, dkp, dkw = geo["met"], geo["dkp"], geo["dkw"]
    out = [np.zeros(o.n1) for _ in range(7)]
    for b, blk in enumerate(p.blocks):
        ni, nj, nk = (int(v) for v in o.nijkp1[0][b])
        o2, o1, off = g.off2d[0][b], g.off1d[0][b], int(o.off[0][b])
        m2 = lambda q: met[q, o2: o2 + ni * nj].reshape(nj, ni)
        with np.errstate(divide="ignore", invalid="ignore"):
            tE = m2(M_RDJU) / m2(M_DIU)
            tN = m2(M_RDIV) / m2(M_DJV)
            tR = m2(M_DIDJP) / m2(M_RP)
        kp = dkp[o1: o1 + nk]
        kw = 1.0 / dkw[o1: o1 + nk]
        r = rho[off: off + ni * nj * nk].reshape(nk, nj, ni)
        I = slice(1, ni - 1)
        J = slice(1, nj - 1)
        K = slice(1, nk - 1)
        rP = r[K, J, I]
        sh = lambda dk, dj, di: r[1 + dk: nk - 1 + dk, 1 + dj: nj - 1 + dj, 1 + di: ni - 1 + di]
        kpc = kp[K][:, None, None]
        c = {}
        c["w"] = tE[J, I][None] * kpc * 2.0 / (sh(0, 0, -1) + rP)
        c["e"] = (tE[J, I] if mutate == "east_factor" else tE[J, 2:])[None] * kpc * 2.0 / (rP + sh(0, 0, 1))
        c["s"] = tN[J, I][None] * kpc * 2.0 / (sh(0, -1, 0) + rP)
        c["n"] = (tN[J, I] if mutate == "north_factor" else tN[2:, I])[None] * kpc * 2.0 / (rP + sh(0, 1, 0))
        c["l"] = tR[J, I][None] * kw[K][:, None, None] * 2.0 / (sh(-1, 0, 0) + rP)
        c["r"] = tR[J, I][None] * (kp[2:] if mutate == "k_factor" else kw[2:])[:, None, None] * 2.0 / (rP + sh(1, 0, 0))
        d = np.zeros_like(rP)
        # side s (W E S N L R): the cells on its plane
        planes = {"w": (slice(None), slice(None), 0), "e": (slice(None), slice(None), -1),
                  "s": (slice(None), 0, slice(None)), "n": (slice(None), -1, slice(None)),
                  "l": (0, slice(None), slice(None)), "r": (-1, slice(None), slice(None))}
        for s, key in enumerate("wesnlr"):
            if blk.bc[s] in SWAP:
                continue
            pl = planes[key]
            if blk.dirichlet[s] and mutate != "no_excess":
                d[pl] += c[key][pl]
            c[key][pl] = 0.0
        ap = d + c["w"] + c["e"] + c["s"] + c["n"] + c["l"] + c["r"]
        for q, v in enumerate([ap, c["w"], c["e"], c["s"], c["n"], c["l"], c["r"]]):
            full = np.zeros((nk, nj, ni))
            full[K, J, I] = v
            out[q][off: off + ni * nj * nk] = full.ravel()
    return out


@pytest.mark.parametrize("cfg,scale", [("c0", 2), ("c1", 8), ("c4", 8), ("c3", 8), ("c2", 16)])
def test_metric_form_is_the_builders_operator(cfg, scale):
    p = P.config(cfg, scale)
    o = O.Oracle(p)
    geo = o.geometry()
    dirich = np.array([b.dirichlet for b in p.blocks], np.int32)
    if p.rho_kind == "uniform":
        rho_lv, rho1 = None, np.ones(o.ntot)
    else:
        rho_lv = o.rho_levels(fields(p, o)["rho"])
        rho1 = rho_lv
    ref = o.coeff(geo, rho=rho_lv, dirichlet=dirich)
    mf = metric_form(o, geo, rho1, p)
    for q in range(7):
        a, m = np.asarray(ref[q][: o.n1]), mf[q]
        scale_q = np.abs(a).max()
        err = np.abs(a - m).max() / scale_q
        assert err < 1e-13, (q, err)


@pytest.mark.parametrize("mutate", ["east_factor", "north_factor", "k_factor", "no_excess"])
def test_metric_form_check_sees_planted_mistakes(mutate):
    """The same comparison on c4 at 1/8 (two blocks, an interface, a stretched grid, a density
    layer, a face-Dirichlet outflow) fails for each planted mistake."""
    p = P.config("c4", 8)
    o = O.Oracle(p)
    geo = o.geometry()
    dirich = np.array([b.dirichlet for b in p.blocks], np.int32)
    rho_lv = o.rho_levels(fields(p, o)["rho"])
    ref = o.coeff(geo, rho=rho_lv, dirichlet=dirich)
    mf = metric_form(o, geo, rho_lv, p, mutate=mutate)
    err = max(np.abs(np.asarray(ref[q][: o.n1]) - mf[q]).max() / np.abs(ref[q]).max()
              for q in range(7))
    assert err > 1e-6, err
