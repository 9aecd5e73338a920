"""The binary16 couplings of the mixed-precision correction's FP32 line relaxation (k_make16 and
the H16 instances of k_zebra_seg / k_resid_cols) restated in numpy on the level-1 operator, and
the property the design rests on checked on the assembled matrices:

* every coupling is rounded once (times the level's power-of-two scale) and read by both cells it
  joins, and the diagonal is formed as the rounded diagonal excess d = ap - sum(a_nb) plus the
  rounded couplings, so the rounded operator is again symmetric with row sums d;
* it is then spectrally within a factor 1 +/- 2^-11 of the true one: every generalized eigenvalue
  of (A16, A) lies in [1 - 2^-11, 1 + 2^-11];
* rounding ap on its own instead breaks that bound by far, which is why the kernels form it.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from helpers import assemble_level1, fields
from fortranutils_amd import problems as P
import oracle as O

U = 2.0 ** -11          # unit roundoff of binary16


def _level1(cfg, scale):
    p = P.config(cfg, scale)
    o = O.Oracle(p)
    fl = fields(p, o)
    dirich = np.array([b.dirichlet for b in p.blocks], np.int32)
    rho = o.rho_levels(fl["rho"]) if p.rho_kind != "uniform" else None
    a = o.coeff(o.geometry(), rho=rho, dirichlet=dirich)
    return o, [np.asarray(x, np.float64) for x in a]


def _round16(a, form_diagonal=True):
    """What the line kernels see (divided by the scale again): binary16 couplings, and the
    diagonal either formed from them and the rounded excess d or rounded on its own."""
    ap, *nb = a                                    # nb: aw ae as an al ar, as k_make16 sums them
    d = ap - sum(nb)
    mx = max(float(np.abs(x).max()) for x in nb + [d])
    sc = 2.0 ** (14 - int(np.ceil(np.log2(mx))))   # largest value near 2^14 (k_range16 / host)
    r16 = lambda x: (x * sc).astype(np.float16).astype(np.float64) / sc
    h = [r16(x) for x in nb]
    aph = r16(d) + sum(h) if form_diagonal else r16(ap)
    return [aph] + h


@pytest.mark.parametrize("cfg,scale", [("c4", 16), ("c0", 4)])
def test_binary16_operator_is_spectrally_close(cfg, scale):
    o, a = _level1(cfg, scale)
    A = assemble_level1(o, a)[0].toarray()
    a16 = _round16(a)
    A16 = assemble_level1(o, a16)[0].toarray()
    assert np.array_equal(A16, A16.T)
    # row sums are the rounded excess: zero wherever the true operator's are
    d = a[0] - sum(a[1:])
    rows = np.abs(A @ np.ones(len(A)))
    zero = rows <= 1e-12 * np.abs(np.diag(A))
    assert zero.sum() > len(A) // 2
    assert np.all(np.abs(A16 @ np.ones(len(A)))[zero] <= 1e-12 * np.abs(np.diag(A16))[zero])
    lam = sl.eigh(A16, A, eigvals_only=True)
    dev = np.abs(lam - 1.0).max()
    print(cfg, "n", len(A), "max |lambda - 1| / 2^-11 =", dev / U, "excess cells", (d != 0).sum())
    assert dev <= U * (1 + 1e-6)
    assert dev > 0.05 * U                         # the rounding is there at all


def test_rounding_the_diagonal_on_its_own_breaks_the_bound():
    o, a = _level1("c4", 16)
    A = assemble_level1(o, a)[0].toarray()
    A16 = assemble_level1(o, _round16(a, form_diagonal=False))[0].toarray()
    lam = sl.eigh(A16, A, eigvals_only=True)
    dev = np.abs(lam - 1.0).max()
    print("ap rounded on its own: max |lambda - 1| / 2^-11 =", dev / U)
    assert dev > 4 * U
