"""Device operator construction (metrics + coeff_p_block + densities + sour_p_block) against the
checker, which itself is bit-identical to the reference Fortran (test_oracle_golden.py)."""
import numpy as np
import pytest

from helpers import fields
from fortranutils_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,scale,ext", [("c0", 4, False), ("c4", 8, False), ("c1", 8, False),
                                           ("c1", 8, True), ("c3", 8, True)])
def test_build_operator_bitwise(cfg, scale, ext):
    import torch
    from fortranutils_amd import lmgx
    p = P.config(cfg, scale)
    o = O.Oracle(p)
    s = lmgx.Solver(lmgx.Topology.from_problem(p))
    fl = fields(p, o)
    geo = o.geometry()
    if ext:
        dirich = np.array([b.dirichlet for b in p.blocks], np.int32)
        s.build_operator(p, rho1=fl["rho"], dirichlet=dirich)
        ref = o.coeff(geo, rho=o.rho_levels(fl["rho"]), dirichlet=dirich)
    else:
        s.build_operator(p)
        ref = o.coeff(geo)
    got = s.get_operator()
    for k, x, y in zip("ap aw ae as an al ar".split(), got, ref):
        assert np.array_equal(x, y), k
    dev = lambda a: torch.from_numpy(a).cuda()
    b = torch.zeros(o.ntot, dtype=torch.float64, device="cuda")
    s.source(dev(fl["u"]), dev(fl["v"]), dev(fl["w"]), dev(fl["rho"]), dev(fl["rhonp1"]),
             p.dtime, b)
    bo = o.sour(geo, fl["u"], fl["v"], fl["w"], fl["rho"], fl["rhonp1"], p.dtime)
    assert np.array_equal(b.cpu().numpy()[: o.n1], bo[: o.n1])
