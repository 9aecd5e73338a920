"""The allocation-free mean-pressure gauge (lmgx_p_mean_subtract_device_noalloc, static device
scratch) against the allocating entry point (lmgx_p_mean_subtract_device): bitwise the same
field and mean on two calls in a row, the second on a smaller problem (the scratch keeps the
first call's partials, which the second must not read), and a clear error for a level 1 that
needs more partials than the scratch holds."""
import numpy as np
import pytest

from fortranutils_amd import problems as P

pytestmark = pytest.mark.gpu


def _solver(p):
    from fortranutils_amd import lmgx
    return lmgx.Solver(lmgx.Topology.from_problem(p))


@pytest.mark.parametrize("first,second", [(("c1", 1), ("c4", 8)), (("c3", 4), ("c0", 4))])
def test_noalloc_mean_gauge_matches_the_allocating_one(first, second):
    import torch
    for cfg, scale in (first, second):
        p = P.config(cfg, scale)
        s = _solver(p)
        print(f"no-alloc gauge: {cfg}/{scale}, n1 = {s.n1}", flush=True)
        g = torch.Generator(device="cuda").manual_seed(5)
        x = 3.0 + torch.randn(s.n1, generator=g, device="cuda", dtype=torch.float64)
        a, b = x.clone(), x.clone()
        m_a = s.p_mean_subtract(a)
        m_b = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        s.p_mean_subtract_noalloc(b, m_b)
        assert torch.equal(a, b)
        assert m_b.item() == m_a
        c = x.clone()
        s.p_mean_subtract_noalloc(c)           # no mean wanted
        assert torch.equal(a, c)
        s.close()


def test_noalloc_rejects_a_level_beyond_the_scratch():
    """Three 1024 x 1024 x 4 blocks: 3 x 16 x 256 = 12,288 column tiles on level 1, beyond the
    scratch's 8192 partials; the call fails before any launch and leaves p alone."""
    import torch
    from fortranutils_amd import lmgx
    n = 3
    topo = lmgx.Topology(np.array([[1026, 1026, 6]] * n), np.ones((n, 6), int),
                         np.zeros((n, 6), int), np.zeros((n, 6), int), nlev=1)
    s = lmgx.Solver(topo)
    x = torch.ones(s.n1, dtype=torch.float64, device="cuda")
    with pytest.raises(lmgx.LmgxError, match="12288 workgroup partials"):
        s.p_mean_subtract_noalloc(x)
    assert bool((x == 1.0).all())
