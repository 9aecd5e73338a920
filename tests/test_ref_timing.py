"""The reference driver's timing-only mode, which bench.py's CPU baseline uses on every x-chain of
the workload at once (oracle/ref_io.run_timing_concurrent): it must run the same solve as the
driver's full mode.  Runs on the CPU; skipped where the reference driver was not built."""
import numpy as np
import pytest

import ref_io
from fortranutils_amd import problems as P

pytestmark = pytest.mark.skipif(not ref_io.available(),
                                reason="reference driver not built (needs /root/reference)")


def test_timing_mode_runs_the_same_solve():
    """Two x-chains of c3 at 1/8, timed concurrently: each history equals, bitwise, the history
    of the reference driver's full-output run of the same chain (the reference's own coeff_p, the same
    seeded source), and the solve times are positive."""
    p = P.config("c3", 8)
    chains = [[0, 1], [2, 3, 4]]
    subs = [P.subproblem(p, c) for c in chains]
    rng = np.random.default_rng(3)
    b1s = [rng.standard_normal(sum(int(np.prod(b.nijkp1)) for b in q.blocks)) for q in subs]
    out = ref_io.run_timing_concurrent(subs, ncycles=2, b1s=b1s)
    assert len(out) == 2
    for q, b1, (t, hist) in zip(subs, b1s, out):
        full = ref_io.run(q, ncycles=2, b1=b1, n1=b1.size, do_post=0)
        assert t > 0.0
        assert hist.shape == full["hist"].shape
        assert np.array_equal(hist, full["hist"])
