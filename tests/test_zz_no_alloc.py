"""The allocation-free shell statistics (psdns_shell_stats_noalloc) against
psdns_shell_stats through the C ABI: the same kernels on a fixed device
array instead of a buffer allocated on first use, so the results agree bit
for bit.  Two calls in a row share the array, the second on a smaller grid
(fewer partial rows), so stale rows of the first call must not leak in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_shell_stats_noalloc_matches_shell_stats(O, P):
    for n, prec in ((64, 8), (32, 4)):
        print("[gpu case] shell_stats_noalloc", n, prec, flush=True)
        s = P.Solver(n, prec, viscos=0.013)
        try:
            s.put_state(O.Grid(n).field(seed=7 + n))
            got = s.shell_stats(n, noalloc=True)
            want = s.shell_stats(n)
            again = s.shell_stats(n // 2, noalloc=True)
        finally:
            s.close()
        assert want[2] > 0 and want[3] > 0
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        assert got[2] == want[2] and got[3] == want[3]
        assert np.array_equal(again[0], want[0][: n // 2]) and again[2] == want[2]
