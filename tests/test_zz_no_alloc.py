"""The allocation-free soln_diff (pgmg_x_diff2_noalloc) against the
allocating one (pgmg_x_diff2), through the C ABI: bit for bit on two calls
in a row, the second on a smaller level, and a level too large for its
device array refused with a clear error."""
import numpy as np
import pytest

import pgmg_amd as P

pytestmark = pytest.mark.gpu


def test_x_diff2_noalloc_equals_x_diff2():
    m, n, o = 70, 45, 40
    print("case", (m, n, o), flush=True)
    with P.Hierarchy(m, n, o, smoother=P.SMOOTH_COLOUR8) as h:
        h.generate_problem(P.FAM_FACE7_UNIFORM, 21)
        h.compute_rats()
        h.v_cycle()  # non-zero x on every level
        rng = np.random.default_rng(22)
        for level in (0, 1, 0):  # the second call on a smaller level
            y = rng.uniform(-1, 1, size=h.shape(level))
            want = h.x_diff2(y, level)
            got = h.x_diff2_noalloc(y, level)
            assert got == want, (level, got, want)
            ref = float(np.sum((y - h.get_x(level)) ** 2))
            assert abs(got - ref) <= 1e-12 * ref


def test_x_diff2_noalloc_refuses_a_level_beyond_its_array():
    m, n, o = 257, 257, 128  # 8.45 M points: more than its 8 Mi doubles hold
    print("case", (m, n, o), flush=True)
    with P.Hierarchy(m, n, o, smoother=P.SMOOTH_COLOUR8, max_levels=2) as h:
        y = np.zeros(h.shape(0))
        with pytest.raises(P.PgmgError, match="allocation-free"):
            h.x_diff2_noalloc(y, 0)
        # a level that fits still works afterwards
        y1 = np.ones(h.shape(1))
        assert h.x_diff2_noalloc(y1, 1) == h.x_diff2(y1, 1)
