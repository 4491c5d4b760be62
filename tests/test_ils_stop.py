"""The reference ILS's target stop rule as an integer bound (no GPU):
ILSGlobalSolverSequential.cpp:97-99 loops while solution * (1.0f + error) <
best, in float; logo2opt_stop_below turns that into the smallest length that
keeps the loop going.  Checked against numpy float32 arithmetic around the
bound, for targets across the uint64 range, and at the edge cases."""
import numpy as np
import pytest

import logo2opt as L


def ends(solution, err, length):
    """The reference's exit test, in float32: !(solution * (1.0f + err) < (float)length)."""
    thr = np.float32(solution) * (np.float32(1.0) + np.float32(err))
    return not (thr < np.float32(length))


@pytest.mark.parametrize("solution,err", [
    (1, 0.0), (100, 0.0), (12345, 0.005), (66048945, 0.0), (66048945, 0.005), (238103054, 0.01),
    (5757191, 0.0001), (2**24 + 1, 0.0), (2**40 + 12345, 0.05), (7, 0.5), (0, 0.0), (0, 0.3),
])
def test_stop_below_matches_float_rule(solution, err):
    sb = L.stop_below(solution, err)
    assert sb >= 1
    assert ends(solution, err, sb - 1)
    assert not ends(solution, err, sb)
    # monotone around the bound
    for d in range(1, 40):
        if sb > d:
            assert ends(solution, err, sb - d)
        assert not ends(solution, err, sb + d)


def test_stop_below_random():
    rs = np.random.default_rng(11)
    for _ in range(2000):
        solution = int(rs.integers(1, 1 << int(rs.integers(1, 50))))
        err = float(np.float32(rs.choice([0.0, 0.005, 0.01, 0.1, 1.0, rs.random()])))
        sb = L.stop_below(solution, err)
        assert ends(solution, err, sb - 1) and not ends(solution, err, sb), (solution, err, sb)


def test_stop_below_edges():
    assert L.stop_below(1, -2.0) == 0  # negative threshold: no length ends the loop
    assert L.stop_below(2**64 - 1, 1.0) == 2**64 - 1  # every length ends it
