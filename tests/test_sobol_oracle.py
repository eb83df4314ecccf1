"""The Sobol / SpaceFillerParticleLayout oracle pinned by the reference's SobolGeneratorSuite
(StgDomain/Utils/tests/SobolGeneratorSuite.c:62-120): for each of the first six generators of the table, the
numbers at indices 3, 776, 1549, ... (index % 773 == 3) of the first 200000, printed "%.4g", must equal
testSobolGeneratorOutput.<degree>-<coefficient>.expected line by line; the Gray-code closed form equals the
recurrence; the rightmost-zero-bit table equals testSobolGeneratorRightmostBitOutput.expected.  The layout is
checked against its definition: counts, in-box coordinates, the lower-element rule on element faces."""
import os

import numpy as np
import pytest

from oracle import sobol_oracle as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sobol")


@pytest.mark.parametrize("entry", range(6))
def test_sequence_matches_reference_suite(entry):
    s, a, _ = S.TABLE[entry]
    with open(os.path.join(GOLD, f"testSobolGeneratorOutput.{s:03d}-{a:03d}.expected")) as f:
        lines = f.read().split("\n")
    assert lines[0].strip().startswith("*" * 5)
    expected = [ln for ln in lines[1:] if ln.strip()]
    x = S.sobol_numbers(entry, 200000)
    got = ["%.4g" % v for i, v in enumerate(x) if i % 773 == 3]
    assert got == expected
    idx = np.arange(1, 200001)
    assert np.array_equal(S.sobol_by_index(entry, idx), S.sobol_ints(entry, 200000))


def test_rightmost_zero_bit_table():
    with open(os.path.join(GOLD, "testSobolGeneratorRightmostBitOutput.expected")) as f:
        rows = [ln for ln in f.read().split("\n")[1:] if ln.strip()]
    assert len(rows) == 30
    for n, row in enumerate(rows):
        c = 0
        while (n >> c) & 1:
            c += 1
        assert row == f"{n:016b} number {n}: {c}"


def test_first_numbers_by_hand():
    # generator 0 (x + 1): m_i = 2 m_{i-1} ^ m_{i-1} = 1, 3, 5, 15, 17, 51; x_n = x_{n-1} ^ m_c / 2^(c+1)
    # with c the rightmost zero bit of n - 1: .1, .1^.11 = .01, .01^.1 = .11, .11^.101 = .011, .111, .001, .101
    assert S.direction_numbers(0)[:6] == [1, 3, 5, 15, 17, 51]
    assert list(S.sobol_numbers(0, 7)) == [0.5, 0.25, 0.75, 0.375, 0.875, 0.125, 0.625]
    # dimension 1 (x^2 + x + 1, m = 1, 1, ...): m_2 = 2 m_1 ^ 4 m_0 ^ m_0 = 7
    assert S.direction_numbers(1)[:4] == [1, 1, 7, 11]


def test_spacefiller_layout_rules():
    nel, lo, hi = (10, 10, 1), (0.0, 0.0, 0.0), (0.9142, 1.0, 1.0)
    off, xi, X = S.spacefiller(2, nel, lo, hi, 20)
    assert off[-1] == 2000 and off[0] == 0 and np.all(np.diff(off) >= 0)
    assert np.all((X >= 0.0) & (X <= np.array(hi[:2])))
    assert np.all(np.abs(xi[:, :2]) <= 1.0 + 1e-14)
    # particle 1 of the sequence sits at (0.5, 0.5) of the box: the vertex shared by elements (4,4), (5,4),
    # (4,5), (5,5), and belongs to the lower one, (4, 4)
    e = np.repeat(np.arange(100), np.diff(off))
    first = np.flatnonzero((X[:, 0] == 0.5 * 0.9142) & (X[:, 1] == 0.5))
    assert len(first) == 1 and e[first[0]] == 44
    assert np.allclose(xi[first[0], :2], [1.0, 1.0], atol=1e-14)
    # the rule along one axis
    assert list(S.element_index(np.array([0.0, 0.25, 0.5, 0.5 + 1e-9, 1.0]), 0.0, 1.0, 4)) == [0, 0, 1, 2, 3]
