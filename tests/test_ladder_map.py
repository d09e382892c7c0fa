"""The CNOT ladder as the pass kernels evaluate it (shift form), pinned to the
oracle's ladder (reference qradient/physical_components/state.py:209-251,
restated in oracle/statevector.py:cnot_ladder, one CNOT at a time).

The forward pass gathers psi'(y) = psi(map(y)) with
map(y) = y ^ ((y & m1) >> 1) ^ ((y & m2) >> 2) (qg_pass.hpp ladder_of; the
host derives m1, m2 from the ladder's columns in qg_api.cpp
ladder_shift_form).  Here the masks are read off the oracle's permutation of
the basis states and the formula is checked on every index up to 16 qubits,
and on a random sample of indices up to 30 qubits."""
import numpy as np
import pytest

from oracle import statevector as sv


def _source_map(n, stacking):
    """src[y]: the index whose amplitude the oracle's ladder moves to y."""
    v = np.arange(2 ** n, dtype=np.float64)
    sv.cnot_ladder(v, stacking)
    return v.astype(np.int64)


def _masks(cols):
    """(m1, m2) of a map whose column k is bit k plus at most bits k-1, k-2."""
    m1 = m2 = 0
    for k, col in enumerate(cols):
        assert (col >> k) & 1, (k, col)
        rest = col ^ (1 << k)
        if k >= 1 and (rest >> (k - 1)) & 1:
            m1 |= 1 << k
            rest ^= 1 << (k - 1)
        if k >= 2 and (rest >> (k - 2)) & 1:
            m2 |= 1 << k
            rest ^= 1 << (k - 2)
        assert rest == 0, 'column {} reaches beyond bit k-2: {:#x}'.format(k, col)
    return m1, m2


def _shift_form(y, m1, m2):
    return y ^ ((y & m1) >> 1) ^ ((y & m2) >> 2)


def _source_index(n, y):
    """The oracle ladder's source index of y, tracked through the bonds: with
    stacking 0 the odd bonds act first, so the source of y is reached by the
    even-bond CNOTs, then the odd-bond ones (each an involution on indices;
    control qubit a is bit n-1-a)."""
    even, odd = sv.ladder_bonds(n)
    x = y
    for (a, b) in even + odd:
        x = np.where((x >> (n - 1 - a)) & 1, x ^ (1 << (n - 1 - b)), x)
    return x


@pytest.mark.parametrize('stacking', [0, 1])
@pytest.mark.parametrize('n', list(range(2, 17)))
def test_shift_form_every_index(n, stacking):
    src = _source_map(n, stacking)
    m1, m2 = _masks([int(src[1 << k]) for k in range(n)])
    y = np.arange(2 ** n, dtype=np.int64)
    np.testing.assert_array_equal(_shift_form(y, m1, m2), src)


def test_index_tracking_matches_the_oracle_ladder():
    n = 12
    y = np.arange(2 ** n, dtype=np.int64)
    np.testing.assert_array_equal(_source_index(n, y), _source_map(n, 0))


@pytest.mark.parametrize('n', [20, 24, 26, 29, 30])
def test_shift_form_large_states_sampled(n):
    """At these sizes a full permutation array would not fit a CPU test, so the
    source index is tracked bond by bond (checked against the full oracle
    ladder at 12 qubits above) on basis states and on random indices."""
    m1, m2 = _masks([int(_source_index(n, np.int64(1 << k))) for k in range(n)])
    y = np.random.RandomState(n).randint(0, 2 ** n, size=1 << 14, dtype=np.int64)
    np.testing.assert_array_equal(_shift_form(y, m1, m2), _source_index(n, y))
