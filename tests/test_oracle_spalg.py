"""Pins of the oracle's SpMatrix algebra (oracle/spalg.c: A +- B, c * A,
SpMatrix(A.adjoint()), A * B with Eigen's rules) against numpy and
hand-derived cases, independent of the oracle's own code.

Dyadic values (multiples of 1/8) make every sum and product exact, so the
dense numpy result must equal the oracle's bit for bit whatever the order;
patterns are checked against the union / transpose / path sets computed
from the masks."""
import numpy as np
import pytest

import oracle


def rand_csc(rng, n, density, zero_frac=0.1, dyadic=True):
    """ColMajor arrays of a random pattern; some stored values are exactly 0."""
    mask = rng.random((n, n)) < density
    if dyadic:
        re = rng.integers(-16, 17, (n, n)) / 8.0
        im = rng.integers(-16, 17, (n, n)) / 8.0
    else:
        re, im = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    M = re + 1j * im
    M[rng.random((n, n)) < zero_frac] = 0.0
    M[~mask] = 0.0
    outer, inner, vals = [0], [], []
    for j in range(n):
        for i in range(n):
            if mask[i, j]:
                inner.append(i)
                vals.append(M[i, j])
        outer.append(len(inner))
    return (np.array(outer, np.int32), np.array(inner, np.int32), np.array(vals, np.complex128)), mask, M


def dense(csc):
    return oracle.csc_to_dense(oracle._c(csc))


def pattern(csc, n):
    o, i, _ = csc
    P = np.zeros((n, n), bool)
    for j in range(n):
        P[i[o[j]:o[j + 1]], j] = True
    return P


def sorted_rows(csc):
    o, i, _ = csc
    return all(np.all(np.diff(i[o[j]:o[j + 1]]) > 0) for j in range(len(o) - 1))


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("sign", [1, -1])
def test_add_union_dense(seed, sign):
    rng = np.random.default_rng(seed)
    n = 17
    a, ma, A = rand_csc(rng, n, 0.25)
    b, mb, B = rand_csc(rng, n, 0.25)
    c = oracle.csr_add(a, b, sign)
    assert sorted_rows(c)
    assert np.array_equal(pattern(c, n), ma | mb)  # union, no pruning
    assert np.array_equal(dense(c), A + sign * B)


@pytest.mark.parametrize("c", [0.375, -2.5, 1.25 - 0.5j, 0.5j])
def test_scale_dense(c):
    rng = np.random.default_rng(7)
    a, ma, A = rand_csc(rng, 11, 0.3)
    s = oracle.csr_scale(a, c)
    assert np.array_equal(pattern(s, 11), ma)
    assert np.array_equal(dense(s), c * A)


def test_adjoint_and_product_dense():
    rng = np.random.default_rng(11)
    n = 19
    a, ma, A = rand_csc(rng, n, 0.2)
    b, mb, B = rand_csc(rng, n, 0.2)
    t = oracle.csr_adjoint(a)
    assert sorted_rows(t) and np.array_equal(pattern(t, n), ma.T)
    assert np.array_equal(dense(t), A.conj().T)
    p = oracle.csr_multiply(a, b)
    assert sorted_rows(p)
    # conservative product: every (i, j) reached through some k is stored,
    # even where the sum cancels
    assert np.array_equal(pattern(p, n), (ma.astype(int) @ mb.astype(int)) > 0)
    assert np.array_equal(dense(p), A @ B)
    # a transposed operand is a different product
    assert not np.array_equal(dense(oracle.csr_multiply(t, b)), A @ B)


def test_product_non_dyadic_close():
    rng = np.random.default_rng(5)
    a, _, A = rand_csc(rng, 23, 0.3, dyadic=False)
    b, _, B = rand_csc(rng, 23, 0.3, dyadic=False)
    assert np.allclose(dense(oracle.csr_multiply(a, b)), A @ B, rtol=0, atol=1e-13)


def _one(n, entries):
    """CSC from (row, col, value) entries given column-major."""
    entries = sorted(entries, key=lambda e: (e[1], e[0]))
    outer = np.zeros(n + 1, np.int32)
    for r, c, v in entries:
        outer[c + 1] += 1
    outer = np.cumsum(outer).astype(np.int32)
    return outer, np.array([e[0] for e in entries], np.int32), np.array([e[2] for e in entries], np.complex128)


def bits(v):
    return np.signbit(np.array([v.real, v.imag]))


def test_signed_zeros_and_explicit_zeros():
    """Eigen's union iterator: a lone entry is op(a, 0) / op(0, b)."""
    mz = complex(-0.0, -0.0)
    a = _one(2, [(0, 0, mz), (1, 1, 2.0)])
    b = _one(2, [(1, 0, complex(1.0, -0.0)), (1, 1, -2.0)])
    s = oracle.csr_add(a, b, 1)
    o, i, v = s
    assert list(i) == [0, 1, 1]
    assert not bits(v[0]).any()                 # -0 + 0 = +0
    assert v[1] == 1.0 and not bits(v[1]).any()  # 0 + (1, -0) = (1, +0)
    assert v[2] == 0.0                           # 2 + (-2): explicit zero kept
    d = oracle.csr_add(a, b, -1)
    _, i, v = d
    assert bits(v[0]).all()                      # (-0) - (+0) = -0: the sign survives a difference
    assert v[1] == -1.0 and not bits(v[1])[1]    # 0 - (1, -0) = (-1, +0)
    assert v[2] == 4.0
    # scaling keeps the pattern and explicit zeros; a real scalar scales both parts
    z = oracle.csr_scale(_one(1, [(0, 0, complex(0.0, 3.0))]), -2.0)
    assert z[2][0] == complex(-0.0, -6.0) and bits(z[2][0])[0]  # (-2) * (+0) = -0
    # double * complex scales both parts; a complex scalar with zero imaginary
    # part goes through the full product: im = 2 (-0) + 0 (1) = +0
    a1 = _one(1, [(0, 0, complex(1.0, -0.0))])
    assert bits(oracle.csr_scale(a1, 2.0)[2][0])[1]
    assert not bits(oracle.csr_scale(a1, complex(2.0, 0.0))[2][0])[1]
