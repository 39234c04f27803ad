"""The oracle's stencil weights against the reference's textbook tables
(test/stencils_tests.jl:40-55) and the properties it asserts in exact
arithmetic (:121-200)."""
from fractions import Fraction as F

import pytest

import oracle_lib as O

FD1 = {2: [F(-1, 2), 0, F(1, 2)],
       4: [F(1, 12), F(-2, 3), 0, F(2, 3), F(-1, 12)],
       6: [F(-1, 60), F(3, 20), F(-3, 4), 0, F(3, 4), F(-3, 20), F(1, 60)],
       8: [F(1, 280), F(-4, 105), F(1, 5), F(-4, 5), 0, F(4, 5), F(-1, 5), F(4, 105), F(-1, 280)]}
FD2 = {2: [1, -2, 1],
       4: [F(-1, 12), F(4, 3), F(-5, 2), F(4, 3), F(-1, 12)],
       6: [F(1, 90), F(-3, 20), F(3, 2), F(-49, 18), F(3, 2), F(-3, 20), F(1, 90)],
       8: [F(-1, 560), F(8, 315), F(-1, 5), F(8, 5), F(-205, 72), F(8, 5), F(-1, 5), F(8, 315),
           F(-1, 560)]}
KO = {2: [F(c, 16) for c in (-1, 4, -6, 4, -1)],
      3: [F(c, 64) for c in (1, -6, 15, -20, 15, -6, 1)],
      4: [F(c, 256) for c in (-1, 8, -28, 56, -70, 56, -28, 8, -1)],
      5: [F(c, 1024) for c in (1, -10, 45, -120, 210, -252, 210, -120, 45, -10, 1)]}


def rational(num, den):
    return [F(n, d) for n, d in zip(num, den)]


@pytest.mark.parametrize("q", [2, 4, 6, 8])
def test_derivative_weights_are_the_textbook_tables(q):
    n1, d1, w1 = O.weights(q, 1)
    n2, d2, w2 = O.weights(q, 2)
    assert rational(n1, d1) == [F(x) for x in FD1[q]]
    assert rational(n2, d2) == [F(x) for x in FD2[q]]
    # rounded once: the double is the correctly rounded exact rational
    for (n, d, w) in ((n1, d1, w1), (n2, d2, w2)):
        for a, b, x in zip(n, d, w):
            assert x == a / b


@pytest.mark.parametrize("q", [2, 4, 6, 8])
def test_each_derivative_is_exact_to_its_degree_and_not_further(q):
    x0, h = F(3, 7), F(2, 5)
    r = q // 2
    for m in (1, 2):
        num, den, _ = O.weights(q, m)
        w = rational(num, den)
        exact_to = q if m == 1 else q + 1
        for deg in range(exact_to + 2):
            approx = sum(wk * (x0 + (k - r) * h) ** deg for k, wk in enumerate(w)) / h ** m
            if deg < m:
                exact = F(0)
            else:
                c = 1
                for i in range(m):
                    c *= deg - i
                exact = c * x0 ** (deg - m)
            if deg <= exact_to:
                assert approx == exact, (q, m, deg)
            else:
                assert approx != exact, (q, m, deg)


@pytest.mark.parametrize("r", [2, 3, 4, 5])
def test_dissipation_is_the_signed_binomial_row(r):
    num, den, w = O.diss_weights(r)
    wr = rational(num, den)
    assert wr == KO[r]
    x0, h = F(3, 7), F(2, 5)
    for deg in range(2 * r):
        assert sum(wk * (x0 + (k - r) * h) ** deg for k, wk in enumerate(wr)) == 0
    assert sum(wk * (x0 + (k - r) * h) ** (2 * r) for k, wk in enumerate(wr)) != 0
    # Nyquist is an eigenvector with eigenvalue exactly -1; the center is negative
    assert sum(wk * (-1) ** (k - r) for k, wk in enumerate(wr)) == -1
    assert wr[r] < 0 and wr == wr[::-1]
