"""Pin the collapsed Gauss-Jacobi rules: exact integration of every monomial of
total degree <= m on the reference simplex (closed form a! b! (c!) / (a+b(+c)+d)!),
and agreement of the product's Golub-Welsch construction with the oracle's scipy
construction."""

from math import factorial

import numpy as np
import pytest

from oracle.quadrature import gauss_jacobi
from nfempy_amd.quadrature import make_quadrature


def exact_monomial(powers):
    d = len(powers)
    num = 1
    for p in powers:
        num *= factorial(p)
    return num / factorial(sum(powers) + d)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("m", [3, 5])
def test_exact_on_monomials(dim, m):
    pts, wts = gauss_jacobi(dim, m)
    assert len(wts) == ((m + 2) // 2) ** dim
    import itertools

    for powers in itertools.product(range(m + 1), repeat=dim):
        if sum(powers) > m:
            continue
        val = np.sum(wts * np.prod(pts ** np.array(powers), axis=1))
        assert abs(val - exact_monomial(powers)) < 1e-14, (powers, val)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("m", [1, 3, 5, 7])
def test_product_matches_oracle(dim, m):
    a = make_quadrature(dim, m)
    b = gauss_jacobi(dim, m)
    assert np.allclose(a[0], b[0], atol=1e-14, rtol=0)
    assert np.allclose(a[1], b[1], atol=1e-15, rtol=1e-13)


def test_not_exact_above_degree():
    # the m=3 triangle rule (4 points) cannot integrate x^4 exactly: guards a wrong np
    pts, wts = gauss_jacobi(2, 3)
    val = np.sum(wts * pts[:, 0] ** 4)
    assert abs(val - exact_monomial((4, 0))) > 1e-6
