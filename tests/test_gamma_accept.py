"""The device Gamma draws decide Marsaglia–Tsang's second acceptance test
(src/statistics.cpp:239) from single-precision estimates with an error bound,
and evaluate the reference's double expression only inside the bound.  The
host build of the same code must agree with the double expression everywhere:
on samples from the draw's own distribution and on samples placed just inside
and just outside the bound.  No device calls."""
import math

import numpy as np
import pytest


def exact(u, t, v, d):
    return math.log(u) < 0.5 * t * t + d * (1.0 - v + math.log(v))


def samples(shape, n, rng):
    """(u, t, v, d) reaching the second test for Gamma(shape), shape >= 1"""
    d = shape - 0.333333333333333
    c = 1.0 / (math.sqrt(d) * 3.0)
    t = rng.standard_normal(n)
    v = 1.0 + c * t
    keep = v > 0
    t, v = t[keep], v[keep] ** 3
    u = rng.random(len(t))
    sq = u < 1.0 - 0.0331 * t ** 4
    return u[~sq], t[~sq], v[~sq], np.full((~sq).sum(), d)


SHAPES = [1.0, 1.25, 2.0, 3.0, 7.0, 40.0, 401.0, 5e3, 1e5, 2e6]


@pytest.mark.parametrize("shape", SHAPES)
def test_natural_samples(vps, shape):
    rng = np.random.default_rng(int(shape * 7) + 1)
    u, t, v, d = samples(shape, 200_000, rng)
    acc, n_exact = vps.debug_mt_accept(u, t, v, d)
    ref = np.array([exact(*x) for x in zip(u, t, v, d)])
    assert (acc == ref).all()
    # the exact form is the rare case
    assert n_exact <= max(2, len(u) // 2000)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("rel", [1e-12, 1e-9, 1e-7, 3e-7, 1e-6, 3e-6, 1e-5, 1e-4])
def test_near_the_boundary(vps, shape, rel):
    """u = exp(rhs) * (1 +- rel): inside the bound the exact form decides,
    just outside it the estimates must decide right"""
    rng = np.random.default_rng(int(shape * 11) + int(-math.log10(rel) * 3))
    u, t, v, d = samples(shape, 20_000, rng)
    rhs = 0.5 * t * t + d * (1.0 - v + np.log(v))
    ok = rhs < -1e-12
    t, v, d, rhs = t[ok], v[ok], d[ok], rhs[ok]
    sign = np.where(rng.random(len(t)) < 0.5, -1.0, 1.0)
    u = np.exp(rhs) * (1.0 + sign * rel)
    acc, n_exact = vps.debug_mt_accept(u, t, v, d)
    ref = np.array([exact(*x) for x in zip(u, t, v, d)])
    assert (acc == ref).all()


def test_edge_inputs(vps):
    u = np.array([0.0, 1e-310, 1e-40, 0.5, 0.5, 0.5, np.nan, 0.999999])
    t = np.array([0.3, 0.3, 0.3, 8.0, -3.0, 0.1, 0.2, 0.0])
    v = np.array([1.1, 1.1, 1.1, 1e-120, 1e120, np.nan, 1.0, 1.0])
    d = np.array([2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, 1e7])
    acc, _ = vps.debug_mt_accept(u, t, v, d)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = [bool(np.log(a) < 0.5 * b * b + e * (1.0 - c + np.log(c))) for a, b, c, e in zip(u, t, v, d)]
    assert acc.tolist() == ref
