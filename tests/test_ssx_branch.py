"""The screened-exchange branch test without a division.  The kernels decide the regular
branch (wdiffr > limittwo and |delw|^2 < limitone, delw = wt conj(wdiff) / wdiffr) first by
|wt conj(wdiff)|^2 < limitone wdiffr^2 (1 - 2^-40), and form delw with the division only
when that fails.  The implication "fast test true => exact test true" rests on the exact
test's roundings moving |delw|^2 by a few parts in 2^53; here it is checked in numpy (same
operations, IEEE double) on terms placed within 1e-10 of the threshold, where a wrong margin
would show, and on random terms.  No GPU."""
import numpy as np

LIMITTWO = 0.25
LIMITONE = 1.0 / (1e-6 * 4.0)


def _exact(wt, wdiff):
    wdiffr = wdiff.real**2 + wdiff.imag**2
    a = wt * np.conj(wdiff)
    rden = 1.0 / wdiffr
    delw = a.real * rden + 1j * (a.imag * rden)
    delwr = delw.real**2 + delw.imag**2
    return (wdiffr > LIMITTWO) & (delwr < LIMITONE)


def _fast(wt, wdiff):
    wdiffr = wdiff.real**2 + wdiff.imag**2
    a = wt * np.conj(wdiff)
    return (wdiffr > LIMITTWO) & ((a.real**2 + a.imag**2) < LIMITONE * (wdiffr * wdiffr) * (1.0 - 2.0**-40))


def test_fast_regular_implies_exact_regular_near_threshold():
    rng = np.random.default_rng(11)
    n = 400_000
    # |delw|^2 = |wt|^2 / |wdiff|^2: put it within 1e-10 (relative) of limitone
    wdiff = (rng.uniform(0.6, 50, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n)))
    rel = 1.0 + rng.uniform(-1e-10, 1e-10, n)
    wt = np.sqrt(LIMITONE * rel) * np.abs(wdiff) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    fast, exact = _fast(wt, wdiff), _exact(wt, wdiff)
    assert not np.any(fast & ~exact)
    # the margin is narrow: most terms more than 2^-40 below the threshold are decided fast
    below = rel < 1.0 - 2e-12
    assert np.mean(fast[below]) > 0.9
    assert np.any(exact & ~fast)  # the slow path still has work near the threshold


def test_fast_regular_implies_exact_regular_random():
    rng = np.random.default_rng(12)
    n = 400_000
    wdiff = rng.normal(0, 3, n) + 1j * rng.normal(0, 3, n)
    wt = rng.normal(0, 1e3, n) + 1j * rng.normal(0, 1e3, n)
    fast, exact = _fast(wt, wdiff), _exact(wt, wdiff)
    assert not np.any(fast & ~exact)
    assert np.mean(fast == exact) > 0.999


def test_fast_test_is_false_for_nan_and_overflow():
    wd = np.array([np.nan + 0j, 1e200 + 0j, 3.0 + 0j])
    wt = np.array([1.0 + 0j, 1e200 + 0j, np.inf + 0j])
    with np.errstate(all="ignore"):
        assert not np.any(_fast(wt, wd))
