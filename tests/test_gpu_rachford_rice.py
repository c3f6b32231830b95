"""GPU Rachford-Rice: the bisection with certified sign skipping returns bit for bit
what the plain bisection returns on the same device (status, beta, x, y, iteration
count), including roots near 0 and 1, near-degenerate K and poor warm-start guesses;
and it agrees with the oracle to the last few ulps (the device exp/expm1 used for the
K-dependent terms may differ from glibc in the last bit)."""
import numpy as np
import pytest

import oracle_bindings as ob


def _cases(nc, n, seed):
    rng = np.random.default_rng(seed)
    z = rng.gamma(1.0, 1.0, size=(nc, n)) + 1e-6
    z /= z.sum(axis=0, keepdims=True)
    kind = rng.integers(0, 4, n)
    lk = rng.normal(0.0, 1.5, size=(nc, n))
    lk[:, kind == 1] *= 1e-4                       # near-degenerate K
    lk[:, kind == 2] *= 8.0                        # wide spread: roots near 0 or 1
    shift = rng.normal(0.0, 2.0, size=n)
    lk[:, kind == 3] += shift[kind == 3]           # one-sided K: no interior root often
    straddle = kind != 3                           # otherwise K_0 > 1 > K_1
    lk[0, straddle] = np.abs(lk[0, straddle]) + 0.05
    lk[1, straddle] = -np.abs(lk[1, straddle]) - 0.05
    guess = rng.uniform(0.0, 1.0, n)
    guess[rng.random(n) < 0.2] = 0.5
    return z, lk, guess


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [2, 3, 6, 12])
def test_rr_skipping_is_bitwise_plain_bisection(gpu, nc):
    from mpmc_hnu_amd import flash

    z, lk, guess = _cases(nc, 50000, seed=nc)
    fast = flash.to_host(flash.rachford_rice_batch(z, lk, guess))
    plain = flash.to_host(flash.rachford_rice_batch(z, lk, np.full(guess.shape, -1.0)))
    for k in ("status", "iterations"):
        assert np.array_equal(fast[k], plain[k]), k
    interior = plain["status"] == 0
    print("interior", interior.sum(), "of", z.shape[1], flush=True)
    assert interior.sum() > 0.25 * z.shape[1]
    assert np.array_equal(fast["beta"][interior], plain["beta"][interior])
    assert np.array_equal(fast["x"][:, interior], plain["x"][:, interior])
    assert np.array_equal(fast["y"][:, interior], plain["y"][:, interior])


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [2, 6, 12])
def test_rr_against_oracle(gpu, nc):
    from mpmc_hnu_amd import flash

    z, lk, guess = _cases(nc, 4000, seed=100 + nc)
    g = flash.to_host(flash.rachford_rice_batch(z, lk, guess))
    n = z.shape[1]
    st = np.zeros(n, np.int32)
    beta = np.zeros(n)
    for c in range(n):
        s, b, x, y, it = ob.rachford_rice(z[:, c], lk[:, c])
        st[c], beta[c] = s, b
    same = g["status"] == st
    assert same.all(), np.nonzero(~same)[0][:8].tolist()
    both = same & (st == 0)
    rel = np.abs(g["beta"][both] - beta[both]) / beta[both]
    assert rel.max() < 1e-12, rel.max()
