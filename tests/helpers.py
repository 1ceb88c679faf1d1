"""Shared helpers of the test suite (case builders; the oracle is the checker)."""
import numpy as np


def domain_box(n, delta, xa):
    mn = [(xa[d] - 1) * delta for d in range(3)]
    mx = [(xa[d] + n[d] - 1) * delta for d in range(3)]
    dim = [n[d] * delta for d in range(3)]
    return mn, mx, dim


def random_mask(shape, p, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.int8)


def iso_voigt(E, nu):
    C = np.zeros((6, 6))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C[:3, :3] = lam
    C[range(3), range(3)] += 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C
