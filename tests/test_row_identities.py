"""Floating-point identities the velocity rows rely on (CPU, IEEE double as on the GPU).

A one-point manifold's row recomputes its tangent lever arms instead of reading them: the manifold
reduction stored TauA = 0 + (rA x n) and TauB = 0 + (rB x (-n)) from the contact's own arm and
normal, and the row has only Nsum = 0 + n and rAavg = (0 + rA) * 1.0 (likewise rB).  Those differ
from the contact's values at most in the sign of zero components; the identity below says that
this moves no bit of the recomputed sums, including for arms and normals with signed zeros,
subnormals and exact cancellations.
"""
import numpy as np


def cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]])


def zero_plus(a):
    return np.float64(0.0) + a


def samples(rng, n):
    pool = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -2.0, 5e-324, -5e-324, 1e-308, 3.0, -3.0])
    for _ in range(n):
        kind = rng.integers(3)
        if kind == 0:
            yield rng.normal(size=3) * 10.0 ** rng.integers(-6, 4)
        elif kind == 1:
            yield rng.choice(pool, size=3)
        else:
            v = rng.normal(size=3)
            v[rng.integers(3)] = rng.choice([0.0, -0.0])
            yield v


def test_one_point_lever_arms_recompute_bit_for_bit():
    rng = np.random.default_rng(11)
    arms, normals = list(samples(rng, 4000)), list(samples(rng, 4000))
    with np.errstate(all="ignore"):
        for r, n in zip(arms, normals):
            stored_a = zero_plus(cross(r, n))
            stored_b = zero_plus(cross(r, -n))
            r_row = zero_plus(r) * np.float64(1.0)
            n_row = zero_plus(n)
            again_a = zero_plus(cross(r_row, n_row))
            again_b = zero_plus(cross(r_row, -n_row))
            assert stored_a.tobytes() == again_a.tobytes(), (r, n)
            assert stored_b.tobytes() == again_b.tobytes(), (r, n)


def test_sign_flip_by_multiplication_is_negation():
    # the manifold reduction orients a normal by multiplying with +-1.0 instead of selecting
    # between n and -n
    vals = np.array([0.0, -0.0, 1.5, -2.25, 5e-324, -1e308, np.inf, -np.inf])
    for v in vals:
        assert (np.float64(-1.0) * v).tobytes() == (-v).tobytes()
        assert (np.float64(1.0) * v).tobytes() == np.float64(v).tobytes()
