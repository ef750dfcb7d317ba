"""The project's random number source for tensor fills and samplers.

The reference uses a global ``std::mt19937_64`` with ``std::normal_distribution``
(PyTAT/PyTAT.hpp:87-126), whose normal variates are implementation-defined.
This package uses numpy's PCG64 instead, seeded through ``TAT.random.seed``,
so seeded inputs are reproducible across machines; parity tests inject the
same storages into the oracle instead of relying on a shared generator.
"""
import numpy as np

_gen = np.random.Generator(np.random.PCG64(0))


def seed(value):
    global _gen
    _gen = np.random.Generator(np.random.PCG64(int(value)))


def generator():
    return _gen


def normal(n, mean=0.0, stddev=1.0, dtype=np.float64):
    return (_gen.standard_normal(n) * stddev + mean).astype(dtype)


def uniform(n, low=0.0, high=1.0, dtype=np.float64):
    return _gen.uniform(low, high, n).astype(dtype)


class _Dist:
    def __init__(self, fn):
        self.fn = fn

    def __call__(self):
        return self.fn()


def uniform_int(low, high):
    """Callable drawing integers in [low, high] (PyTAT.hpp:96-104)."""
    return _Dist(lambda: int(_gen.integers(low, high + 1)))


def uniform_real(low, high):
    return _Dist(lambda: float(_gen.uniform(low, high)))


def normal_real(mean, stddev):
    return _Dist(lambda: float(_gen.normal(mean, stddev)))
