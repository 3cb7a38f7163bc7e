"""Shared helpers for the parity tests: seeded inputs fed identically to the
HIP library and to the oracle."""
import numpy as np


def two_stream(n_per_species, L, v0=0.1, sigma=0.01, seed=1, nspecies=2):
    """Two counter-streaming electron beams (+-v0 in x, Gaussian spread sigma
    in x and y), uniform positions; returns per-species (ids, x, y, ux, uy)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nspecies):
        n = n_per_species
        x = rng.uniform(0, L[0], n)
        y = rng.uniform(0, L[1], n)
        sign = 1.0 if s % 2 == 0 else -1.0
        ux = sign * v0 + rng.normal(0, sigma, n)
        uy = rng.normal(0, sigma, n)
        out.append((np.arange(n, dtype=np.int64), x, y, ux, uy))
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def feed(sim, parts, add_method):
    for s, (ids, x, y, ux, uy) in enumerate(parts):
        getattr(sim, add_method)(s, ids, x, y, ux, uy)
