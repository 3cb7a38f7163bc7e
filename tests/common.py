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


def seeded_two_stream(nx, ny, ppc, v0, sigma, mode, amp, seed):
    """Two counter-streaming electron beams (+-v0 in x, spread sigma) with
    uniform seeded positions and a velocity perturbation amp*sin(k x),
    k = 2 pi mode / nx, that seeds one Fourier mode above the noise."""
    rng = np.random.default_rng(seed)
    n = nx * ny * ppc // 2
    k = 2 * np.pi * mode / nx
    out = []
    for s in range(2):
        x = rng.uniform(0, nx, n)
        y = rng.uniform(0, ny, n)
        sign = 1.0 if s == 0 else -1.0
        ux = sign * v0 + rng.normal(0, sigma, n) + amp * np.sin(k * x)
        uy = rng.normal(0, sigma, n)
        out.append((np.arange(n, dtype=np.int64), x, y, ux, uy))
    return out


def growth_window(log_energy, lo=3.5, hi=0.5):
    """Iterations [a, b) of the exponential phase of a mode's log-energy
    history: from its last crossing of (max - lo) before the maximum to its
    first crossing of (max - hi)."""
    le = np.asarray(log_energy)
    top = int(np.argmax(le))
    peak = le[top]
    a = top
    while a > 0 and le[a - 1] > peak - lo:
        a -= 1
    b = a
    while b < top and le[b] < peak - hi:
        b += 1
    return a, b


def growth_rate(log_energy, dt, window):
    """Amplitude growth rate: half the least-squares slope of the mode's
    log-energy over the window."""
    a, b = window
    t = np.arange(a, b) * dt
    slope = np.polyfit(t, np.asarray(log_energy)[a:b], 1)[0]
    return 0.5 * slope


def c3_workload(grid, ppc=64):
    """Electrons + mobile ions at 1:1836 on a grid x grid periodic domain with
    unit cells: half the particles are electrons (q = -1, m = n_e per cell so
    that w_p = 1) counter-streaming at +-0.1 with a 0.01 Gaussian spread, half
    are ions (q = +1, m = 1836 m_e, spread 0.01/sqrt(1836)); B = 0, dt = 0.05."""
    n_cells = grid * grid
    n_e = n_cells * ppc // 2
    m_e = n_e / n_cells
    return dict(grid=grid, L=float(grid), n=(n_e, n_cells * ppc - n_e), q=(-1.0, 1.0),
                m=(m_e, 1836.0 * m_e), v0=(0.1, 0.0), sigma=(0.01, 0.01 / np.sqrt(1836.0)),
                dt=0.05)


def c3_species(w, s, seed, ids=True):
    """Species s of a c3_workload: (ids, x, y, ux, uy), positions uniform,
    streams alternating +-v0 by index."""
    rng = np.random.default_rng([seed, s])
    n, L = w["n"][s], w["L"]
    x = rng.uniform(0.0, L, n)
    y = rng.uniform(0.0, L, n)
    ux = rng.normal(0.0, w["sigma"][s], n)
    if w["v0"][s]:
        ux[0::2] += w["v0"][s]
        ux[1::2] -= w["v0"][s]
    uy = rng.normal(0.0, w["sigma"][s], n)
    return (np.arange(n, dtype=np.int64) if ids else np.empty(0, np.int64)), x, y, ux, uy
