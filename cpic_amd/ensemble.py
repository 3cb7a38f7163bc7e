"""The batched-ensemble workload: many independent two-stream plasmas of
uneven grids and densities in one handle (cpic_batch_create).

Per plasma: a grid edge drawn from GRIDS, particles per cell drawn uniformly
from 16-128, drift v0 swept linearly over 0.05-0.2 and thermal spread sigma
over 0.005-0.02; unit cells (L = grid edge), two counter-streaming electron
species with m = ppc (w_p = 1), dt = 0.1.  Tile shapes and the push's
particles per thread come from the autotuner's table (cpic_amd/tuned).

Used by bench.py (`--workload ensemble` and the default line's ensemble
object), by the GPU tests of batches and by the autotuner, so that all three
run the same mix.
"""
import numpy as np

import cpic_amd as ca
from cpic_amd import autotune

GRIDS = (32, 64, 128, 256)


def workload(nplasmas=256, seed=256, grids=GRIDS):
    """Per plasma: grid edge, particles per cell, drift v0, spread sigma, dt."""
    rng = np.random.default_rng(seed)
    edges = rng.choice(grids, nplasmas)
    ppcs = rng.integers(16, 129, nplasmas)
    v0 = np.linspace(0.05, 0.2, nplasmas)
    sigma = np.linspace(0.005, 0.02, nplasmas)
    return [dict(grid=int(g), ppc=int(c), v0=float(v), sigma=float(s), dt=0.1)
            for g, c, v, s in zip(edges, ppcs, v0, sigma)]


def particles(p, seed, i):
    """Plasma i's two species [(x, y, ux, uy), ...] (seeded by (seed, i))."""
    rng = np.random.default_rng([seed, i])
    parts, _ = autotune.two_stream(rng, p["grid"], p["ppc"], p["v0"], p["sigma"])
    return parts


def batch_ppt(plist, table):
    """One particles-per-thread setting per batch: the tuned value of the grid
    edge that carries the most particles."""
    load = {}
    for p in plist:
        load[p["grid"]] = load.get(p["grid"], 0) + p["grid"] ** 2 * p["ppc"]
    return autotune.ppt_for(max(load, key=load.get), 32, table)


def descs(plist, table=None, flags=0, ppt=None, tiles=None):
    """cpic_dev_desc_t of every plasma (FP32 particles); tiles maps a grid
    edge to a tile shape (default: the table's)."""
    table = autotune.load() if table is None else table
    ppt = batch_ppt(plist, table) if ppt is None else ppt
    out = []
    for p in plist:
        g = p["grid"]
        tile = tiles[g] if tiles and g in tiles else autotune.tile_for(g, 32, table)
        out.append(ca.make_desc(g, g, (float(g), float(g)), p["dt"], [-1.0, -1.0],
                                [float(p["ppc"]), float(p["ppc"])], precision=32, tile=tile,
                                flags=flags, ppt=ppt))
    return out


def upload(b, plist, seed, ids=False, parts=None, log=None):
    """Uploads every plasma's particles (generated, or from parts[i]);
    returns the particle count.  With ids, particle k of a species gets
    index k."""
    ntot = 0
    for i, p in enumerate(plist):
        pp = parts[i] if parts is not None else particles(p, seed, i)
        for s, (x, y, ux, uy) in enumerate(pp):
            idv = np.arange(len(x), dtype=np.int64) if ids else np.empty(0, np.int64)
            b.upload(s, idv, x, y, ux, uy, plasma=i)
            ntot += len(x)
        if log and i % 32 == 31:
            log(f"ensemble: uploaded {i + 1} plasmas, {ntot} particles")
    return ntot
