"""Tile-shape and particles-per-thread autotuner for batched runs.

For each grid edge and precision, times the full step (solve + push +
migration + deposit, captured-graph replay) of a batch of identical
two-stream plasmas under every candidate (tile shape, particles per thread
and chunk of the push kernel) and keeps the fastest.
Results are cached as JSON next to this file (tuned/tiles.json), keyed by
"fp<precision>/<grid edge>", with the measured rates, so a later run (or
another process) reuses them without timing again.

The reference fixes its blocking by configuration (`simulation.plasma_chunks`,
`simulation.pblock_nmax`, src/sim.c:62-66); here the tile edge is the
analogous knob, and it is chosen by measurement per grid size.

    python -m cpic_amd.autotune --grids 32 64 128 256 --precision 32
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import cpic_amd as ca

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned", "tiles.json")
CANDIDATES = ((8, 8), (16, 8), (16, 16), (32, 16), (16, 32), (32, 32))
PPT = (2, 4)
DEFAULT_TILE = (16, 16)
DEFAULT_PPT = 2


def two_stream(rng, grid, ppc, v0, sigma):
    """Two electron species of grid^2*ppc/2 particles each (+-v0 in x,
    Gaussian spread sigma), uniform positions, unit cells; m = ppc gives
    w_p = 1 with q = -1, e0 = 1."""
    n = grid * grid * ppc // 2
    out = []
    for s in range(2):
        x = rng.uniform(0.0, grid, n)
        y = rng.uniform(0.0, grid, n)
        ux = rng.normal(0.0, sigma, n) + (v0 if s == 0 else -v0)
        uy = rng.normal(0.0, sigma, n)
        out.append((x, y, ux, uy))
    return out, float(ppc)


def load(path=CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tile_for(grid, precision, table=None):
    """Tuned tile for a grid edge (the default when not tuned)."""
    table = load() if table is None else table
    e = table.get(f"fp{precision}/{grid}")
    return tuple(e["tile"]) if e else DEFAULT_TILE


def ppt_for(grid, precision, table=None):
    """Tuned particles per thread for a grid edge (the default when not tuned)."""
    table = load() if table is None else table
    e = table.get(f"fp{precision}/{grid}")
    return int(e.get("ppt", DEFAULT_PPT)) if e else DEFAULT_PPT


def time_candidate(grid, tile, precision, parts, m, nplasmas, steps, warmup, dt=0.1, ppt=0):
    descs = [ca.make_desc(grid, grid, (float(grid), float(grid)), dt, [-1.0, -1.0], [m, m],
                          precision=precision, tile=tile, flags=ca.F_NO_DIAG, ppt=ppt)
             for _ in range(nplasmas)]
    b = ca.Batch(descs)
    empty = np.empty(0, np.int64)
    for p in range(nplasmas):
        for s, (x, y, ux, uy) in enumerate(parts):
            b.upload(s, empty, x, y, ux, uy, plasma=p)
    b.pre_step()
    b.run(0, warmup)
    b.sync()
    t0 = time.perf_counter()
    b.run(warmup, steps)
    b.sync()
    el = time.perf_counter() - t0
    b.close()
    n = nplasmas * sum(len(x) for x, _, _, _ in parts)
    return n * steps / el


def tune(grids, precision=32, ppc=64, target_particles=1 << 25, steps=20, warmup=6,
         candidates=CANDIDATES, seed=5, path=CACHE, log=print):
    table = load(path)
    for grid in grids:
        rng = np.random.default_rng([seed, grid])
        parts, m = two_stream(rng, grid, ppc, 0.1, 0.01)
        nplasmas = max(1, target_particles // (grid * grid * ppc))
        rates = {}
        for tile in candidates:
            if tile[0] > grid or tile[1] > grid:
                continue
            for ppt in PPT:
                r = time_candidate(grid, tile, precision, parts, m, nplasmas, steps, warmup,
                                   ppt=ppt)
                rates[f"{tile[0]}x{tile[1]}/{ppt}"] = r
                log(f"autotune fp{precision} grid {grid} x{nplasmas} tile {tile} ppt {ppt}: "
                    f"{r:.4g} particle-steps/s")
        best = max(rates, key=rates.get)
        shape, ppt = best.split("/")
        table[f"fp{precision}/{grid}"] = {
            "tile": [int(v) for v in shape.split("x")], "ppt": int(ppt), "ppc": ppc,
            "plasmas": nplasmas, "rates": rates,
        }
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return table


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--grids", type=int, nargs="+", default=[32, 64, 128, 256])
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--out", default=CACHE)
    args = ap.parse_args()
    tune(args.grids, args.precision, args.ppc, path=args.out,
         log=lambda *a: print(*a, file=sys.stderr, flush=True))
    print(json.dumps(load(args.out)))


if __name__ == "__main__":
    main()
