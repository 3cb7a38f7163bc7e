"""Tile-shape and particles-per-thread autotuner.

Two modes, both timing the full step (solve + push + migration + deposit,
captured-graph replay) and keeping the fastest candidate (tile shape, and
particles per thread and chunk of the push kernel):

- `tune`: per grid edge and precision, a batch of identical two-stream
  plasmas (one plasma at 2048^2): the single-run workloads.
- `tune_ensemble`: the batched-ensemble mix itself (cpic_amd.ensemble): the
  tile of each grid edge is tuned inside the mixed batch, one edge at a
  time with the others at their current choice, then the batch's particles
  per thread.

Results are cached as JSON next to this file (tuned/tiles.json), keyed by
"fp<precision>/<grid edge>", with the measured rates and what they were
measured on ("tuned_on"), so a later run (or another process) reuses them.

The reference fixes its blocking by configuration (`simulation.plasma_chunks`,
`simulation.pblock_nmax`, src/sim.c:62-66); here the tile edge is the
analogous knob, and it is chosen by measurement per grid size.

    python -m cpic_amd.autotune --grids 2048 --precision 32
    python -m cpic_amd.autotune --ensemble
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
            "tuned_on": f"{nplasmas} identical {grid}^2 two-stream plasmas at {ppc} ppc",
        }
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return table


def ensemble_sample(every=4):
    """The ensemble mix the tuner measures: every `every`-th plasma of the
    256-plasma workload (the same grid, density, v0 and sigma draws)."""
    from cpic_amd import ensemble
    return ensemble.workload(256)[::every]


def tune_ensemble(plist=None, seed=2024, steps=20, warmup=6, candidates=CANDIDATES,
                  path=CACHE, log=print):
    """Tunes fp32 tiles per grid edge and the batch's particles per thread on
    the ensemble mix (plist, default ensemble_sample()), and writes them to
    the table with the rates and the mix they were measured on."""
    from cpic_amd import ensemble
    plist = ensemble_sample() if plist is None else plist
    table = load(path)
    parts = [ensemble.particles(p, seed, i) for i, p in enumerate(plist)]
    ntot = sum(len(x) for pp in parts for x, _, _, _ in pp)
    edges = sorted({p["grid"] for p in plist},
                   key=lambda g: -sum(p["grid"] ** 2 * p["ppc"] for p in plist if p["grid"] == g))
    current = {g: tuple(tile_for(g, 32, table)) for g in edges}
    ppt = ensemble.batch_ppt(plist, table)

    def rate(tiles, ppt_):
        b = ca.Batch(ensemble.descs(plist, table, flags=ca.F_NO_DIAG, ppt=ppt_, tiles=tiles))
        ensemble.upload(b, plist, seed, parts=parts)
        b.pre_step()
        b.run(0, warmup)
        b.sync()
        t0 = time.perf_counter()
        b.run(warmup, steps)
        b.sync()
        el = time.perf_counter() - t0
        b.close()
        return ntot * steps / el

    rates = {g: {} for g in edges}
    for g in edges:
        for tile in candidates:
            if tile[0] > g or tile[1] > g:
                continue
            r = rate({**current, g: tile}, ppt)
            rates[g][f"{tile[0]}x{tile[1]}/{ppt}"] = r
            log(f"autotune ensemble: edge {g} tile {tile} ppt {ppt}: {r:.4g} particle-steps/s")
        best = max(rates[g], key=rates[g].get)
        current[g] = tuple(int(v) for v in best.split("/")[0].split("x"))
    ppt_rates = {}
    for p in PPT:
        ppt_rates[p] = rate(current, p)
        log(f"autotune ensemble: tiles {current} ppt {p}: {ppt_rates[p]:.4g} particle-steps/s")
    ppt = max(ppt_rates, key=ppt_rates.get)
    every = 256 // len(plist) if len(plist) else 0
    which = "all" if every == 1 else f"every {every}th"
    mix = (f"ensemble mix: {len(plist)} plasmas ({which} of the 256-plasma workload), "
           f"{ntot} particles, grids {sorted(edges)}")
    for g in edges:
        table[f"fp32/{g}"] = {
            "tile": list(current[g]), "ppt": int(ppt), "plasmas": len(plist),
            "rates": rates[g], "ppt_rates": {str(k): v for k, v in ppt_rates.items()},
            "tuned_on": mix,
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
    ap.add_argument("--ensemble", action="store_true",
                    help="tune fp32 tiles and particles per thread on the ensemble mix")
    ap.add_argument("--every", type=int, default=4,
                    help="with --ensemble: tune on every N-th plasma of the 256 (1: all)")
    args = ap.parse_args()
    lg = lambda *a: print(*a, file=sys.stderr, flush=True)   # noqa: E731
    if args.ensemble:
        tune_ensemble(plist=ensemble_sample(args.every), path=args.out, log=lg)
    else:
        tune(args.grids, args.precision, args.ppc, path=args.out, log=lg)
    print(json.dumps(load(args.out)))


if __name__ == "__main__":
    main()
