"""The tuned tile table (cpic_amd/tuned/tiles.json) and its lookups, on the
CPU: every grid the bench workloads use has an entry the library accepts
(power-of-two tile edges no larger than the grid or 64, 2 or 4 particles per
thread), the recorded choice is the fastest measured candidate, and grids
without an entry fall back to the library defaults."""
import os
import tempfile

from cpic_amd import autotune

BENCH_GRIDS = {32: (32,), 64: (32,), 128: (32,), 256: (32,), 2048: (32, 64)}


def pow2(v):
    return v > 0 and v & (v - 1) == 0


def test_table_covers_bench_grids_with_valid_choices():
    table = autotune.load()
    for grid, precs in BENCH_GRIDS.items():
        for prec in precs:
            key = f"fp{prec}/{grid}"
            assert key in table, key
            e = table[key]
            tx, ty = e["tile"]
            assert pow2(tx) and pow2(ty) and tx <= min(grid, 64) and ty <= min(grid, 64), key
            assert e["ppt"] in autotune.PPT, key
            best = max(e["rates"], key=e["rates"].get)
            if "ppt_rates" in e:
                # tuned inside the ensemble mix: tile first, then the batch's ppt
                assert best.split("/")[0] == f"{tx}x{ty}", (key, best)
                assert max(e["ppt_rates"], key=e["ppt_rates"].get) == str(e["ppt"]), key
            else:
                assert best == f"{tx}x{ty}/{e['ppt']}", (key, best)
            assert autotune.tile_for(grid, prec, table) == (tx, ty)
            assert autotune.ppt_for(grid, prec, table) == e["ppt"]


def test_table_is_the_tuners_output_for_the_workloads_that_use_it():
    """The ensemble's grids were tuned inside the ensemble mix (one edge at a
    time, then the batch's particles per thread), the 2048^2 entries on the
    single-plasma workload: each entry records what it was measured on."""
    table = autotune.load()
    for grid in (32, 64, 128, 256):
        e = table[f"fp32/{grid}"]
        assert e["tuned_on"].startswith("ensemble mix"), grid
        assert "ppt_rates" in e
    ppts = {table[f"fp32/{g}"]["ppt"] for g in (32, 64, 128, 256)}
    assert len(ppts) == 1          # one particles-per-thread setting per batch
    for prec in (32, 64):
        e = table[f"fp{prec}/2048"]
        assert e["tuned_on"].startswith("1 identical 2048^2"), prec


def test_untuned_grid_uses_defaults():
    table = {"fp32/64": {"tile": [16, 32], "ppt": 4, "rates": {"16x32/4": 1.0}}}
    assert autotune.tile_for(512, 32, table) == autotune.DEFAULT_TILE
    assert autotune.ppt_for(512, 32, table) == autotune.DEFAULT_PPT
    assert autotune.tile_for(64, 64, table) == autotune.DEFAULT_TILE
    assert autotune.tile_for(64, 32, table) == (16, 32)


def test_missing_cache_file_is_an_empty_table():
    with tempfile.TemporaryDirectory() as d:
        assert autotune.load(os.path.join(d, "none.json")) == {}

