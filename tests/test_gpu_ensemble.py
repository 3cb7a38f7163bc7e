"""The batched-ensemble workload (cpic_amd.ensemble: independent two-stream
plasmas of uneven grids and densities, FP32 particles, FP64 Poisson solve,
tile shapes and particles per thread from the autotuner's table, steps
replayed from captured graphs, tiles taken largest first).

- 16 plasmas drawn from the workload's generator on 32^2, 64^2 and 128^2 at
  16-128 particles per cell: after 20 steps every plasma's rho, phi, E,
  particle state and field-energy history are within 1e-5 relative L2 of
  the FP64 oracle's run of that plasma alone.
- The full 256-plasma batch (422,682,624 particles) for 20 steps: per
  plasma, particle counts from the device's cell histogram unchanged, total
  charge equal to the particle count (unit charges) to rounding, finite and
  positive field energies; no particle outside its tile."""
import time

import numpy as np
import pytest

import cpic_amd as ca
import pyoracle as po
from common import rel_l2

pytestmark = pytest.mark.gpu

SEED = 2024


def test_ensemble_sample_each_plasma_vs_fp64_oracle():
    from cpic_amd import ensemble
    plist = ensemble.workload(16, seed=99, grids=(32, 64, 128))
    assert {p["grid"] for p in plist} == {32, 64, 128}
    b = ca.Batch(ensemble.descs(plist, flags=ca.F_IDS))
    ensemble.upload(b, plist, SEED, ids=True)
    b.pre_step()
    b.run(0, 20)
    b.sync()
    worst = {}
    for i, p in enumerate(plist):
        g = p["grid"]
        o = po.Oracle(g, g, (float(g), float(g)), p["dt"], [-1.0, -1.0],
                      [float(p["ppc"])] * 2)
        for s, (x, y, ux, uy) in enumerate(ensemble.particles(p, SEED, i)):
            o.add(s, np.arange(len(x), dtype=np.int64), x, y, ux, uy)
        o.pre_step()
        W = []
        for it in range(20):
            o.stage("field_E")
            W.append(o.field_energy())
            o.stage("plasma_E")
            o.stage("plasma_r")
            o.stage("field_rho")
            o.iter = it + 1
        errs = {f: rel_l2(b.field(f, plasma=i), o.field(f)) for f in ("rho", "phi", "Ex", "Ey")}
        errs["W"] = rel_l2(b.energy(plasma=i)["field"][:20], W)
        for s in range(2):
            a, d = o.particles(s), b.particles(s, plasma=i)
            order = np.argsort(a["id"], kind="stable")
            assert np.array_equal(a["id"][order], d["id"])
            for k in ("x", "y", "ux", "uy"):
                errs[f"{s}.{k}"] = rel_l2(d[k], a[k][order])
        for k, v in errs.items():
            worst[k] = max(worst.get(k, 0.0), v)
        assert max(errs.values()) <= 1e-5, (i, p, errs)
    print("worst rel L2 over 16 plasmas vs the FP64 oracle:",
          {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    b.close()


def test_ensemble_full_batch_invariants():
    from cpic_amd import ensemble
    plist = ensemble.workload(256)
    t0 = time.time()
    b = ca.Batch(ensemble.descs(plist))
    ntot = ensemble.upload(b, plist, SEED)
    assert ntot == sum(p["grid"] ** 2 * p["ppc"] // 2 * 2 for p in plist)
    print(f"setup {time.time() - t0:.1f} s, {ntot} particles", flush=True)
    b.pre_step()
    b.run(0, 20)
    b.sync()
    print(f"20 steps {time.time() - t0:.1f} s", flush=True)
    st = b.stats()
    assert st["misplaced"] == 0 and st["nparticles"] == ntot, st
    for i, p in enumerate(plist):
        n = p["grid"] ** 2 * p["ppc"] // 2
        for s in range(2):
            assert int(b.cell_counts(s, plasma=i).sum()) == n, (i, s)
        rho = b.field("rho", plasma=i)
        # -q/e0 = 1 per electron: the deposit sums to the particle count
        assert abs(rho.sum() - 2 * n) <= 1e-9 * 2 * n, (i, rho.sum())
        W = b.energy(maxn=20, plasma=i)["field"]
        assert len(W) == 20 and np.all(np.isfinite(W)) and np.all(W > 0), i
    b.close()
