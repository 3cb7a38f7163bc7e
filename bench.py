#!/usr/bin/env python3
"""Throughput of the full 2D electrostatic PIC step on one MI355X.

Workload (mixed-precision two-stream at scale): periodic 2048 x 2048 grid,
unit cells, 268,435,456 particles (64 per cell) in two species -- electrons
(q=-1, m=32, i.e. w_p = 1) counter-streaming at +-0.1 with a 0.01 Gaussian
spread, and mobile ions (q=+1, m=1836*32, spread 0.01/sqrt(1836)) -- B = 0,
dt = 0.05.  One step = FFT Poisson solve (FP64) -> E -> CIC gather -> Boris
push -> periodic wrap + tile migration -> CIC deposit.  Particles are FP32
(tile-relative positions); the same workload in FP64 is timed for
comparison.  Inputs are generated on the host (synthetic, seeded), uploaded,
and resident in HBM before the timed region.

Prints one JSON line.  `value` = particle-steps/s of the FP32 step, timed
over K steps of the library's production path (`cpic_dev_run`: steps
replayed two at a time from a captured HIP graph); the per-kernel times
(roofline) come from K further steps launched one by one between HIP events.
The same line carries the FP64 step, the FP32 accuracy against the FP64 run
(1000 steps, full size) and against the FP64 oracle (the same workload scaled
to 256^2, 1000 steps; the oracle runs in the CPU-baseline leg), the step time
late in the run (steps 900-1000), the HBM streaming ceiling of a plain
read+write kernel, and the batched-ensemble and differentiable-run figures
below as `ensemble` and `adjoint` objects.

`--workload adjoint` times the differentiable FP64 run (256^2, 4,194,304
particles, 200 steps, checkpoints every 20: forward and reverse sweep).

`--workload ensemble` instead times a batch of 256 independent two-stream
plasmas in one handle (grids drawn from 32^2..256^2, 16-128 particles per
cell, v0 swept 0.05-0.2 and sigma 0.005-0.02, FP32 particles + FP64 FFT
Poisson, tile shapes from the autotuner), 300 steps replayed from captured
HIP graphs, and prints its own JSON line.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "particle-steps/s per full PIC step at 2048^2, 64 ppc; FP32 vs FP64 GPU; %HBM"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_species(rng, n, L, v0, sigma):
    x = rng.uniform(0.0, L, n)
    y = rng.uniform(0.0, L, n)
    ux = rng.normal(0.0, sigma, n)
    if v0:
        ux[0::2] += v0
        ux[1::2] -= v0
    uy = rng.normal(0.0, sigma, n)
    return x, y, ux, uy


def workload(grid, ppc):
    n_cells = grid * grid
    n_e = n_cells * ppc // 2
    n_i = n_cells * ppc - n_e
    m_e = n_e / n_cells          # w_p = 1 with q = -1, e0 = 1, unit cells
    return dict(grid=grid, L=float(grid), n=(n_e, n_i), q=(-1.0, 1.0),
                m=(m_e, 1836.0 * m_e), v0=(0.1, 0.0),
                sigma=(0.01, 0.01 / np.sqrt(1836.0)), dt=0.05)


def make_device(ca, w, precision, flags, seed=2024, tile=(0, 0), ppt=0):
    d = ca.Device(w["grid"], w["grid"], (w["L"], w["L"]), w["dt"], list(w["q"]), list(w["m"]),
                  nchunks=1, precision=precision, flags=flags, tile=tile, ppt=ppt)
    for s in range(2):
        rng = np.random.default_rng([seed, s])
        x, y, ux, uy = gen_species(rng, w["n"][s], w["L"], w["v0"][s], w["sigma"][s])
        d.upload(s, np.empty(0, np.int64), x, y, ux, uy)
        del x, y, ux, uy
    return d


def time_steps(d, it0, warmup, steps):
    d.pre_step()
    d.run(it0, warmup)
    d.sync()
    d.reset_kernel_times()
    t0 = time.perf_counter()
    d.run(it0 + warmup, steps)
    d.sync()
    t1 = time.perf_counter()
    return (t1 - t0), d.kernel_times()


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def oracle_sample(po, w, frac, steps, seed=7):
    """Times the oracle's steps on the full grid with 1/frac of the
    particles: (field seconds per step, particle seconds per step, sample
    size)."""
    grid = w["grid"]
    o = po.Oracle(grid, grid, (w["L"], w["L"]), w["dt"], list(w["q"]), list(w["m"]), nchunks=64)
    nsample = 0
    for s in range(2):
        n = w["n"][s] // frac
        rng = np.random.default_rng([seed, s])
        x, y, ux, uy = gen_species(rng, n, w["L"], w["v0"][s], w["sigma"][s])
        o.add(s, np.arange(n, dtype=np.int64), x, y, ux, uy)
        nsample += n
        del x, y, ux, uy
    o.pre_step()
    o.step()                                  # iteration 0
    tf = tp = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        o.stage("field_E")
        t1 = time.perf_counter()
        o.stage("plasma_E")
        o.stage("plasma_r")
        o.stage("field_rho")
        t2 = time.perf_counter()
        o.iter = o.iter + 1
        tf += t1 - t0
        tp += t2 - t1
    del o
    return tf / steps, tp / steps, nsample


def oracle_history(po, w, seed, steps):
    """The oracle's field-energy history over `steps` steps of workload w from
    the seeded input make_device uploads (same generator, same seed)."""
    grid = w["grid"]
    o = po.Oracle(grid, grid, (w["L"], w["L"]), w["dt"], list(w["q"]), list(w["m"]), nchunks=min(64, grid))
    for s in range(2):
        rng = np.random.default_rng([seed, s])
        x, y, ux, uy = gen_species(rng, w["n"][s], w["L"], w["v0"][s], w["sigma"][s])
        o.add(s, np.arange(len(x), dtype=np.int64), x, y, ux, uy)
        del x, y, ux, uy
    o.pre_step()
    hist = []
    for it in range(steps):
        o.stage("field_E")
        hist.append(o.field_energy())
        o.stage("plasma_E")
        o.stage("plasma_r")
        o.stage("field_rho")
        o.iter = it + 1
    del o
    return np.array(hist)


def cpu_baseline(w, frac=8, steps=10, accuracy=None):
    """The FP64 oracle (C restatement of the reference loop, OpenMP over 64
    chunks) on the host cores: the full grid with 1/frac of the particles;
    the full workload's step time is extrapolated as t_fields +
    frac * t_particles.  Also one core on a 1/256 sample.  With `accuracy`
    (from fp32_oracle_leg_gpu), also the oracle's run of that scaled
    workload, whose field-energy history is the FP32 accuracy reference
    (stored in accuracy["oracle"]) and whose rate is reported beside."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    acc_run = None
    if accuracy is not None:
        wa = accuracy["w"]
        t0 = time.perf_counter()
        accuracy["oracle"] = oracle_history(po, wa, accuracy["seed"], accuracy["steps"])
        el = time.perf_counter() - t0
        acc_run = {"workload": f"the same workload at {wa['grid']}^2, {sum(wa['n'])} particles, "
                               f"{accuracy['steps']} steps (set-up included)",
                   "seconds": el,
                   "particle_steps_per_s": sum(wa["n"]) * accuracy["steps"] / el,
                   "cores": threads}
        log(f"oracle accuracy run: {el:.1f} s")
    tf, tp, ns = oracle_sample(po, w, frac, steps)
    t_full = tf + frac * tp
    ntot = sum(w["n"])
    single = None
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        tf1, tp1, ns1 = oracle_sample(po, w, 256, 2)
        gomp.omp_set_num_threads(threads)
        single = {"value": ntot / (tf1 + 256 * tp1), "cores": 1,
                  "ns_per_particle_step": tp1 / ns1 * 1e9, "fields_ms": tf1 * 1e3}
    except OSError:
        pass
    out = {
        "value": ntot / t_full,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP64 on {w['grid']}^2 with {ns} particles (1/{frac} of the workload), "
                   f"{steps} steps: fields {tf * 1e3:.1f} ms/step, particles {tp * 1e3:.1f} "
                   f"ms/step ({tp / ns * 1e9:.2f} ns/particle-step); full step extrapolated "
                   f"to {t_full:.2f} s; CPU: {cpu_model()}, {threads} threads (nproc "
                   f"{os.cpu_count()}), radix-2 FFT"),
    }
    if single:
        out["single_core"] = single
    if acc_run:
        out["accuracy_reference_run"] = acc_run
    # the reference's own best single-node figure for this grid, as published
    # (perf/strong-scaling-fixed/csv/time.csv:2): not re-measured here
    out["published_reference"] = {
        "value": 1e8 / 3.184, "unit": "particle-steps/s", "kind": "published, not re-measured",
        "config": "2048^2, 1e8 particles, FP64, full step 3.184 s/iter",
        "hardware": "MareNostrum 4, 1 node, 48 cores",
        "source": "perf/strong-scaling-fixed/csv/time.csv:2"}
    return out


def fp32_oracle_leg_gpu(ca, grid, steps, seed):
    """The FP32 run of the workload scaled to grid^2 (same particles per
    cell, species and constants; the full run's tile shape and particles per
    thread): its field-energy history over `steps` steps.  The oracle's
    history of the same input is computed in the CPU-baseline leg."""
    from cpic_amd import autotune
    w = workload(grid, 64)
    tile, ppt = autotune.tile_for(2048, 32), autotune.ppt_for(2048, 32)
    d = make_device(ca, w, 32, 0, seed=seed, tile=tile, ppt=ppt)
    d.pre_step()
    d.run(0, steps)
    d.sync()
    hist = d.energy(maxn=steps)["field"][:steps]
    d.close()
    return dict(w=w, seed=seed, steps=steps, gpu=np.array(hist), tile=list(tile), ppt=ppt)


def acc_report(acc):
    """FP32 field-energy error against the oracle, per step."""
    g, o = acc["gpu"], acc.get("oracle")
    if o is None:
        return None
    n = min(len(g), len(o))
    rel = np.abs(g[:n] - o[:n]) / np.abs(o[:n])
    above = np.nonzero(rel > 1e-5)[0]
    w = acc["w"]
    return {"steps": int(n), "reference": "FP64 oracle (CPU) on the same seeded input",
            "workload": f"two-stream {w['grid']}^2, 64 ppc, {sum(w['n'])} particles, "
                        f"electrons + ions (1:1836), FP32 particles + FP64 FFT Poisson",
            "tile": acc["tile"], "ppt": acc["ppt"],
            "max_rel": float(rel.max()), "max_rel_first_100": float(rel[:100].max()),
            "first_step_above_1e-5": int(above[0]) if len(above) else None,
            "energy_growth": float(o[n - 1] / o[0])}


def stream_ceiling(total_bytes, achieved, tile_slots=None):
    """Best rate of a hand-written read+write stream (lib/libcpic_probe.so:
    16 B per element, dwordx4 loads and stores, 1-8 loads in flight per
    thread) moving the push's algorithmic bytes, half read and half written."""
    import ctypes
    try:
        probe = ctypes.CDLL(os.path.join(ROOT, "cpic_amd", "lib", "libcpic_probe.so"))
    except OSError as e:
        return {"ceiling_GBs": None, "ceiling_error": str(e)}
    probe.cpic_probe_stream.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]
    half = total_bytes // 2 // 16 * 16
    rates = {}
    for u in (1, 2, 4, 8):
        ms = ctypes.c_double()
        if probe.cpic_probe_stream(half, u, 10, ctypes.byref(ms)) == 0 and ms.value > 0:
            rates[str(u)] = 2 * half / (ms.value / 1e3) / 1e9
    if not rates:
        return {"ceiling_GBs": None}
    best = max(rates.values())
    out = {"ceiling_GBs": best,
           "ceiling_kernel": (f"read {half} B + write {half} B, 16 B per element (dwordx4), "
                              f"best of 1/2/4/8 loads in flight per thread"),
           "ceiling_GBs_by_loads_in_flight": rates,
           "frac_of_ceiling": achieved / best if achieved else None}
    # the push's own load/store pattern without its arithmetic
    probe.cpic_probe_stream_groups.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    if probe.cpic_probe_stream_groups(half, 0, 10, ctypes.byref(ms)) == 0 and ms.value > 0:
        moved = 2 * (half // 16 // 1024 * 1024 * 16)
        pat = moved / (ms.value / 1e3) / 1e9
        out.update(pattern_ceiling_GBs=pat,
                   pattern_kernel=("the push's grouped FP32 state: dwordx4 loads of 4 slots per "
                                   "thread and component, one dword store per slot and component, "
                                   "no arithmetic"),
                   frac_of_pattern_ceiling=achieved / pat if achieved else None)
    # the same pattern streamed bucket by bucket, one workgroup per bucket as
    # the push runs (tile_slots particles per bucket, chunks of 1024)
    if tile_slots:
        probe.cpic_probe_stream_tiles.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_double)]
        ts = max(1024, int(round(tile_slots / 1024)) * 1024)
        ms = ctypes.c_double()
        if probe.cpic_probe_stream_tiles(half, ts, 10, ctypes.byref(ms)) == 0 and ms.value > 0:
            moved = 2 * (half // 16 // ts * ts * 16)
            tpat = moved / (ms.value / 1e3) / 1e9
            out.update(tile_pattern_ceiling_GBs=tpat,
                       tile_pattern_kernel=(f"the same pattern, one 256-thread workgroup per "
                                            f"bucket of {ts} slots streamed in chunks of 1024, in "
                                            f"XCD bands as the push orders them"),
                       frac_of_tile_pattern_ceiling=achieved / tpat if achieved else None)
    log(f"stream ceiling: {rates}, push pattern: {out.get('pattern_ceiling_GBs')}, "
        f"per tile: {out.get('tile_pattern_ceiling_GBs')}")
    return out


def traffic_from_profiles(precision, name=None):
    """HBM bytes per push launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), if present."""
    path = os.path.join(ROOT, "profiles", name or f"pmc_push_fp{precision}_2048.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            try:
                return float(line.split("=", 1)[1].split()[0])
            except (IndexError, ValueError):
                return None
    return None


def copy_calibration(nbytes=1 << 30, reps=3):
    """Plain device-to-device copies of a known size (the runtime's copy
    kernel), used to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {reps} x {nbytes} bytes device to device")


def run_ensemble(args, ca):
    """The batched ensemble (cpic_amd.ensemble): returns its JSON object."""
    from cpic_amd import autotune, ensemble
    plist = ensemble.workload(args.plasmas)
    table = autotune.load()
    t0 = time.time()
    ppt = ensemble.batch_ppt(plist, table)
    b = ca.Batch(ensemble.descs(plist, table, ppt=ppt))
    ntot = ensemble.upload(b, plist, args.seed, log=log)
    b.pre_step()
    b.run(0, args.warmup)
    b.sync()
    log(f"ensemble: {len(plist)} plasmas, {ntot} particles, init {time.time() - t0:.1f} s")
    steps = args.ens_steps - args.warmup
    t0 = time.perf_counter()
    b.run(args.warmup, steps)
    b.sync()
    el = time.perf_counter() - t0
    # per-kernel times over a few more steps (launched one by one)
    b.set_profiling(True)
    b.reset_kernel_times()
    b.run(args.ens_steps, 10)
    b.sync()
    kt = b.kernel_times()
    b.set_profiling(False)
    nodes = sum(p["grid"] ** 2 for p in plist)
    push_n, push_ms = kt.get("push_gather_deposit", (0, 0.0))
    avg_push = push_ms / max(push_n, 1) / 1e3
    alg = ntot * 32 + nodes * 24
    achieved = alg / avg_push / 1e9 if avg_push > 0 else None
    value = ntot * steps / el
    grids = {g: sum(1 for p in plist if p["grid"] == g) for g in ensemble.GRIDS}
    out = {
        "metric": "particle-steps/s, batched ensemble of independent two-stream plasmas",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": el / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded two-stream electrons, host-generated)",
        "config": {"workload": f"{len(plist)} two-stream plasmas in one batch, grids "
                               f"{dict((f'{g}^2', n) for g, n in grids.items())}, 16-128 ppc, "
                               f"{ntot} particles, FP32 particles + FP64 FFT Poisson, "
                               f"captured-graph replay",
                   "plasmas": len(plist), "particles": ntot, "nodes": nodes,
                   "tiles": {str(g): list(autotune.tile_for(g, 32, table))
                             for g in ensemble.GRIDS},
                   "ppt": ppt,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": traffic_from_profiles(32, "pmc_push_fp32_ensemble_256plasmas.txt"),
                     "kernel": "push_gather_deposit", "alg_bytes_per_launch": alg,
                     "avg_launch_ms": avg_push * 1e3},
        "kernels_ms": {k: ms / max(n, 1) for k, (n, ms) in kt.items()},
    }
    b.close()
    return out


def run_adjoint(args):
    """Differentiable run: 256^2, 4,194,304 electrons (64 ppc) two-stream,
    B = 0, 200 steps, loss = sum of field energies, gradients of every
    initial position and velocity with checkpoints every 20 steps (FP64).
    Times the forward and the backward (recompute + reverse sweep)
    separately.  The initial state and the gradients cross the C ABI as host
    arrays: `value` counts the device time of both calls (HIP events from the
    end of the uploads to the start of the downloads); the wall-clock rate,
    host staging and PCIe transfers included, is reported beside it."""
    import ctypes
    import cpic_amd as ca
    nx, ppc, nsteps, K = 256, 64, 200, 20
    n = nx * nx * ppc // 2
    rng = np.random.default_rng(args.seed)
    x = rng.uniform(0.0, nx, 2 * n)
    y = rng.uniform(0.0, nx, 2 * n)
    sign = np.concatenate([np.ones(n), -np.ones(n)])
    ux = sign * 0.1 + rng.normal(0.0, 0.01, 2 * n)
    uy = rng.normal(0.0, 0.01, 2 * n)
    sp = np.concatenate([np.zeros(n), np.ones(n)]).astype(np.int32)
    desc = ca.make_desc(nx, nx, (float(nx), float(nx)), 0.1, [-1.0, -1.0], [float(ppc)] * 2,
                        precision=64)
    h = ctypes.c_void_p()
    ca.check(ca.lib().cpic_adj_create(ctypes.byref(desc), 2 * n,
                                      sp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      ctypes.byref(h)))
    loss = ctypes.c_double()
    g = [np.empty(2 * n) for _ in range(4)]

    def fwd():
        ca.check(ca.lib().cpic_adj_forward(h, ca._dp(x), ca._dp(y), ca._dp(ux), ca._dp(uy), nsteps,
                                           K, ctypes.byref(loss), None))

    def bwd():
        ca.check(ca.lib().cpic_adj_backward(h, 1.0, *(ca._dp(a) for a in g)))

    fwd()
    bwd()                                      # warm-up
    t0 = time.perf_counter()
    fwd()
    t1 = time.perf_counter()
    bwd()
    t2 = time.perf_counter()
    fms, bms = ctypes.c_double(), ctypes.c_double()
    ca.check(ca.lib().cpic_adj_device_ms(h, ctypes.byref(fms), ctypes.byref(bms)))
    ca.lib().cpic_adj_destroy(h)
    ps = 2 * n * nsteps
    dev_s = (fms.value + bms.value) / 1e3
    out = {
        "metric": "particle-steps/s of the differentiable run (forward + reverse sweep)",
        "value": ps / dev_s, "unit": "particle-steps/s", "n_gpus": 1, "steps": nsteps,
        "warmup": 1, "ms_per_step": dev_s / nsteps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded two-stream electrons, host-generated)",
        "config": {"workload": f"differentiable two-stream {nx}^2, {2 * n} particles, {nsteps} "
                               f"steps, loss = sum of field energies, checkpoint every {K}",
                   "particles": 2 * n, "grid": [nx, nx], "parallelism": "single GPU"},
        "forward_device_s": fms.value / 1e3, "backward_device_s": bms.value / 1e3,
        "forward_particle_steps_per_s": ps / (fms.value / 1e3),
        "backward_particle_steps_per_s": ps / (bms.value / 1e3),
        "wall_clock_incl_host_transfers": {
            "forward_s": t1 - t0, "backward_s": t2 - t1,
            "particle_steps_per_s": ps / (t2 - t0)},
        "loss": loss.value,
    }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=50, help="0 skips the FP64 comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--tile", type=int, default=0,
                    help="tile edge in cells (0 = the tuned shape for the grid, else the default)")
    ap.add_argument("--ppt", type=int, default=0, choices=(0, 2, 4),
                    help="push particles per thread (0 = tuned for the grid, else the default)")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps over which FP32 and FP64 field energies are compared (0 = skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 1 GiB device-to-device copies (PMC byte-counter calibration)")
    ap.add_argument("--workload", choices=("twostream", "ensemble", "adjoint"),
                    default="twostream")
    ap.add_argument("--plasmas", type=int, default=256, help="ensemble: number of plasmas")
    ap.add_argument("--ens-steps", type=int, default=300, help="ensemble: steps incl. warmup")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-ensemble", action="store_true",
                    help="default workload: skip the ensemble object")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="default workload: skip the adjoint object")
    ap.add_argument("--no-ceiling", action="store_true",
                    help="skip the HBM streaming-ceiling probe")
    ap.add_argument("--oracle-grid", type=int, default=256,
                    help="grid edge of the FP32-vs-oracle accuracy leg (0 = skip)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: the library runs on one GPU; multi-GPU runs are not implemented "
              "(the path would shard by Y slabs of tiles with halo and leaver exchange)",
              file=sys.stderr)
        sys.exit(2)

    import cpic_amd as ca
    if args.calibrate_copy:
        copy_calibration()
    if args.workload == "ensemble":
        print(json.dumps(run_ensemble(args, ca)), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps(run_adjoint(args)), flush=True)
        return
    w = workload(args.grid, args.ppc)
    ntot = sum(w["n"])
    ng = args.grid * args.grid

    from cpic_amd import autotune
    res = {}
    for prec, steps in ((32, args.steps), (64, args.fp64_steps)):
        if steps <= 0:
            continue
        t0 = time.time()
        tile = (args.tile, args.tile) if args.tile else autotune.tile_for(args.grid, prec)
        ppt = args.ppt or (0 if args.tile else autotune.ppt_for(args.grid, prec))
        d = make_device(ca, w, prec, ca.F_PROFILE, tile=tile, ppt=ppt)
        log(f"fp{prec}: init {time.time() - t0:.1f} s")
        # the library's production path (steps replayed from captured
        # graphs, no kernel timers) for the throughput ...
        d.set_profiling(False)
        d.pre_step()
        d.run(0, args.warmup)
        d.sync()
        t0 = time.perf_counter()
        d.run(args.warmup, steps)
        d.sync()
        el = time.perf_counter() - t0
        # ... then as many steps again, launched one by one between HIP
        # events, for the per-kernel times
        d.set_profiling(True)
        d.reset_kernel_times()
        t0 = time.perf_counter()
        d.run(args.warmup + steps, steps)
        d.sync()
        el_prof = time.perf_counter() - t0
        kt = d.kernel_times()
        st = d.stats()
        log(f"fp{prec}: {steps} steps in {el:.3f} s, {el / steps * 1e3:.3f} ms/step "
            f"({el_prof / steps * 1e3:.3f} ms/step with kernel timers), {st}")
        for k, (n, ms) in sorted(kt.items(), key=lambda kv: -kv[1][1]):
            log(f"   {k:24s} {n:6d} launches {ms / max(n, 1):9.4f} ms avg")
        hist = None
        snaps = {}
        late = None
        if args.accuracy_steps > 0:
            # same seed in both precisions: rho and E snapshots at fixed
            # iterations, the field-energy history up to accuracy_steps; the
            # last 100 steps are timed on the production path (late window:
            # after many migrations, not right after the upload's sort)
            d.set_profiling(False)
            done = args.warmup + 2 * steps
            for at in sorted({max(done, 200), args.accuracy_steps}):
                if at > done:
                    lw = min(100, at - done) if at == args.accuracy_steps else 0
                    if at - lw > done:
                        d.run(done, at - lw - done)
                    d.sync()
                    if lw:
                        t0 = time.perf_counter()
                        d.run(at - lw, lw)
                        d.sync()
                        late = dict(first=at - lw, steps=lw,
                                    ms=(time.perf_counter() - t0) / lw * 1e3)
                    done = at
                d.sync()
                snaps[done] = {f: d.field(f) for f in ("rho", "Ex")}
            hist = d.energy()["field"][:args.accuracy_steps]
        res[prec] = dict(elapsed=el, elapsed_prof=el_prof, steps=steps, kt=kt, stats=st,
                         hist=hist, snaps=snaps, ppt=ppt or 2, late=late)
        d.close()
        del d

    r32 = res[32]
    value = ntot * r32["steps"] / r32["elapsed"]
    push_n, push_ms = r32["kt"].get("push_gather_deposit", (0, 0.0))
    avg_push = push_ms / max(push_n, 1) / 1e3
    alg_bytes = ntot * 32 + ng * 24
    achieved = alg_bytes / avg_push / 1e9 if avg_push > 0 else None
    traffic = traffic_from_profiles(32)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["elapsed"] / r32["steps"] * 1e3,
        "ms_per_step_with_kernel_timers": r32["elapsed_prof"] / r32["steps"] * 1e3,
        "ms_per_step_late": r32["late"]["ms"] if r32["late"] else None,
        "late_window": (f"steps {r32['late']['first']}-{r32['late']['first'] + r32['late']['steps']}"
                        f" of the accuracy leg, graph replay") if r32["late"] else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-stream electrons + mobile ions, host-generated)",
        "config": {"workload": f"two-stream {args.grid}^2, {args.ppc} ppc, {ntot} particles, "
                               f"electrons + ions (1:1836), FP32 particles + FP64 FFT Poisson",
                   "grid": [args.grid, args.grid], "particles": ntot, "ppc": args.ppc,
                   "dt": w["dt"], "tile": r32["stats"]["tile"], "ppt": r32["ppt"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": traffic,
                     "kernel": "push_gather_deposit",
                     "alg_bytes_per_launch": alg_bytes,
                     "avg_launch_ms": avg_push * 1e3},
        "kernels_ms": {k: ms / max(n, 1) for k, (n, ms) in r32["kt"].items()},
    }
    if 64 in res:
        r64 = res[64]
        v64 = ntot * r64["steps"] / r64["elapsed"]
        n64, ms64 = r64["kt"].get("push_gather_deposit", (0, 0.0))
        a64 = ms64 / max(n64, 1) / 1e3
        out["fp64"] = {"value": v64, "ms_per_step": r64["elapsed"] / r64["steps"] * 1e3,
                       "ms_per_step_late": r64["late"]["ms"] if r64["late"] else None,
                       "push_achieved_GBs": (ntot * 64 + ng * 24) / a64 / 1e9 if a64 else None,
                       "tile": r64["stats"]["tile"], "ppt": r64["ppt"]}
        out["fp32_over_fp64"] = value / v64
        h32, h64 = res[32]["hist"], r64["hist"]
        if h32 is not None and h64 is not None and len(h32) and len(h64):
            n = min(len(h32), len(h64))
            rel = np.abs(h32[:n] - h64[:n]) / np.abs(h64[:n])
            above = np.nonzero(rel > 1e-5)[0]
            out["fp32_field_energy_error"] = {
                "steps": int(n), "reference": "FP64 GPU run of the same seed",
                "max_rel": float(rel.max()),
                "max_rel_first_100": float(rel[:100].max()),
                "first_step_above_1e-5": int(above[0]) if len(above) else None,
                "energy_growth": float(h64[n - 1] / h64[0]),
            }
        common = sorted(set(res[32]["snaps"]) & set(r64["snaps"]))
        if common:
            def rl2(a, b):
                return float(np.linalg.norm(a - b) / np.linalg.norm(b))
            out["fp32_field_error"] = {
                "reference": "FP64 GPU run of the same seed",
                "rho_rel_l2": {str(k): rl2(res[32]["snaps"][k]["rho"], r64["snaps"][k]["rho"])
                               for k in common},
                "Ex_rel_l2": {str(k): rl2(res[32]["snaps"][k]["Ex"], r64["snaps"][k]["Ex"])
                              for k in common},
            }
    if not args.no_ceiling:
        # the push runs one workgroup per (tile, species) bucket
        tile = r32["stats"]["tile"]
        nbuckets = 2 * (args.grid // tile[0]) * (args.grid // tile[1]) if tile else 0
        out["roofline"].update(stream_ceiling(alg_bytes, achieved,
                                              tile_slots=ntot / nbuckets if nbuckets else None))
    acc = None
    if args.oracle_grid > 0 and args.accuracy_steps > 0:
        acc = fp32_oracle_leg_gpu(ca, args.oracle_grid, args.accuracy_steps, args.seed)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(w, accuracy=acc)
        except Exception as e:  # the baseline must not hide the GPU result
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        if acc is not None:
            out["fp32_field_energy_error_vs_oracle"] = acc_report(acc)
    if not args.no_ensemble:
        e = run_ensemble(args, ca)
        out["ensemble"] = {k: e[k] for k in ("value", "unit", "ms_per_step", "steps", "dtype")}
        out["ensemble"].update(workload=e["config"]["workload"],
                               push_frac=e["roofline"]["frac"],
                               push_avg_launch_ms=e["roofline"]["avg_launch_ms"])
    if not args.no_adjoint:
        a = run_adjoint(args)
        out["adjoint"] = {k: a[k] for k in ("value", "unit", "ms_per_step", "steps", "dtype",
                                            "forward_device_s", "backward_device_s")}
        out["adjoint"].update(workload=a["config"]["workload"],
                              wall_clock_particle_steps_per_s=a["wall_clock_incl_host_transfers"][
                                  "particle_steps_per_s"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
