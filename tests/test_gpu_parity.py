"""HIP path vs the FP64 oracle, through the C ABI (include/cpic_hip.h).

Bar: bit-exact for cell indices, per-cell counts, migration lists and, on a
shared input state, for the gathered E, pushed velocities and positions;
1e-12 relative L2 for rho, E, particle state and the field-energy history
over the first 20 steps (FP64 mode)."""
import os

import numpy as np
import pytest

import cpic_amd as ca
import pyoracle as po
from common import growth_rate, growth_window, rel_l2, seeded_two_stream, two_stream

pytestmark = pytest.mark.gpu

FLAGS = ca.F_IDS | ca.F_PARTICLE_E


def make_pair(nx, ny, L, dt, q, m, parts, nchunks=1, precision=64, tile=(0, 0), flags=FLAGS,
              B=(0.0, 0.0, 0.0), e0=1.0, slack=0.0, ppt=0):
    o = po.Oracle(nx, ny, L, dt, q, m, e0=e0, B=B, nchunks=nchunks)
    d = ca.Device(nx, ny, L, dt, q, m, e0=e0, B=B, nchunks=nchunks, precision=precision,
                  tile=tile, flags=flags, slack=slack, ppt=ppt)
    for s, (ids, x, y, ux, uy) in enumerate(parts):
        o.add(s, ids, x, y, ux, uy)
        d.upload(s, ids, x, y, ux, uy)
    return o, d


def c1(seed=1, n=5120, precision=64, tile=(0, 0), nchunks=1, slack=0.0, flags=FLAGS, ppt=0):
    """64x64 two-stream: 2 x 5120 electrons, q=-1, m such that w_p = 1."""
    L = (64.0, 64.0)
    m = 2 * n / (L[0] * L[1])
    parts = two_stream(n, L, 0.1, 0.01, seed)
    return make_pair(64, 64, L, 0.1, [-1.0, -1.0], [m, m], parts, nchunks=nchunks,
                     precision=precision, tile=tile, slack=slack, flags=flags, ppt=ppt)


def assert_same_particles(o, d, ns, exact=True, tol=1e-12):
    for s in range(ns):
        a = o.particles(s)
        b = d.particles(s)
        assert np.array_equal(a["id"], b["id"])
        for k in ("x", "y", "ux", "uy"):
            if exact:
                assert np.array_equal(a[k], b[k]), (s, k, np.max(np.abs(a[k] - b[k])))
            else:
                assert rel_l2(b[k], a[k]) < tol, (s, k, rel_l2(b[k], a[k]))


def test_library_exports_and_device():
    d = ca.Device(16, 16, (16.0, 16.0), 0.1, [-1.0], [1.0])
    st = d.stats()
    assert st["tile"] == [16, 16] and st["precision"] == 64
    d.close()


def test_pre_step_fields_match_oracle():
    o, d = c1()
    o.pre_step()
    d.pre_step()
    for f in ("rho", "phi", "Ex", "Ey"):
        assert rel_l2(d.field(f), o.field(f)) < 1e-13, f


@pytest.mark.parametrize("it", [0, 1])
def test_push_bitwise_on_shared_state(it):
    """Same E field and particles: gather, Boris push, wrap are bit-exact."""
    o, d = c1(seed=2)
    o.pre_step()
    d.pre_step()
    d.set_field("Ex", o.field("Ex"))
    d.set_field("Ey", o.field("Ey"))
    o.iter = it
    o.stage("plasma_E")
    o.stage("plasma_r")
    d.stage("plasma_E", it)
    d.stage("plasma_r", it)
    for s in range(2):
        a, b = o.particles(s), d.particles(s)
        assert np.array_equal(a["id"], b["id"])
        for k in ("Ex", "Ey", "x", "y", "ux", "uy"):
            assert np.array_equal(a[k], b[k]), (s, k, np.max(np.abs(a[k] - b[k])))


def test_migration_and_cells_bitwise():
    """Cell indices, per-cell counts and the list of tile leavers (with
    destinations) equal what the oracle's positions give."""
    o, d = c1(seed=3, tile=(8, 8))
    o.pre_step()
    d.pre_step()
    d.set_field("Ex", o.field("Ex"))
    d.set_field("Ey", o.field("Ey"))
    before = {s: o.particles(s) for s in range(2)}
    o.iter = 1
    o.stage("plasma_E")
    o.stage("plasma_r")
    d.stage("plasma_E", 1)
    d.stage("plasma_r", 1)
    mig = d.migration()
    d.stage("field_rho", 1)
    exp_ids, exp_dst = [], []
    for s in range(2):
        a = o.particles(s)
        cells = o.cells(a["x"], a["y"])
        b = d.particles(s)
        assert np.array_equal(b["cell"], cells)
        counts = np.bincount(cells, minlength=64 * 64).reshape(64, 64)
        assert np.array_equal(d.cell_counts(s), counts)
        c0 = o.cells(before[s]["x"], before[s]["y"])
        t0 = (c0 // 64 // 8) * 8 + (c0 % 64) // 8
        t1 = (cells // 64 // 8) * 8 + (cells % 64) // 8
        moved = t0 != t1
        exp_ids += [(s, i) for i in a["id"][moved]]
        exp_dst += list(zip([s] * moved.sum(), a["id"][moved], t1[moved]))
    got = sorted(zip(mig["species"].tolist(), mig["id"].tolist(), mig["dst"].tolist()))
    assert got == sorted((int(s), int(i), int(t)) for s, i, t in exp_dst)
    assert len(got) > 0


def run_both(o, d, nsteps):
    """Runs pre_step + nsteps on both, returning per-step field energies."""
    o.pre_step()
    d.pre_step()
    W = []
    for it in range(nsteps):
        o.stage("field_E")
        W.append(o.field_energy())
        o.stage("plasma_E")
        o.stage("plasma_r")
        o.stage("field_rho")
        o.iter = it + 1
        d.step(it)
    return np.array(W)


@pytest.mark.parametrize("tile", [(16, 16), (8, 8), (4, 16), (16, 8), (32, 16), (64, 64)])
def test_20_steps_fp64_rel_l2(tile):
    o, d = c1(seed=4, tile=tile)
    W = run_both(o, d, 20)
    d.sync()
    for f in ("rho", "phi", "Ex", "Ey"):
        assert rel_l2(d.field(f), o.field(f)) < 1e-12, f
    assert_same_particles(o, d, 2, exact=False, tol=1e-12)
    e = d.energy()
    assert len(e["field"]) == 20
    assert rel_l2(e["field"], W) < 1e-12


def test_20_steps_chunks_and_tight_capacity():
    """nchunks sets umax only; a capacity with no slack exercises the
    fallback path (leavers kept in their source tile) without changing
    results."""
    o, d = c1(seed=5, tile=(4, 4), nchunks=4, slack=-1.0)
    run_both(o, d, 20)
    d.sync()
    st = d.stats()
    assert st["leavers"] > 0
    assert rel_l2(d.field("rho"), o.field("rho")) < 1e-12
    assert_same_particles(o, d, 2, exact=False, tol=1e-12)


@pytest.mark.parametrize("tile,ppt", [((16, 16), 2), ((8, 8), 4), ((64, 64), 2)])
def test_20_steps_push_computes_E_from_phi(tile, ppt):
    """Without per-particle E output nothing reads the E array during a step,
    so the push computes its tiles' E from phi: the same 1e-12 bar over 20
    steps, and E read afterwards (computed on demand) matches the oracle's."""
    o, d = c1(seed=4, tile=tile, flags=ca.F_IDS, ppt=ppt)
    W = run_both(o, d, 20)
    d.sync()
    for f in ("rho", "phi", "Ex", "Ey"):
        assert rel_l2(d.field(f), o.field(f)) < 1e-12, f
    assert_same_particles(o, d, 2, exact=False, tol=1e-12)
    assert rel_l2(d.energy()["field"], W) < 1e-12


@pytest.mark.parametrize("precision", [64, 32])
def test_push_E_from_phi_bitwise_equals_E_array(precision):
    """One step from a shared rho gives the same particles bit for bit
    whether the push reads the E array (per-particle E output requested),
    computes E from phi, or phi is overwritten between field_E and the push
    (E keeps the value field_E gave it, as in the reference)."""
    got = []
    rho = None
    for flags, clobber in ((ca.F_IDS | ca.F_PARTICLE_E, False), (ca.F_IDS, False),
                           (ca.F_IDS, True)):
        o, d = c1(seed=8, tile=(8, 8), flags=flags, precision=precision)
        o.pre_step()
        d.pre_step()
        if rho is None:
            rho = o.field("rho")
        d.set_field("rho", rho)
        d.stage("field_E", 1)
        if clobber:
            d.set_field("phi", np.zeros_like(rho))
        d.stage("plasma_E", 1)
        d.stage("plasma_r", 1)
        got.append([d.particles(s) for s in range(2)])
        d.close()
    for run in got[1:]:
        for s in range(2):
            for k in ("id", "x", "y", "ux", "uy"):
                assert np.array_equal(run[s][k], got[0][s][k]), (s, k)


def test_rebuild_midway():
    o, d = c1(seed=6)
    o.pre_step()
    d.pre_step()
    for it in range(10):
        o.step()
        d.step(it)
    d.rebuild()
    for it in range(10, 20):
        o.step()
        d.step(it)
    assert rel_l2(d.field("rho"), o.field("rho")) < 1e-12
    assert_same_particles(o, d, 2, exact=False, tol=1e-12)


def test_kinetic_energy_diag():
    o, d = c1(seed=7)
    o.pre_step()
    d.pre_step()
    m = o.params.m[0]
    K = []
    for it in range(5):
        u0 = [o.particles(s) for s in range(2)]
        o.step()
        u1 = [o.particles(s) for s in range(2)]
        K.append(sum(0.25 * m * (np.sum(a["ux"] ** 2 + a["uy"] ** 2) + np.sum(b["ux"] ** 2 + b["uy"] ** 2))
                     for a, b in zip(u0, u1)))
        d.step(it)
    e = d.energy()
    assert rel_l2(e["kinetic"], K) < 1e-12


def test_magnetised_fp64():
    """Bz with w_c/w_p = 0.5 (Boris rotation active), 32^2, 20 steps."""
    L = (32.0, 32.0)
    n = 4096
    m = 2 * n / (L[0] * L[1])
    parts = two_stream(n, L, 0.1, 0.01, seed=8)
    Bz = 0.5 * m          # w_c = |q| B / m
    o, d = make_pair(32, 32, L, 0.05, [-1.0, -1.0], [m, m], parts, B=(0.0, 0.0, Bz),
                     tile=(8, 8))
    run_both(o, d, 20)
    assert rel_l2(d.field("Ex"), o.field("Ex")) < 1e-12
    assert_same_particles(o, d, 2, exact=False, tol=1e-12)


def test_harmonic_fixture_gpu(golden_dir):
    """The reference's harm.r0x/harm.E0x (1024^2, 1200 steps) on the GPU."""
    r0x = np.loadtxt(os.path.join(golden_dir, "harm.r0x"))
    e0x = np.loadtxt(os.path.join(golden_dir, "harm.E0x"))
    d = ca.Device(1024, 1024, (8.0, 8.0), 5e-3, [-1.0, -1.0], [1e-2, 1e-2], flags=FLAGS)
    d.upload(0, [0], [1.0], [4.0], [0.0], [0.0])
    d.upload(1, [0], [7.0], [4.0], [0.0], [0.0])
    d.pre_step()
    rs, es = [], []
    for it in range(len(r0x)):
        d.step(it)
        p = d.particles(0)
        rs.append(p["x"][0])
        es.append(p["Ex"][0])
    np.testing.assert_allclose(rs, r0x, rtol=0, atol=5.1e-7)
    np.testing.assert_allclose(es, e0x, rtol=6e-7, atol=0)


def test_constant_speed_gpu():
    """test/constant-speed.c: 50,000 steps, velocity constant to 1e-10."""
    d = ca.Device(64, 64, (32.0, 32.0), 0.5, [-1.0], [1e-2], nchunks=4, flags=FLAGS)
    d.upload(0, [0], [4.0], [16.0], [1.33333], [0.282828])
    d.pre_step()
    d.step(0)
    u = d.particles(0)
    ref = (u["ux"][0], u["uy"][0])
    it = 1
    for _ in range(50):
        d.run(it, 1000)
        it += 1000
        p = d.particles(0)
        assert abs(p["ux"][0] - ref[0]) < 1e-10 and abs(p["uy"][0] - ref[1]) < 1e-10


def test_fp32_close_to_fp64():
    o, d = c1(seed=9, precision=32)
    W = run_both(o, d, 20)
    d.sync()
    assert rel_l2(d.field("rho"), o.field("rho")) < 1e-5
    assert rel_l2(d.field("Ex"), o.field("Ex")) < 1e-4
    assert_same_particles(o, d, 2, exact=False, tol=1e-5)
    e = d.energy()
    err = rel_l2(e["field"], W)
    print(f"FP32 field-energy history vs the FP64 oracle, 20 steps: rel L2 {err:.2e}")
    assert err < 1e-5


def test_umax_violation_reported():
    d = ca.Device(16, 16, (16.0, 16.0), 1.0, [-1.0], [1.0], nchunks=4)
    d.upload(0, [0], [1.0], [1.0], [4.5], [0.0])
    d.pre_step()
    d.step(0)
    with pytest.raises(ca.CpicError) as ei:
        d.sync()
    assert ei.value.code == -3


def test_failed_resort_leaves_a_handle_that_refuses_work(monkeypatch):
    """A re-sort that fails part way (here a particle-buffer allocation,
    made to fail from the host) frees a buffer the push writes: the call
    returns CPIC_ENOMEM and every later call on the handle but destroy
    returns CPIC_ESTATE, without launching a kernel."""
    o, d = c1(seed=4)
    d.pre_step()
    for it in range(3):
        d.step(it)
    d.sync()
    # eight times the particles of species 0: the next commit must allocate
    ids, x, y, ux, uy = two_stream(8 * 5120, (64.0, 64.0), 0.1, 0.01, 5)[0]
    d.upload(0, ids, x, y, ux, uy)
    monkeypatch.setenv("CPIC_FAIL_PARTS_ALLOC", "1")
    with pytest.raises(ca.CpicError) as ei:
        d.rebuild()
    assert ei.value.code == -2
    monkeypatch.delenv("CPIC_FAIL_PARTS_ALLOC")
    calls = [lambda: d.step(3), lambda: d.run(3, 4), d.sync, lambda: d.field("rho"),
             lambda: d.particles(1), d.rebuild, d.pre_step, lambda: d.stage("plasma_r", 3),
             lambda: d.upload(1, ids, x, y, ux, uy), d.energy, d.solve_xy]
    for k, call in enumerate(calls):
        with pytest.raises(ca.CpicError) as ei:
            call()
        assert ei.value.code == -4, k
    d.close()


def test_bucket_over_the_slot_limit_refused():
    """One (tile, species) bucket holds at most 2^27 slots (the push
    addresses a bucket with 32-bit byte offsets): a commit whose bucket would
    need more (110 M particles in one 4x4 tile, 1.25x capacity) returns
    CPIC_ENOMEM before it changes the particle storage, and the handle keeps
    working with an upload that fits."""
    d = ca.Device(4, 4, (4.0, 4.0), 0.1, [-1.0], [1.0], precision=32, tile=(4, 4))
    rng = np.random.default_rng(21)
    n = 110_000_000
    empty = np.empty(0, np.int64)
    x = rng.uniform(0.0, 4.0, n)
    y = rng.uniform(0.0, 4.0, n)
    u = np.zeros(n)
    d.upload(0, empty, x, y, u, u)
    del x, y
    with pytest.raises(ca.CpicError) as ei:
        d.pre_step()
    assert ei.value.code == -2
    k = 1000
    d.upload(0, empty, rng.uniform(0.0, 4.0, k), rng.uniform(0.0, 4.0, k), np.zeros(k),
             np.zeros(k))
    d.pre_step()
    d.run(0, 3)
    d.sync()
    assert len(d.particles(0)["x"]) == k
    d.close()


def test_in_plane_B_rejected():
    with pytest.raises(ca.CpicError) as ei:
        ca.Device(16, 16, (16.0, 16.0), 0.1, [-1.0], [1.0], B=(0.0, 1.0, 0.0))
    assert ei.value.code == -1


@pytest.mark.parametrize("precision", [64, 32])
def test_two_stream_linear_growth_rate(precision):
    """Linear-phase growth rate of a seeded two-stream mode (64x8 grid, 1024
    ppc, v0 = 1, mode 6 near the fastest-growing wavenumber, 300 steps):
    the rate fitted from the GPU's rho history is within 1e-3 (relative) of
    the rate fitted from the oracle's, over the same window."""
    nx, ny, ppc, M = 64, 8, 1024, 6
    parts = seeded_two_stream(nx, ny, ppc, 1.0, 0.01, M, 0.003, seed=5)
    m = float(ppc)
    o, d = make_pair(nx, ny, (float(nx), float(ny)), 0.1, [-1.0, -1.0], [m, m], parts,
                     precision=precision, flags=0)
    o.pre_step()
    d.pre_step()
    lo, lg = [], []
    for it in range(300):
        o.stage("field_E")
        lo.append(np.log(np.abs(np.fft.rfft2(o.field("rho"))[0, M]) ** 2))
        lg.append(np.log(np.abs(np.fft.rfft2(d.field("rho"))[0, M]) ** 2))
        o.stage("plasma_E")
        o.stage("plasma_r")
        o.stage("field_rho")
        o.iter = it + 1
        d.step(it)
    w = growth_window(lo)
    assert w[1] - w[0] >= 20, w
    go = growth_rate(lo, 0.1, w)
    gg = growth_rate(lg, 0.1, w)
    k = 2 * np.pi * M / nx
    wb = 1 / np.sqrt(2)
    theory = np.sqrt(-(k * k + wb * wb - wb * np.sqrt(4 * k * k + wb * wb)))
    print(f"fp{precision}: window {w}, gamma oracle {go:.6f} gpu {gg:.6f} "
          f"(rel {abs(gg - go) / go:.2e}); cold-beam theory {theory:.4f}", flush=True)
    assert 0.5 * theory < go < 1.2 * theory
    assert abs(gg - go) <= 1e-3 * go


def test_c2_scale_fp64_matches_oracle():
    """The FP64 drop-in path at scale: 512^2, 2 x 8,388,608 counter-streaming
    electrons (64 per cell), Bz with w_c/w_p = 0.5, dt = 0.05; 20 steps
    against the oracle at 1e-12 (rho, phi, E, particles, field energy), then
    on to 500 steps on the GPU with charge and particle-count invariants."""
    import time
    nx = 512
    n = nx * nx * 64 // 2
    L = (float(nx), float(nx))
    m = 64.0
    t0 = time.time()
    parts = two_stream(n, L, 0.1, 0.01, seed=77)
    o, d = make_pair(nx, nx, L, 0.05, [-1.0, -1.0], [m, m], parts, B=(0.0, 0.0, 32.0),
                     nchunks=8, flags=ca.F_IDS)
    del parts
    print(f"setup {time.time() - t0:.1f} s", flush=True)
    W = run_both(o, d, 20)
    d.sync()
    print(f"20 steps on both {time.time() - t0:.1f} s", flush=True)
    for f in ("rho", "phi", "Ex", "Ey"):
        assert rel_l2(d.field(f), o.field(f)) < 1e-12, f
    for s in range(2):
        a, b = o.particles(s), d.particles(s)
        assert np.array_equal(a["id"], b["id"])
        for k in ("x", "y", "ux", "uy"):
            assert rel_l2(b[k], a[k]) < 1e-12, (s, k)
    del o
    e = d.energy()
    assert rel_l2(e["field"], W) < 1e-12
    d.run(20, 480)
    d.sync()
    rho = d.field("rho")
    assert abs(rho.sum() - 2 * n) < 1e-9 * 2 * n       # -q/e0 = 1 per electron
    assert d.count(0) == n and d.count(1) == n
    assert np.all(np.isfinite(d.energy()["field"][:500]))
    print(f"500 steps {time.time() - t0:.1f} s", flush=True)


@pytest.mark.parametrize("precision,tol,tile", [(64, 1e-12, (8, 8)), (32, 1e-5, (8, 8)),
                                                (32, 1e-5, (32, 16)), (32, 1e-5, (16, 8))])
def test_four_particles_per_thread(precision, tol, tile):
    """The push kernel's ppt = 4 instantiation (autotuner choice) gives the
    same results as the default, at the tuned tile shapes too."""
    o, d = c1(seed=12, precision=precision, ppt=4, tile=tile)
    W = run_both(o, d, 20)
    d.sync()
    assert rel_l2(d.field("rho"), o.field("rho")) < tol
    assert_same_particles(o, d, 2, exact=False, tol=tol)
    assert rel_l2(d.energy()["field"], W) < max(tol, 1e-12) * 10


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 1e-5)])
def test_fast_streams_overflow_leaver_stage_and_buffer(precision, tol):
    """Beams at 2 cells per step on 16x16 tiles: about 1,000 leavers per
    bucket per step (more than the push's LDS stage of 256 holds, so the
    per-wavefront fallback runs) and ~35,000 per step against a leaver
    buffer of 16,384 (the rest stay in their source bucket, misplaced, and
    are gathered and deposited through the global-memory path).  Results
    are unchanged."""
    n, L = 131072, (64.0, 64.0)
    m = 2 * n / (L[0] * L[1])
    parts = two_stream(n, L, 20.0, 2.0, seed=21)
    o, d = make_pair(64, 64, L, 0.1, [-1.0, -1.0], [m, m], parts, precision=precision,
                     tile=(16, 16), flags=ca.F_IDS, slack=0.25)
    # leavers per (tile, species) bucket in the first moving step, from the
    # oracle's positions
    o.pre_step()
    o.step()
    before = [o.particles(s) for s in range(2)]
    o.step()
    most = 0
    for s in range(2):
        a = o.particles(s)
        c0, c1_ = o.cells(before[s]["x"], before[s]["y"]), o.cells(a["x"], a["y"])
        t0 = (c0 // 64 // 16) * 4 + (c0 % 64) // 16
        t1 = (c1_ // 64 // 16) * 4 + (c1_ % 64) // 16
        most = max(most, np.bincount(t0[t0 != t1], minlength=16).max())
    assert most > 256
    o, d = make_pair(64, 64, L, 0.1, [-1.0, -1.0], [m, m], parts, precision=precision,
                     tile=(16, 16), flags=ca.F_IDS, slack=0.25)
    W = run_both(o, d, 20)
    d.sync()
    st = d.stats()
    assert st["misplaced"] > 0
    assert rel_l2(d.field("rho"), o.field("rho")) < tol
    assert_same_particles(o, d, 2, exact=False, tol=tol)
    assert rel_l2(d.energy()["field"], W) < max(tol, 1e-12) * 10
    if precision == 32:
        return           # FP32 positions may sit across a cell edge from FP64 ones
    for s in range(2):
        a = o.particles(s)
        assert np.array_equal(d.cell_counts(s), np.bincount(
            o.cells(a["x"], a["y"]), minlength=64 * 64).reshape(64, 64))


def test_empty_single_and_ragged_species():
    """Species with 999 particles (no whole wavefront or block), one with a
    single particle and one with none: 20 steps match the oracle at 1e-12;
    the empty species downloads empty."""
    L = (64.0, 64.0)
    parts = two_stream(999, L, 0.1, 0.01, 31)
    one = (np.array([7], np.int64), np.array([10.3]), np.array([20.7]), np.array([0.05]),
           np.array([-0.02]))
    empty = (np.empty(0, np.int64),) + (np.empty(0),) * 4
    parts = [parts[0], parts[1], one, empty]
    m = 2 * 999 / (L[0] * L[1])
    o, d = make_pair(64, 64, L, 0.1, [-1.0, -1.0, 1.0, 1.0], [m, m, 50.0, 50.0], parts,
                     tile=(16, 16))
    W = run_both(o, d, 20)
    d.sync()
    for f in ("rho", "phi", "Ex", "Ey"):
        assert rel_l2(d.field(f), o.field(f)) < 1e-12, f
    assert_same_particles(o, d, 3, exact=False, tol=1e-12)
    assert d.count(3) == 0 and d.particles(3)["x"].size == 0
    assert rel_l2(d.energy()["field"], W) < 1e-12


@pytest.mark.parametrize("nx,ny", [(8192, 8192), (8192, 2), (2, 8192)])
def test_solve_at_maximum_and_extreme_grids(nx, ny):
    """The MFT solve at the largest grid the library accepts (8192^2: one
    row per row-pass block, one 128 KB column per column-pass block, the
    LDS twiddle factors at N = 8192) and at the thinnest ones: for rho a sum
    of two Fourier modes, phi = G(k) rho mode by mode (closed form).  The
    solve is ill-conditioned at these sizes: the rounding of the sampled
    modes (~eps) reaches the lowest modes, where |G| is largest (about
    (N / 2 pi)^2 = 1.7e6 at N = 8192), so the bound is 64 eps max|G| sum|a|
    (numpy's pocketfft misses the closed form by the same 1e-11 at 8192 x 2;
    a wrong G, mode or normalisation misses it by O(|phi|))."""
    import math
    d = ca.Device(nx, ny, (float(nx), float(ny)), 0.1, [-1.0], [1.0])
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    modes = [(min(3, nx // 2), min(5, ny // 2), 1.0), (nx // 2, 1 if ny > 2 else 0, 0.25)]
    rho = np.zeros((ny, nx))
    phi = np.zeros((ny, nx))
    n_max = max(nx, ny)
    gmax = 1.0 / (2.0 - 2.0 * math.cos(2 * math.pi / n_max))
    tol = 64 * np.finfo(float).eps * gmax * sum(abs(a) for _, _, a in modes)
    for kx, ky, a in modes:
        if kx == 0 and ky == 0:
            continue
        g = np.cos(2 * np.pi * (kx * x / nx + ky * y / ny)) * a
        G = 1.0 / (2.0 * (math.cos(2 * math.pi * kx / nx) + math.cos(2 * math.pi * ky / ny)) - 4.0)
        rho += g
        phi += G * g
    del x, y, g
    d.set_field("rho", rho)
    d.solve_xy()
    got = d.field("phi")
    err = np.max(np.abs(got - phi))
    assert err < tol, (nx, ny, err, tol)
    assert np.max(np.abs(phi)) > 1e3 * tol          # the check is not vacuous
    d.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_long_run_conservation_with_fast_beams(precision):
    """3,000 steps of fast beams (0.5 cell per step) on 8x8 tiles, replayed
    from captured graphs with a re-sort after step 1,500: the particle count
    of each species, the total charge sum(rho) = N and finiteness hold at
    every check; leavers cross tiles every step.  About 7 % of the particles
    leave their tile per step, more than the initial leaver buffer (N/32)
    holds: the first sync grows it and re-sorts, so the second 500 steps
    misplace far fewer particles than the first."""
    n, L = 262144, (128.0, 128.0)
    m = 2 * n / (L[0] * L[1])
    parts = two_stream(n, L, 5.0, 0.5, seed=41)
    d = ca.Device(128, 128, L, 0.1, [-1.0, -1.0], [m, m], precision=precision, tile=(8, 8))
    for s, (ids, x, y, ux, uy) in enumerate(parts):
        d.upload(s, ids, x, y, ux, uy)
    d.pre_step()
    it = 0
    tol = 1e-9 if precision == 64 else 1e-5
    misplaced = []
    for block in range(6):
        d.run(it, 500)
        it += 500
        d.sync()
        misplaced.append(d.stats()["misplaced"])
        assert d.count(0) == n and d.count(1) == n, it
        rho = d.field("rho")
        assert np.all(np.isfinite(rho)), it
        assert abs(rho.sum() - 2 * n) < tol * 2 * n, (it, rho.sum())   # -q/e0 = 1 each
        if block == 2:
            d.rebuild()
    st = d.stats()
    assert st["leavers"] > 100 * n // 8
    assert misplaced[0] > 0 and misplaced[1] - misplaced[0] < misplaced[0] // 4, misplaced
    W = d.energy()["field"][:it]
    assert np.all(np.isfinite(W))
    d.close()
