"""Pins the FP64 oracle before anything is checked against it.

Every function of the oracle is checked against either the reference's own
output files (tests/golden: E.csv, harm.r0x, harm.E0x written by cpic's
test/harmonic run), a known answer from the reference's test sources, an
independent implementation (numpy's pocketfft), or what the mathematics fixes
(closed forms, invariants, brute force on tiny inputs).
"""
import math
import os

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.skipif(not os.path.exists(po.LIBPATH), reason="oracle not built")


def G_half(nx, ny):
    """MFT coefficients of src/solver.c:269-280, (ny, nx//2+1) like numpy's rfft2."""
    kx = np.arange(nx // 2 + 1)
    ky = np.arange(ny)
    den = 2.0 * (np.cos(2 * np.pi / nx * kx)[None, :] + np.cos(2 * np.pi / ny * ky)[:, None]) - 4.0
    den[0, 0] = 1.0
    G = 1.0 / den
    G[0, 0] = 0.0
    return G


# ----------------------------------------------------------------- golden ---

def harmonic_oracle(n):
    """conf/harmonic.conf: two electrons at (1,4) and (7,4), L=8, dt=5e-3."""
    o = po.Oracle(n, n, (8.0, 8.0), 5e-3, q=[-1.0, -1.0], m=[1e-2, 1e-2], e0=1.0, nchunks=1)
    o.add(0, [0], [1.0], [4.0], [0.0], [0.0])
    o.add(1, [0], [7.0], [4.0], [0.0], [0.0])
    o.pre_step()
    return o


def test_E_csv_64(golden_dir):
    """test/harmonic/E.csv: the reference's 64x64 solve, row y=4, 6 digits."""
    csv = np.genfromtxt(os.path.join(golden_dir, "E.csv"), delimiter=",", skip_header=1)
    o = harmonic_oracle(64)
    row = 32
    np.testing.assert_allclose(o.field("rho")[row], csv[:, 7], atol=0)
    np.testing.assert_allclose(o.field("phi")[row], csv[:, 6], rtol=5e-6, atol=5e-7)
    np.testing.assert_allclose(o.field("Ex")[row], csv[:, 0], rtol=5e-6, atol=5e-6)
    assert np.max(np.abs(o.field("Ey")[row])) < 1e-14


def test_harmonic_1200_steps(golden_dir):
    """test/harmonic/harm.{r0x,E0x}: 1024^2, 1200 steps of the full loop
    (solve, E, gather, push with the iter-0 half step, wrap, deposit)."""
    r0x = np.loadtxt(os.path.join(golden_dir, "harm.r0x"))
    e0x = np.loadtxt(os.path.join(golden_dir, "harm.E0x"))
    o = harmonic_oracle(1024)
    rs, es = [], []
    for _ in range(len(r0x)):
        o.step()
        p = o.particles(0)
        rs.append(p["x"][0])
        es.append(p["Ex"][0])
    # values are printed with %e: 7 significant digits
    np.testing.assert_allclose(rs, r0x, rtol=0, atol=5.1e-7)
    np.testing.assert_allclose(es, e0x, rtol=6e-7, atol=0)


def test_iter0_half_step_by_hand(golden_dir):
    """harm.r0x[1] = 1 + dt * u^{1/2}, u^{-1/2} = -(dt/2)(q/m)E: the iteration-0
    half back-step, checked by hand against the reference's data."""
    o = harmonic_oracle(1024)
    o.step()                       # iter 0: velocity back half a step, no move
    p = o.particles(0)
    E = p["Ex"][0]
    dt, qm = 5e-3, -1.0 / 1e-2
    assert p["x"][0] == 1.0
    assert abs(p["ux"][0] - (-(dt / 2) * qm * E)) < 1e-15
    o.step()                       # iter 1: full step
    p = o.particles(0)
    assert abs(p["x"][0] - (1.0 + dt * ((dt / 2) * qm * E))) < 1e-12


# ---------------------------------------------------------- interpolation ---

def test_cic_known_answer():
    """test/interpolate.c.disabled:20-26: rel=(1/4,3/4) -> w00=3/16, w10=1/16,
    w01=9/16, w11=3/16, i0=(1,2)."""
    w, i0 = po.weights(1.25, 2.75, 1.0, 1.0)
    assert list(i0) == [1, 2]
    assert w[0, 0] == 3 / 16 and w[1, 0] == 1 / 16
    assert w[0, 1] == 9 / 16 and w[1, 1] == 3 / 16


def test_weights_partition_of_unity_and_brute_force():
    rng = np.random.default_rng(1)
    for _ in range(200):
        dx = rng.choice([0.125, 0.5, 1.0, 100 / 2048])
        x, y = rng.uniform(0, 64 * dx, 2)
        w, i0 = po.weights(x, y, dx, dx)
        assert abs(w.sum() - 1.0) < 1e-14
        rx, ry = x / dx - math.floor(x / dx), y / dx - math.floor(y / dx)
        assert i0[0] == math.floor(x / dx) and i0[1] == math.floor(y / dx)
        brute = np.array([[(1 - rx) * (1 - ry), (1 - rx) * ry], [rx * (1 - ry), rx * ry]])
        np.testing.assert_allclose(w, brute, atol=1e-12)


def make_random(nx, ny, L, n, seed, nspecies=1, nchunks=1, q=-1.0, m=1.0):
    rng = np.random.default_rng(seed)
    o = po.Oracle(nx, ny, L, 0.1, q=[q] * nspecies, m=[m] * nspecies, nchunks=nchunks)
    for s in range(nspecies):
        o.add(s, np.arange(n), rng.uniform(0, L[0], n), rng.uniform(0, L[1], n),
              rng.normal(0, 0.1, n), rng.normal(0, 0.1, n))
    return o, rng


def test_deposit_charge_conservation_and_node_particle():
    o, _ = make_random(16, 8, (16.0, 8.0), 500, 3, nchunks=4)
    o.comm_plasma(1)
    o.stage("field_rho")
    rho = o.field("rho")
    assert abs(rho.sum() - 500 * 1.0) < 1e-11        # c = -q/e0 = 1 per particle
    # a particle exactly on a node deposits only to that node
    o2 = po.Oracle(8, 8, (8.0, 8.0), 0.1, q=[-2.0], m=[1.0], e0=0.5)
    o2.add(0, [0], [3.0], [5.0], [0.0], [0.0])
    o2.comm_plasma(1)
    o2.stage("field_rho")
    r2 = o2.field("rho")
    assert r2[5, 3] == 4.0 and np.count_nonzero(r2) == 1


def test_gather_deposit_adjoint():
    """sum_g rho_g f_g == sum_p c_p (gather f)(x_p) for any grid field f."""
    o, rng = make_random(16, 16, (16.0, 16.0), 300, 5, nchunks=2)
    o.comm_plasma(1)
    o.stage("field_rho")
    rho = o.field("rho")
    f = rng.normal(size=(16, 16))
    o.set_field("Ex", f)
    o.set_field("Ey", np.zeros((16, 16)))
    o.stage("plasma_E")
    p = o.particles(0)
    lhs = float(np.sum(rho * f))
    rhs = float(np.sum(1.0 * p["Ex"]))
    assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(lhs))


def test_gather_ghost_row_and_wrap():
    """A particle in the last row/column reads row 0 / column 0 (ghost row
    equals row 0, X wraps by remod)."""
    o = po.Oracle(4, 4, (4.0, 4.0), 0.1, q=[-1.0], m=[1.0])
    o.add(0, [0], [3.5], [3.5], [0.0], [0.0])
    f = np.arange(16, dtype=float).reshape(4, 4)
    o.set_field("Ex", f)
    o.set_field("Ey", f.T.copy())
    o.comm_plasma(1)
    o.stage("plasma_E")
    p = o.particles(0)
    expect = 0.25 * (f[3, 3] + f[0, 3] + f[3, 0] + f[0, 0])
    assert abs(p["Ex"][0] - expect) < 1e-14


# ------------------------------------------------------------------ solver ---

@pytest.mark.parametrize("nx,ny", [(2, 2), (4, 4), (8, 16), (16, 8), (64, 64), (128, 32)])
def test_mft_matches_numpy_pocketfft(nx, ny):
    rng = np.random.default_rng(nx * 131 + ny)
    rho = rng.normal(size=(ny, nx))
    phi = po.mft_solve(rho)
    ref = np.fft.irfft2(G_half(nx, ny) * np.fft.rfft2(rho), s=(ny, nx))
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(phi - ref)) < 1e-13 * scale


@pytest.mark.parametrize("kx,ky", [(1, 0), (0, 3), (5, 7), (16, 16), (31, 1)])
def test_mft_single_mode_closed_form(kx, ky):
    nx = ny = 64
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    rho = np.cos(2 * np.pi * (kx * x / nx + ky * y / ny))
    phi = po.mft_solve(rho)
    g = 1.0 / (2.0 * (math.cos(2 * math.pi * kx / nx) + math.cos(2 * math.pi * ky / ny)) - 4.0)
    assert np.max(np.abs(phi - g * rho)) < 1e-12


def test_mft_laplacian_residual():
    nx, ny = 32, 64
    rho = np.random.default_rng(7).normal(size=(ny, nx))
    phi = po.mft_solve(rho)
    lap = (np.roll(phi, 1, 1) + np.roll(phi, -1, 1) + np.roll(phi, 1, 0) + np.roll(phi, -1, 0)
           - 4 * phi)
    assert np.max(np.abs(lap - (rho - rho.mean()))) < 1e-12
    assert abs(phi.mean()) < 1e-14


def test_fft_against_dft():
    rng = np.random.default_rng(9)
    for n in (1, 2, 4, 8, 32, 256):
        z = rng.normal(size=n) + 1j * rng.normal(size=n)
        np.testing.assert_allclose(po.fft(z, -1), np.fft.fft(z), atol=1e-12)
        np.testing.assert_allclose(po.fft(z, +1), np.fft.ifft(z) * n, atol=1e-12)


def test_E_stencil_single_mode():
    """E = -grad(phi) by central differences: for phi = sin(k x) the stencil
    gives Ex = -cos(k x) sin(k dx)/dx exactly in the mathematics."""
    nx = ny = 32
    o = po.Oracle(nx, ny, (16.0, 16.0), 0.1, q=[-1.0], m=[1.0])
    dx = 0.5
    k = 2 * np.pi * 3 / nx
    x = np.arange(nx)
    rho_mode = np.cos(k * x)[None, :].repeat(ny, 0)
    o.set_field("rho", rho_mode)
    o.stage("field_E")
    phi = o.field("phi")
    g = 1.0 / (2.0 * (math.cos(k) + 1.0) - 4.0)
    np.testing.assert_allclose(phi, g * rho_mode, atol=1e-13)
    ex = o.field("Ex")
    np.testing.assert_allclose(ex, (g * np.sin(k * x) * math.sin(k) / dx)[None, :].repeat(ny, 0),
                               atol=1e-12)
    assert np.max(np.abs(o.field("Ey"))) < 1e-13


def test_field_energy_parseval():
    """-(e0/2) sum rho*phi equals the spectral sum -(e0/2) sum_k G|rho_k|^2/N."""
    nx, ny = 32, 16
    o, rng = make_random(nx, ny, (32.0, 16.0), 2000, 11)
    o.comm_plasma(1)
    o.stage("field_rho")
    o.stage("field_E")
    rho = o.field("rho")
    rk = np.fft.rfft2(rho)
    w = np.full(nx // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    spec = -0.5 * np.sum(w[None, :] * G_half(nx, ny) * np.abs(rk) ** 2) / (nx * ny)
    assert abs(o.field_energy() - spec) < 1e-12 * abs(spec)
    assert o.field_energy() > 0


# ------------------------------------------------------------------- mover ---

def test_boris_E_only():
    r, u = po.boris([1.0, 2.0, 0.0], [0.3, -0.2, 0.0], [0.5, -1.5, 0.0], [0, 0, 0], -1.0, 2.0, 0.1)
    du = (-1.0 / 2.0) * np.array([0.5, -1.5]) * 0.1
    np.testing.assert_allclose(u[:2], np.array([0.3, -0.2]) + du, atol=1e-16)
    np.testing.assert_allclose(r[:2], np.array([1.0, 2.0]) + 0.1 * u[:2], atol=1e-16)


@pytest.mark.parametrize("q,Bz", [(-1.0, 1.0), (1.0, 0.7), (-1.0, -3.0)])
def test_boris_B_only_rotation(q, Bz):
    m, dt = 1.3, 0.05
    u = np.array([0.4, -0.1, 0.0])
    r = np.zeros(3)
    theta = -2.0 * math.atan(q * Bz * dt / (2 * m))   # du/dt = (q/m) u x B
    speed = np.hypot(*u[:2])
    for n in range(1, 200):
        r, u = po.boris(r, u, [0, 0, 0], [0, 0, Bz], q, m, dt)
        assert abs(np.hypot(*u[:2]) - speed) < 1e-14
        ang = math.atan2(u[1], u[0]) - math.atan2(-0.1, 0.4)
        d = (ang - n * theta + math.pi) % (2 * math.pi) - math.pi
        assert abs(d) < 1e-12
    assert u[2] == 0.0


def test_boris_in_plane_B_gives_uz():
    """With B in the plane the rotation creates u_z, which the reference's
    umax_Z = 0 rejects (src/sim.c:200, src/mover.c:113)."""
    _, u = po.boris([0, 0, 0], [0.4, 0.0, 0.0], [0, 0, 0], [0, 1.0, 0], -1.0, 1.0, 0.1)
    assert u[2] != 0.0
    o = po.Oracle(8, 8, (8.0, 8.0), 0.1, q=[-1.0], m=[1.0], B=(0.0, 1.0, 0.0))
    o.add(0, [0], [1.0], [1.0], [0.4], [0.0])
    o.pre_step()
    with pytest.raises(po.OracleError, match="max velocity"):
        o.step()


def test_umax_violation():
    """check_velocity aborts when |u| > chunksize*dx/dt (src/mover.c:98-137)."""
    o = po.Oracle(16, 16, (16.0, 16.0), 1.0, q=[-1.0], m=[1.0], nchunks=4)
    o.add(0, [0], [1.0], [1.0], [4.5], [0.0])     # umax_x = 4 * 1 / 1 = 4
    o.pre_step()
    with pytest.raises(po.OracleError, match="max velocity"):
        o.step()


def test_constant_speed_invariant():
    """test/constant-speed.c: one electron in conf/constant-speed.conf keeps
    its velocity to 1e-10 (zero CIC self-force with a symmetric G)."""
    o = po.Oracle(64, 64, (32.0, 32.0), 0.5, q=[-1.0], m=[1e-2], nchunks=4)
    u0 = (1.33333, 0.282828)
    o.add(0, [0], [4.0], [16.0], [u0[0]], [u0[1]])
    o.pre_step()
    o.step()                  # iter 0 pushes back half a step with E = 0
    p = o.particles(0)
    u_ref = (p["ux"][0], p["uy"][0])
    assert abs(u_ref[0] - u0[0]) < 1e-10 and abs(u_ref[1] - u0[1]) < 1e-10
    for _ in range(40):
        o.step(50)
        p = o.particles(0)
        assert abs(p["ux"][0] - u_ref[0]) < 1e-10
        assert abs(p["uy"][0] - u_ref[1]) < 1e-10
        assert 0 <= p["x"][0] < 32 and 0 <= p["y"][0] < 32


# --------------------------------------------------------------- migration ---

def test_global_exchange_places_every_particle():
    nc = 8
    o, _ = make_random(32, 16, (32.0, 16.0), 1000, 13, nspecies=2, nchunks=nc)
    o.comm_plasma(1)
    counts = o.chunk_counts()
    assert counts.sum() == 2000
    for s in range(2):
        p = o.particles(s)
        pos = dict(zip(p["id"], p["x"]))
        for c in range(nc):
            ids = o.chunk_ids(c, s)
            xs = np.array([pos[i] for i in ids])
            assert np.all(xs >= c * 4.0) and np.all(xs <= (c + 1) * 4.0)


def test_local_exchange_and_wrap_conserve_particles():
    nc = 4
    o, rng = make_random(32, 32, (32.0, 32.0), 3000, 17, nchunks=nc)
    o.pre_step()
    for _ in range(30):
        o.step()
    assert o.count(0) == 3000
    p = o.particles(0)
    assert np.all(p["x"] >= 0) and np.all(p["x"] < 32) and np.all(p["y"] >= 0) and np.all(p["y"] < 32)
    assert sorted(p["id"]) == list(range(3000))
    for c in range(nc):
        ids = o.chunk_ids(c, 0)
        pos = dict(zip(p["id"], p["x"]))
        xs = np.array([pos[i] for i in ids])
        assert np.all(xs >= c * 8.0) and np.all(xs <= (c + 1) * 8.0)


def test_wrap_to_L_maps_to_zero():
    """Documented deviation: x = -tiny wraps to L in FP64; it is mapped to 0."""
    o = po.Oracle(8, 8, (8.0, 8.0), 1.0, q=[-1.0], m=[1.0])
    o.add(0, [0], [1e-17], [4.0], [-2e-17], [0.0])
    o.pre_step()
    o.set_field("Ex", np.zeros((8, 8)))
    o.set_field("Ey", np.zeros((8, 8)))
    o.iter = 1
    o.stage("plasma_E")
    o.stage("plasma_r")
    p = o.particles(0)
    assert p["x"][0] == 0.0


def test_growth_rate_fit_recovers_known_rate():
    """The fit used for the linear-phase growth-rate comparison recovers the
    rate of a synthetic exp(2 gamma t) energy with noise and saturation."""
    from common import growth_rate, growth_window
    rng = np.random.default_rng(0)
    dt, gamma = 0.1, 0.283
    t = np.arange(400) * dt
    amp = np.exp(gamma * np.minimum(t, 25.0)) * (1 + 0.3 * np.cos(1.7 * t) * np.exp(-0.5 * t))
    le = np.log(amp ** 2 + 1.0) + 1e-3 * rng.normal(size=t.size)
    w = growth_window(le)
    assert w[1] - w[0] > 20
    assert abs(growth_rate(le, dt, w) - gamma) < 1e-3 * gamma * 10
