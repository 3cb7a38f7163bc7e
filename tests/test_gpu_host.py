"""The C host driver on the GPU: the reference's own test programs
(test/harmonic.c, test/constant-speed.c) replayed through sim_load/sim_step,
and the `cpic` command line on configuration files."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from common import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cpic_amd", "lib", "libcpic_host.so")
BIN = os.path.join(ROOT, "cpic_amd", "bin", "cpic")
CONF = os.path.join(ROOT, "tests", "golden", "conf")

pytestmark = pytest.mark.gpu


def host():
    h = C.CDLL(LIB)
    h.sim_load.restype = C.c_void_p
    h.sim_load.argtypes = [C.c_char_p, C.c_int]
    h.sim_step.argtypes = [C.c_void_p]
    h.sim_free.argtypes = [C.c_void_p]
    h.sim_particle.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double)]
    h.sim_get_iter.restype = C.c_int64
    h.sim_get_iter.argtypes = [C.c_void_p]
    h.sim_get_cycles.restype = C.c_int64
    h.sim_get_cycles.argtypes = [C.c_void_p]
    return h


def tracked_copy(src, d, extra=""):
    text = open(src).read().replace("simulation = {", "simulation = {\n\ttrack_particles = true\n" + extra, 1)
    path = os.path.join(d, os.path.basename(src))
    with open(path, "w") as f:
        f.write(text)
    return path


def particle(h, s, sp, idx):
    out = (C.c_double * 6)()
    assert h.sim_particle(s, sp, idx, out) == 0
    return list(out)


def test_harmonic_through_host_driver(golden_dir):
    """test/harmonic.c: after every sim_step, r[X] and E[X] of electron 0
    match harm.r0x / harm.E0x."""
    r0x = np.loadtxt(os.path.join(golden_dir, "harm.r0x"))
    e0x = np.loadtxt(os.path.join(golden_dir, "harm.E0x"))
    h = host()
    with tempfile.TemporaryDirectory() as d:
        s = h.sim_load(tracked_copy(os.path.join(CONF, "harmonic.conf"), d).encode(), 1)
    assert s
    assert h.sim_get_iter(s) == 0 and h.sim_get_cycles(s) == 1200
    rs, es = [], []
    while h.sim_get_iter(s) < h.sim_get_cycles(s):
        assert h.sim_step(s) == 0
        p = particle(h, s, 0, 0)
        rs.append(p[0])
        es.append(p[4])
    h.sim_free(s)
    np.testing.assert_allclose(rs, r0x, rtol=0, atol=5.1e-7)
    np.testing.assert_allclose(es, e0x, rtol=6e-7, atol=0)


def test_constant_speed_through_host_driver():
    """test/constant-speed.c: |u - u0| <= 1e-10 at every step (5,000 of the
    50,000 here; the full run is in test_gpu_parity)."""
    h = host()
    with tempfile.TemporaryDirectory() as d:
        s = h.sim_load(tracked_copy(os.path.join(CONF, "constant-speed.conf"), d).encode(), 1)
    assert s
    u0 = particle(h, s, 0, 0)[2:4]
    for _ in range(5000):
        assert h.sim_step(s) == 0
        if h.sim_get_iter(s) % 250 == 0:
            u = particle(h, s, 0, 0)[2:4]
            assert abs(u[0] - u0[0]) < 1e-10 and abs(u[1] - u0[1]) < 1e-10
    h.sim_free(s)


def test_cyclotron_through_host_driver():
    """test/cyclotron.c (conf/cyclotron.conf: one electron, q = -1, m = 1, in
    B = (0, 0, -1) on a 4x4 grid, dt = 0.05, 5,000 steps) through
    sim_load/sim_step.  With E = 0 the Boris step rotates u by exactly
    -2 atan(q B dt / 2m) per step and keeps |u|, and the positions lie on a
    circle of radius dt |u| / (2 sin(theta/2)).  The particle's own field is
    zero only in exact arithmetic: with e0 = 8.85e-12 its rounding residual
    (E ~ 1e-7 at the particle, the same in the oracle) bounds the checks at
    1e-7 relative.  The centre is taken from the orbit itself (the
    reference's test computes it on the wrong side of the particle).
    The trajectory also matches the oracle's step by step."""
    import pyoracle as po
    h = host()
    with tempfile.TemporaryDirectory() as d:
        s = h.sim_load(tracked_copy(os.path.join(CONF, "cyclotron.conf"), d).encode(), 1)
    assert s
    assert h.sim_get_cycles(s) == 5000
    traj = []
    while h.sim_get_iter(s) < h.sim_get_cycles(s):
        assert h.sim_step(s) == 0
        traj.append(particle(h, s, 0, 0)[:4])
    h.sim_free(s)
    traj = np.array(traj)
    o = po.Oracle(4, 4, (32.0, 32.0), 0.05, [-1.0], [1.0], e0=8.85e-12, B=(0.0, 0.0, -1.0),
                  nchunks=4)
    o.add(0, np.array([0]), np.array([2.0]), np.array([16.0]), np.array([0.0]), np.array([10.0]))
    o.pre_step()
    ref = []
    for _ in range(5000):
        o.step()
        p = o.particles(0)
        ref.append([p["x"][0], p["y"][0], p["ux"][0], p["uy"][0]])
    ref = np.array(ref)
    assert np.max(np.abs(traj - ref)) < 1e-5, np.max(np.abs(traj - ref), axis=0)
    speed = np.hypot(traj[:, 2], traj[:, 3])
    assert np.max(np.abs(speed - 10.0)) <= 1e-7 * 10.0
    theta = -2.0 * np.arctan((-1.0) * (-1.0) * 0.05 / (2.0 * 1.0))
    dang = np.diff(np.unwrap(np.arctan2(traj[:, 3], traj[:, 2])))
    assert np.max(np.abs(dang - theta)) <= 1e-7, (dang.min(), dang.max(), theta)
    # circumcentre of three positions after the first move
    (ax, ay), (bx, by), (cx, cy) = traj[1, :2], traj[200, :2], traj[400, :2]
    dd = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
    ux_ = ((ax * ax + ay * ay) * (by - cy) + (bx * bx + by * by) * (cy - ay) +
           (cx * cx + cy * cy) * (ay - by)) / dd
    uy_ = ((ax * ax + ay * ay) * (cx - bx) + (bx * bx + by * by) * (ax - cx) +
           (cx * cx + cy * cy) * (bx - ax)) / dd
    R = 0.05 * 10.0 / (2 * np.sin(abs(theta) / 2))
    dist = np.hypot(traj[1:, 0] - ux_, traj[1:, 1] - uy_)
    assert np.max(np.abs(dist - R)) <= 1e-6 * R, (np.max(np.abs(dist - R)), R)
    # clockwise about a centre to the right of the start (x0 + R, y0)
    assert abs(ux_ - (2.0 + R)) < 1e-3 and abs(uy_ - 16.0) < 0.5


def test_cpic_cli_stats_line():
    """cpic prints the reference's stats line when sampling (src/sim.c:465)."""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(ROOT, "cpic_amd", "conf", "two-stream-64.conf")
        text = open(src).read().replace("stop_SEM = 0.0", "stop_SEM = 0.5").replace(
            "cycles = 200", "cycles = 40")
        path = os.path.join(d, "c.conf")
        with open(path, "w") as f:
            f.write(text)
        r = subprocess.run([BIN, "-q", path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("stats ")]
    assert lines, r.stdout
    fields = dict(kv.split("=") for kv in lines[-1].split()[1:])
    assert set(fields) == {"iter", "last", "mean", "std", "sem", "rsem", "mem", "solver"}
    assert "Simulation runs now" in r.stdout


def test_cpic_cli_umax_abort_is_clean():
    """A velocity above umax fails with a clear error and a non-zero exit
    (the reference aborts in check_velocity, src/mover.c:123-136)."""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(ROOT, "cpic_amd", "conf", "two-stream-64.conf")
        text = open(src).read().replace("drift_velocity = [0.1, 0.0]", "drift_velocity = [200.0, 0.0]")
        path = os.path.join(d, "c.conf")
        with open(path, "w") as f:
            f.write(text)
        r = subprocess.run([BIN, "-q", path], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "max velocity exceeded" in r.stderr


def test_cpic_cli_bad_conf():
    r = subprocess.run([BIN, os.path.join(CONF, "mpi.conf")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and 'pblock_nmax' in r.stderr


def test_field_output_reference_format():
    """output.path: bin/<iter>/{rho,phi,E_X,E_Y}.bin (padded FP64 matrices,
    aligned) and xdmf/fields-iter<iter>.xdmf whose HyperSlabs select the
    interior (src/output.c:400-635), every output.period.field iterations;
    the interiors equal the oracle's fields of that iteration."""
    import xml.etree.ElementTree as ET
    import pyoracle as po
    from test_host import generate
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "out")
        src = os.path.join(ROOT, "cpic_amd", "conf", "two-stream-64.conf")
        text = open(src).read().replace("cycles = 200", "cycles = 5")
        text += (f'\noutput = {{\n\tpath = "{out}"\n\tslices = 4\n\talignment = 4096\n'
                 f'\tperiod = {{ field = 2 }}\n}}\n')
        path = os.path.join(d, "c.conf")
        with open(path, "w") as f:
            f.write(text)
        r = subprocess.run([BIN, "-q", path], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        o = po.Oracle(64, 64, (64.0, 64.0), 0.1, [-1.0, -1.0], [2.5, 2.5], nchunks=4)
        for s in range(2):
            ret, x, y, ux, uy = generate(path, s, 5120)
            assert ret == 0
            o.add(s, np.arange(5120, dtype=np.int64), x, y, ux, uy)
        o.pre_step()
        names = {"rho": "rho", "phi": "phi", "E_X": "Ex", "E_Y": "Ey"}
        shapes = {"rho": (65, 66, 0), "phi": (67, 66, 1), "E_X": (65, 64, 0), "E_Y": (65, 64, 0)}
        for it in range(5):
            o.stage("field_E")
            xd = os.path.join(out, "xdmf", f"fields-iter{it}.xdmf")
            if it % 2:
                assert not os.path.exists(xd)
            else:
                root = ET.parse(xd).getroot()
                grid = root.find("Domain/Grid")
                assert grid.find("Topology").get("NumberOfElements") == "1 64 64"
                geo = grid.findall("Geometry/DataItem")
                assert geo[1].text.split() == ["0.000000", "1.000000", "1.000000"]
                attrs = grid.findall("Attribute")
                assert [a.get("Name") for a in attrs] == ["phi", "rho", "E_X", "E_Y"]
                for a in attrs:
                    name = a.get("Name")
                    rows, cols, dy = shapes[name]
                    slab = a.find("DataItem")
                    assert slab.get("Dimensions") == "1 64 64"
                    sel, data = slab.findall("DataItem")
                    assert sel.text.split()[:3] == ["0", str(dy), "0"]
                    assert data.get("Dimensions") == f"1 {rows} {cols}"
                    assert data.text.strip() == f"../bin/{it}/{name}.bin"
                    fn = os.path.join(out, "bin", str(it), f"{name}.bin")
                    nbytes = rows * cols * 8
                    assert os.path.getsize(fn) == (nbytes + 4095) // 4096 * 4096
                    m = np.fromfile(fn, dtype=np.float64, count=rows * cols).reshape(rows, cols)
                    interior = m[dy:dy + 64, :64]
                    ref = o.field(names[name])
                    assert rel_l2(interior, ref) < 1e-12, (it, name)
                    if name == "phi":
                        assert np.array_equal(m[0, :64], interior[63])
                        assert np.array_equal(m[65, :64], interior[0])
                        assert np.array_equal(m[66, :64], interior[1])
                    elif name != "rho":
                        assert np.array_equal(m[64], interior[0])
            o.stage("plasma_E")
            o.stage("plasma_r")
            o.stage("field_rho")
            o.iter = it + 1
