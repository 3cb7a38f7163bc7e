"""C host driver (cpic_amd/host): configuration reader and particle init
methods, checked on the CPU through ctypes (no device calls)."""
import ctypes as C
import os
import tempfile
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cpic_amd", "lib", "libcpic_host.so")
CONF = os.path.join(ROOT, "tests", "golden", "conf")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="host library not built")


class Cfg(C.Structure):
    _fields_ = [("root", C.c_void_p), ("err", C.c_char * 256), ("err_line", C.c_int),
                ("path", C.c_char * 4096)]


_h = None


def host():
    global _h
    if _h is None:
        h = C.CDLL(LIB)
        d = C.POINTER(C.c_double)
        h.cfg_read_file.argtypes = [C.POINTER(Cfg), C.c_char_p]
        h.cfg_read_string.argtypes = [C.POINTER(Cfg), C.c_char_p]
        h.cfg_destroy.argtypes = [C.POINTER(Cfg)]
        h.cfg_lookup.restype = C.c_void_p
        h.cfg_lookup.argtypes = [C.POINTER(Cfg), C.c_char_p]
        h.cfg_length.argtypes = [C.c_void_p]
        h.cfg_lookup_int64.argtypes = [C.POINTER(Cfg), C.c_char_p, C.POINTER(C.c_int64)]
        h.cfg_lookup_float.argtypes = [C.POINTER(Cfg), C.c_char_p, d]
        h.cfg_lookup_bool.argtypes = [C.POINTER(Cfg), C.c_char_p, C.POINTER(C.c_int)]
        h.cfg_lookup_string.argtypes = [C.POINTER(Cfg), C.c_char_p, C.POINTER(C.c_char_p)]
        h.cfg_lookup_array_float.argtypes = [C.POINTER(Cfg), C.c_char_p, d, C.c_int64]
        h.host_generate.argtypes = [C.c_char_p, C.c_int, C.c_int64, d, d, d, d]
        _h = h
    return _h


class Conf:
    def __init__(self, text=None, path=None):
        self.c = Cfg()
        if path:
            self.ret = host().cfg_read_file(C.byref(self.c), path.encode())
        else:
            self.ret = host().cfg_read_string(C.byref(self.c), text.encode())

    def __del__(self):
        if self.ret == 0:
            host().cfg_destroy(C.byref(self.c))

    @property
    def error(self):
        return self.c.err.decode(), self.c.err_line

    def i(self, path):
        v = C.c_int64()
        assert host().cfg_lookup_int64(C.byref(self.c), path.encode(), C.byref(v)) == 0, path
        return v.value

    def f(self, path):
        v = C.c_double()
        assert host().cfg_lookup_float(C.byref(self.c), path.encode(), C.byref(v)) == 0, path
        return v.value

    def s(self, path):
        v = C.c_char_p()
        assert host().cfg_lookup_string(C.byref(self.c), path.encode(), C.byref(v)) == 0, path
        return v.value.decode()

    def b(self, path):
        v = C.c_int()
        assert host().cfg_lookup_bool(C.byref(self.c), path.encode(), C.byref(v)) == 0, path
        return v.value

    def arr(self, path, n):
        out = (C.c_double * n)()
        assert host().cfg_lookup_array_float(C.byref(self.c), path.encode(), out, n) == 0, path
        return list(out)

    def has(self, path):
        return bool(host().cfg_lookup(C.byref(self.c), path.encode()))

    def length(self, path):
        return host().cfg_length(host().cfg_lookup(C.byref(self.c), path.encode()))


def generate(path, sp, n):
    x, y, ux, uy = (np.empty(n) for _ in range(4))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    ret = host().host_generate(path.encode(), sp, n, dp(x), dp(y), dp(ux), dp(uy))
    return ret, x, y, ux, uy


# ------------------------------------------------------------- reader ---

def test_reference_confs_parse():
    c = Conf(path=os.path.join(CONF, "harmonic.conf"))
    assert c.ret == 0, c.error
    assert c.length("species") == 2
    assert c.s("species.[0].name") == "electron"
    assert c.f("species.[0].mass") == 1.0e-2
    assert c.s("species.[1].init_method") == "position delta"
    assert c.arr("species.[1].position_init", 2) == [7.0, 4.0]
    assert c.i("grid.points.[0]") == 1024
    assert c.f("simulation.time_step") == 5.0e-3
    assert c.i("simulation.enable_fftw_threads") == 0
    assert c.f("constants.light_speed") == 2.99792458e+8
    assert c.i("output.alignment") == 4096
    assert c.arr("plot.size", 2) == [800.0, 800.0]
    cy = Conf(path=os.path.join(CONF, "cyclotron.conf"))
    assert cy.ret == 0, cy.error
    assert cy.arr("field.magnetic", 3) == [0.0, 0.0, -1.0]
    assert cy.f("constants.vacuum_permittivity") == 8.85e-12
    for name in ("constant-speed.conf", "mpi.conf", "two-streams.conf"):
        assert Conf(path=os.path.join(CONF, name)).ret == 0, name


def test_syntax_features():
    text = r'''
    # comment
    a = 1; b : 0x1F; c = -3L, d = 2.5e-3 // trailing comment
    /* block
       comment */ e = "x\ty" "z\x41";
    g = { h = true; i = [1.0, 2, 3.5] ; j = ( { k = "v" }, 4, [1, 2] ) };
    big = 1234567890123LL;
    '''
    c = Conf(text)
    assert c.ret == 0, c.error
    assert c.i("a") == 1 and c.i("b") == 31 and c.i("c") == -3 and c.f("d") == 2.5e-3
    assert c.s("e") == "x\tyzA"
    assert c.b("g.h") == 1 and c.arr("g.i", 3) == [1.0, 2.0, 3.5]
    assert c.s("g.j.[0].k") == "v" and c.i("g.j.[1]") == 4 and c.length("g.j.[2]") == 2
    assert c.i("big") == 1234567890123
    assert c.f("a") == 1.0       # integers are accepted where floats are asked
    assert not c.has("g.nope")


def test_syntax_errors_report_line():
    c = Conf("a = 1;\nb = [1, \"x\"];\n")
    assert c.ret != 0 and c.error[1] == 2
    c = Conf("a = {\n b = 2;\n")
    assert c.ret != 0 and "missing" in c.error[0]
    c = Conf("a = 1; a = 2;")
    assert c.ret != 0 and "duplicate" in c.error[0]


def test_include_relative_to_file():
    with tempfile.TemporaryDirectory() as d:
        os.mkdir(os.path.join(d, "sub"))
        with open(os.path.join(d, "sub", "inc.conf"), "w") as f:
            f.write('x = 5;\n@include "more.conf"\n')
        with open(os.path.join(d, "sub", "more.conf"), "w") as f:
            f.write('y = "deep";\n')
        with open(os.path.join(d, "sub", "main.conf"), "w") as f:
            f.write('g = { @include "inc.conf" };\nz = 1;\n')
        c = Conf(path=os.path.join(d, "sub", "main.conf"))
        assert c.ret == 0, c.error
        assert c.i("g.x") == 5 and c.s("g.y") == "deep" and c.i("z") == 1


# --------------------------------------------------------- init methods ---

def write_conf(d, species, nchunks=4, seed=138, points=(64, 64), L=(32.0, 32.0)):
    sp = ",\n".join("{ " + "; ".join(f"{k} = {v}" for k, v in s.items()) + " }" for s in species)
    text = f"""
species = ( {sp} );
field = {{ magnetic = [0.0, 0.0, 0.0] }};
grid = {{ points = [{points[0]}, {points[1]}] }};
simulation = {{
  dimensions = 2; solver = "MFT"; enable_fftw_threads = 0; plasma_chunks = {nchunks};
  random_seed = {seed}; cycles = 10; time_step = 0.1; space_length = [{L[0]}, {L[1]}];
  sampling_period = {{ energy = 0; field = 1; particle = 1 }};
  stop_SEM = 0.0; realtime_plot = 0; pblock_nmax = 1024;
}};
constants = {{ light_speed = 1.0; vacuum_permittivity = 1.0 }};
"""
    path = os.path.join(d, "t.conf")
    with open(path, "w") as f:
        f.write(text)
    return path


def test_position_delta_closed_form():
    """init_position_delta: r = fmod(r0 + dr*i, L), u = drift (src/particle.c:91-168)."""
    ret, x, y, ux, uy = generate(os.path.join(CONF, "constant-speed.conf"), 0, 1)
    assert ret == 0
    assert (x[0], y[0], ux[0], uy[0]) == (4.0, 16.0, 1.33333, 0.282828)
    with tempfile.TemporaryDirectory() as d:
        path = write_conf(d, [dict(name='"e"', particles=10, charge=-1.0, mass=1.0,
                                   drift_velocity="[0.5, -0.25]", init_method='"position delta"',
                                   position_delta="[7.5, -3.25]", position_init="[1.0, 2.0]")])
        ret, x, y, ux, uy = generate(path, 0, 10)
        assert ret == 0
        i = np.arange(10)
        np.testing.assert_array_equal(x, np.mod(1.0 + 7.5 * i, 32.0))
        np.testing.assert_array_equal(y, np.mod(2.0 - 3.25 * i, 32.0))
        assert np.all(ux == 0.5) and np.all(uy == -0.25)


def test_random_position_rng_order():
    """init_randpos draws rand() for rx, ry, ux, uy in chunk -> species ->
    index order after srand(seed) (src/sim.c:153, src/plasma.c:62-128,
    src/particle.c:24-88); restated here with glibc's own rand()."""
    libc = C.CDLL("libc.so.6")
    nc, n, seed, L, v = 4, (13, 9), 77, (32.0, 16.0), ((0.3, 0.1), (0.0, 2.0))
    with tempfile.TemporaryDirectory() as d:
        sp = [dict(name='"e%d"' % s, particles=n[s], charge=-1.0, mass=1.0,
                   drift_velocity=f"[{v[s][0]}, {v[s][1]}]", init_method='"random position"')
              for s in range(2)]
        path = write_conf(d, sp, nchunks=nc, seed=seed, L=L)
        got = [generate(path, s, n[s])[1:] for s in range(2)]
    libc.srand(seed)
    RAND_MAX = 2147483647

    def uni(a, b):
        # rand()/(RAND_MAX+1.0) * (b-a) + a with one rounding, as clang
        # contracts it; Fraction arithmetic is exact and float() rounds once
        r = libc.rand() / (RAND_MAX + 1.0)
        return float(Fraction(r) * Fraction(b - a) + Fraction(a))

    exp = [[np.empty(n[s]) for _ in range(4)] for s in range(2)]
    for c in range(nc):
        for s in range(2):
            for i in range(c, n[s], nc):
                exp[s][0][i] = uni(0.0, L[0])
                exp[s][1][i] = uni(0.0, L[1])
                exp[s][2][i] = uni(-v[s][0], v[s][0])
                exp[s][3][i] = uni(-v[s][1], v[s][1])
    for s in range(2):
        for k in range(4):
            np.testing.assert_array_equal(got[s][k], exp[s][k])


def test_two_stream_moments_and_chunk_independence():
    with tempfile.TemporaryDirectory() as d:
        spec = [dict(name='"e"', particles=200000, charge=-1.0, mass=1.0,
                     drift_velocity="[0.1, 0.0]", thermal_velocity="[0.01, 0.02]",
                     init_method='"two stream"')]
        p1 = write_conf(d, spec, nchunks=4)
        r1 = generate(p1, 0, 200000)
        p2 = write_conf(d, spec, nchunks=16)
        r2 = generate(p2, 0, 200000)
    assert r1[0] == 0 and r2[0] == 0
    for a, b in zip(r1[1:], r2[1:]):
        np.testing.assert_array_equal(a, b)
    x, y, ux, uy = r1[1:]
    assert 0 <= x.min() and x.max() < 32 and 0 <= y.min() and y.max() < 32
    assert abs(ux[0::2].mean() - 0.1) < 1e-4 and abs(ux[1::2].mean() + 0.1) < 1e-4
    assert abs(ux[0::2].std() - 0.01) < 1e-4 and abs(uy.std() - 0.02) < 2e-4
    assert abs(uy.mean()) < 2e-4
    assert abs(x.mean() - 16) < 0.1


def test_config_errors_like_reference():
    """Missing mandatory keys fail like sim_read_config's READ (src/sim.c:46-48)."""
    ret = generate(os.path.join(CONF, "mpi.conf"), 0, 2000000)[0]
    assert ret != 0        # no simulation.pblock_nmax
    ret = generate(os.path.join(CONF, "two-streams.conf"), 0, 100)[0]
    assert ret != 0        # stale keys, no simulation.dimensions
    with tempfile.TemporaryDirectory() as d:
        path = write_conf(d, [dict(name='"e"', particles=10, charge=-1.0, mass=1.0,
                                   drift_velocity="[0.0, 0.0]", init_method='"nope"')])
        assert generate(path, 0, 10)[0] != 0
        path = write_conf(d, [dict(name='"e"', particles=10, charge=-1.0, mass=1.0,
                                   drift_velocity="[0.0, 0.0]", init_method='"random position"')],
                          nchunks=3)
        assert generate(path, 0, 10)[0] != 0      # 64 % 3 != 0 (src/sim.c:124-130)
