"""ctypes wrapper of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always to check results or as the CPU baseline,
never as the measured product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liboracle.so")

MAX_SPECIES = 8
RHO, PHI, EX, EY = 0, 1, 2, 3
FIELDS = {"rho": RHO, "phi": PHI, "Ex": EX, "Ey": EY}


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64),
        ("L", C.c_double * 2),
        ("dt", C.c_double), ("e0", C.c_double),
        ("B", C.c_double * 3),
        ("nchunks", C.c_int64),
        ("nspecies", C.c_int32),
        ("q", C.c_double * MAX_SPECIES),
        ("m", C.c_double * MAX_SPECIES),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"oracle library missing: {LIBPATH} (run make -C oracle)")
        L = C.CDLL(LIBPATH)
        P = C.POINTER
        d = P(C.c_double)
        i64 = P(C.c_int64)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P(OrcParams), P(C.c_int)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_add_particles.argtypes = [C.c_void_p, C.c_int, C.c_int64, i64, d, d, d, d]
        L.orc_set_particles.argtypes = [C.c_void_p, C.c_int, C.c_int64, i64, d, d, d, d]
        for f in ("orc_pre_step", "orc_step", "orc_stage_field_E", "orc_stage_plasma_E",
                  "orc_stage_plasma_r", "orc_stage_field_rho"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, C.c_int64]
        L.orc_comm_plasma.argtypes = [C.c_void_p, C.c_int]
        L.orc_iter.restype = C.c_int64
        L.orc_iter.argtypes = [C.c_void_p]
        L.orc_set_iter.argtypes = [C.c_void_p, C.c_int64]
        L.orc_count.restype = C.c_int64
        L.orc_count.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_particles.argtypes = [C.c_void_p, C.c_int, i64, d, d, d, d, d, d]
        L.orc_chunk_counts.argtypes = [C.c_void_p, i64]
        L.orc_chunk_ids.argtypes = [C.c_void_p, C.c_int64, C.c_int, i64]
        L.orc_get_field.argtypes = [C.c_void_p, C.c_int, d]
        L.orc_set_field.argtypes = [C.c_void_p, C.c_int, d]
        L.orc_field_energy.restype = C.c_double
        L.orc_field_energy.argtypes = [C.c_void_p]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [C.c_void_p]
        L.orc_weights.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                  d, i64]
        L.orc_boris.argtypes = [d, d, d, d, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_mft_solve.argtypes = [C.c_int64, C.c_int64, d, d]
        L.orc_fft.argtypes = [C.c_int64, d, d, C.c_int]
        L.orc_cells.argtypes = [C.c_void_p, C.c_int64, d, d, i64]
        L.orc_strerror.restype = C.c_char_p
        L.orc_strerror.argtypes = [C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class OracleError(RuntimeError):
    pass


def _check(ret):
    if ret != 0:
        raise OracleError(lib().orc_strerror(ret).decode() + f" ({ret})")


class Oracle:
    """One simulation in the oracle: the reference's sim_t, one rank."""

    def __init__(self, nx, ny, L, dt, q, m, e0=1.0, B=(0.0, 0.0, 0.0), nchunks=1):
        p = OrcParams()
        p.nx, p.ny = int(nx), int(ny)
        p.L[0], p.L[1] = float(L[0]), float(L[1])
        p.dt, p.e0 = float(dt), float(e0)
        for i in range(3):
            p.B[i] = float(B[i])
        p.nchunks = int(nchunks)
        p.nspecies = len(q)
        for i, (qq, mm) in enumerate(zip(q, m)):
            p.q[i], p.m[i] = float(qq), float(mm)
        err = C.c_int(0)
        self._h = lib().orc_create(C.byref(p), C.byref(err))
        if not self._h:
            _check(err.value or -1)
        self.nx, self.ny, self.nspecies = int(nx), int(ny), len(q)
        self.params = p

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    def add(self, species, ids, x, y, ux, uy):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        x, y, ux, uy = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, ux, uy))
        _check(lib().orc_add_particles(self._h, species, len(ids), _ip(ids), _dp(x), _dp(y),
                                       _dp(ux), _dp(uy)))

    def set_particles(self, species, ids, x, y, ux, uy):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        x, y, ux, uy = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, ux, uy))
        _check(lib().orc_set_particles(self._h, species, len(ids), _ip(ids), _dp(x), _dp(y),
                                       _dp(ux), _dp(uy)))

    def pre_step(self):
        _check(lib().orc_pre_step(self._h))

    def step(self, n=1):
        _check(lib().orc_run(self._h, int(n)))

    def stage(self, name):
        _check(getattr(lib(), "orc_stage_" + name)(self._h))

    def comm_plasma(self, global_exchange):
        _check(lib().orc_comm_plasma(self._h, int(global_exchange)))

    @property
    def iter(self):
        return lib().orc_iter(self._h)

    @iter.setter
    def iter(self, v):
        lib().orc_set_iter(self._h, int(v))

    def count(self, species):
        return lib().orc_count(self._h, species)

    def particles(self, species):
        n = self.count(species)
        ids = np.empty(n, np.int64)
        arrs = [np.empty(n) for _ in range(6)]
        _check(lib().orc_get_particles(self._h, species, _ip(ids), *[_dp(a) for a in arrs]))
        keys = ("x", "y", "ux", "uy", "Ex", "Ey")
        out = {"id": ids}
        out.update(dict(zip(keys, arrs)))
        return out

    def chunk_counts(self):
        out = np.empty(self.params.nchunks * self.nspecies, np.int64)
        _check(lib().orc_chunk_counts(self._h, _ip(out)))
        return out.reshape(self.params.nchunks, self.nspecies)

    def chunk_ids(self, chunk, species):
        n = int(self.chunk_counts()[chunk, species])
        out = np.empty(n, np.int64)
        _check(lib().orc_chunk_ids(self._h, chunk, species, _ip(out)))
        return out

    def field(self, which):
        out = np.empty((self.ny, self.nx))
        _check(lib().orc_get_field(self._h, FIELDS[which], _dp(out)))
        return out

    def set_field(self, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.ny, self.nx)
        _check(lib().orc_set_field(self._h, FIELDS[which], _dp(arr)))

    def field_energy(self):
        return lib().orc_field_energy(self._h)

    def kinetic_energy(self):
        return lib().orc_kinetic_energy(self._h)

    def cells(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty(len(x), np.int64)
        lib().orc_cells(self._h, len(x), _dp(x), _dp(y), _ip(out))
        return out


def weights(x, y, dx, dy):
    w = np.empty((2, 2))
    i0 = np.empty(2, np.int64)
    lib().orc_weights(float(x), float(y), float(dx), float(dy), _dp(w), _ip(i0))
    return w, i0


def boris(r, u, E, B, q, m, dt, set_r=True):
    r = np.array(r, dtype=np.float64)
    u = np.array(u, dtype=np.float64)
    E = np.array(E, dtype=np.float64)
    B = np.array(B, dtype=np.float64)
    lib().orc_boris(_dp(r), _dp(u), _dp(E), _dp(B), float(q), float(m), float(dt),
                    int(set_r))
    return r, u


def mft_solve(rho):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    ny, nx = rho.shape
    phi = np.empty_like(rho)
    _check(lib().orc_mft_solve(nx, ny, _dp(rho), _dp(phi)))
    return phi


def fft(z, sign=-1):
    re = np.ascontiguousarray(np.real(z), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(z), dtype=np.float64).copy()
    _check(lib().orc_fft(len(re), _dp(re), _dp(im), int(sign)))
    return re + 1j * im
