"""Python access to the gfx950 PIC library (lib/libcpic_hip.so) through its C ABI.

This is plumbing for tests and benchmarks: every call goes straight to the
C entry points declared in include/cpic_hip.h.  There is no fallback: if the
HIP library is missing, importing the handle raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "lib", "libcpic_hip.so")

MAX_SPECIES = 8
F_IDS, F_PARTICLE_E, F_PROFILE, F_NO_DIAG, F_NO_GRAPH = 0x1, 0x2, 0x4, 0x8, 0x10
RHO, PHI, EX, EY = 0, 1, 2, 3
FIELDS = {"rho": RHO, "phi": PHI, "Ex": EX, "Ey": EY}
ERRORS = {0: "CPIC_OK", -1: "CPIC_EINVAL", -2: "CPIC_ENOMEM", -3: "CPIC_EUMAX",
          -4: "CPIC_ESTATE", -5: "CPIC_EHIP", -6: "CPIC_ENODEV", -7: "CPIC_ENAN",
          -8: "CPIC_ESIZE"}
# Bytes of the static per-GPU workspace of cpic_solve_xy_noalloc.
NOALLOC_WORKSPACE_BYTES = 34078720

# Every symbol include/cpic_hip.h declares.
EXPORTS = (
    "cpic_dev_strerror", "cpic_dev_create", "cpic_dev_destroy", "cpic_dev_upload_particles",
    "cpic_dev_download_particles", "cpic_dev_count", "cpic_dev_pre_step", "cpic_dev_field_E",
    "cpic_dev_plasma_E", "cpic_dev_plasma_r", "cpic_dev_field_rho", "cpic_dev_step",
    "cpic_dev_run", "cpic_dev_sync", "cpic_dev_solve_xy", "cpic_solve_xy_noalloc",
    "cpic_solver_rho_size",
    "cpic_dev_download_field", "cpic_dev_upload_field", "cpic_dev_energy", "cpic_dev_migration",
    "cpic_dev_cell_counts", "cpic_dev_rebuild", "cpic_dev_stats", "cpic_dev_kernel_times",
    "cpic_dev_reset_kernel_times", "cpic_dev_set_profiling", "cpic_batch_create", "cpic_batch_size",
    "cpic_batch_upload_particles", "cpic_batch_download_particles", "cpic_batch_count",
    "cpic_batch_download_field", "cpic_batch_upload_field", "cpic_batch_energy",
    "cpic_batch_cell_counts", "cpic_adj_create", "cpic_adj_destroy", "cpic_adj_forward",
    "cpic_adj_backward", "cpic_adj_state", "cpic_adj_device_ms", "cpic_adj_set_field_history",
)


class Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64),
        ("L", C.c_double * 2), ("dt", C.c_double), ("e0", C.c_double),
        ("B", C.c_double * 3),
        ("nspecies", C.c_int32), ("precision", C.c_int32),
        ("q", C.c_double * MAX_SPECIES), ("m", C.c_double * MAX_SPECIES),
        ("nchunks", C.c_int64),
        ("tile", C.c_int32 * 2), ("flags", C.c_int32),
        ("slack", C.c_double),
        ("device", C.c_int32), ("ppt", C.c_int32),
    ]


class Particles(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("id", C.POINTER(C.c_int64)),
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("ux", C.POINTER(C.c_double)), ("uy", C.POINTER(C.c_double)),
        ("Ex", C.POINTER(C.c_double)), ("Ey", C.POINTER(C.c_double)),
        ("cell", C.POINTER(C.c_int64)),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("nparticles", C.c_int64), ("capacity", C.c_int64), ("nbuckets", C.c_int64),
        ("leavers", C.c_int64), ("fallbacks", C.c_int64), ("misplaced", C.c_int64),
        ("rebuilds", C.c_int64), ("tile", C.c_int32 * 2), ("precision", C.c_int32),
    ]


_lib = None


def lib():
    """Loads libcpic_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"HIP library missing: {LIBPATH} (run `make -C cpic_amd`)")
        L = C.CDLL(LIBPATH)
        P = C.POINTER
        d, i64, i32, vp = P(C.c_double), P(C.c_int64), P(C.c_int32), C.c_void_p
        L.cpic_dev_strerror.restype = C.c_char_p
        L.cpic_dev_strerror.argtypes = [C.c_int]
        L.cpic_dev_create.argtypes = [P(Desc), P(vp)]
        L.cpic_dev_destroy.argtypes = [vp]
        L.cpic_dev_upload_particles.argtypes = [vp, C.c_int32, C.c_int64, i64, d, d, d, d]
        L.cpic_dev_download_particles.argtypes = [vp, C.c_int32, P(Particles)]
        L.cpic_dev_count.argtypes = [vp, C.c_int32, i64]
        L.cpic_dev_pre_step.argtypes = [vp]
        for f in ("field_E", "plasma_E", "plasma_r", "field_rho", "step"):
            getattr(L, "cpic_dev_" + f).argtypes = [vp, C.c_int64]
        L.cpic_dev_run.argtypes = [vp, C.c_int64, C.c_int64]
        L.cpic_dev_sync.argtypes = [vp]
        L.cpic_dev_solve_xy.argtypes = [vp]
        L.cpic_solve_xy_noalloc.argtypes = [C.c_int64, C.c_int64, vp, vp, vp]
        L.cpic_solver_rho_size.restype = C.c_int64
        L.cpic_solver_rho_size.argtypes = [C.c_int64, C.c_int64, i64, i64]
        L.cpic_dev_download_field.argtypes = [vp, C.c_int32, d]
        L.cpic_dev_upload_field.argtypes = [vp, C.c_int32, d]
        L.cpic_dev_energy.argtypes = [vp, C.c_int64, d, d, d, d, i64]
        L.cpic_dev_migration.argtypes = [vp, i64, i64, i32, i32, i32]
        L.cpic_dev_cell_counts.argtypes = [vp, C.c_int32, i64]
        L.cpic_dev_rebuild.argtypes = [vp]
        L.cpic_dev_stats.argtypes = [vp, P(Stats)]
        L.cpic_dev_kernel_times.argtypes = [vp, i32, P(C.c_char_p), i64, d]
        L.cpic_dev_reset_kernel_times.argtypes = [vp]
        L.cpic_dev_set_profiling.argtypes = [vp, C.c_int32]
        L.cpic_batch_create.argtypes = [P(Desc), C.c_int32, P(vp)]
        L.cpic_batch_size.argtypes = [vp, i32]
        L.cpic_batch_upload_particles.argtypes = [vp, C.c_int32, C.c_int32, C.c_int64, i64,
                                                  d, d, d, d]
        L.cpic_batch_download_particles.argtypes = [vp, C.c_int32, C.c_int32, P(Particles)]
        L.cpic_batch_count.argtypes = [vp, C.c_int32, C.c_int32, i64]
        L.cpic_batch_download_field.argtypes = [vp, C.c_int32, C.c_int32, d]
        L.cpic_batch_upload_field.argtypes = [vp, C.c_int32, C.c_int32, d]
        L.cpic_batch_energy.argtypes = [vp, C.c_int32, C.c_int64, d, d, d, d, i64]
        L.cpic_batch_cell_counts.argtypes = [vp, C.c_int32, C.c_int32, i64]
        L.cpic_adj_create.argtypes = [P(Desc), C.c_int64, i32, P(vp)]
        L.cpic_adj_destroy.argtypes = [vp]
        L.cpic_adj_forward.argtypes = [vp, d, d, d, d, C.c_int64, C.c_int32, d, d]
        L.cpic_adj_backward.argtypes = [vp, C.c_double, d, d, d, d]
        L.cpic_adj_state.argtypes = [vp, d, d, d, d]
        L.cpic_adj_device_ms.argtypes = [vp, d, d]
        L.cpic_adj_set_field_history.argtypes = [vp, C.c_int]
        _lib = L
    return _lib


class CpicError(RuntimeError):
    def __init__(self, code):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {lib().cpic_dev_strerror(code).decode()}")


def check(ret):
    if ret != 0:
        raise CpicError(ret)
    return ret


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64)) if a is not None else None


def make_desc(nx, ny, L, dt, q, m, e0=1.0, B=(0.0, 0.0, 0.0), nchunks=1, precision=64,
              tile=(0, 0), flags=0, slack=0.0, device=0, ppt=0):
    """A cpic_dev_desc_t from the reference's configuration values."""
    desc = Desc()
    desc.nx, desc.ny = int(nx), int(ny)
    desc.L[0], desc.L[1] = float(L[0]), float(L[1])
    desc.dt, desc.e0 = float(dt), float(e0)
    for i in range(3):
        desc.B[i] = float(B[i])
    desc.nspecies = len(q)
    desc.precision = int(precision)
    for i, (qq, mm) in enumerate(zip(q, m)):
        desc.q[i], desc.m[i] = float(qq), float(mm)
    desc.nchunks = int(nchunks)
    desc.tile[0], desc.tile[1] = int(tile[0]), int(tile[1])
    desc.flags = int(flags)
    desc.slack = float(slack)
    desc.device = int(device)
    desc.ppt = int(ppt)
    return desc


class Device:
    """One simulation on one GPU (a cpic_dev_t holding one plasma).

    Methods that address one plasma take `plasma` (0 for a single run);
    stage calls act on every plasma of the handle."""

    def __init__(self, nx, ny, L, dt, q, m, e0=1.0, B=(0.0, 0.0, 0.0), nchunks=1,
                 precision=64, tile=(0, 0), flags=0, slack=0.0, device=0, ppt=0):
        desc = make_desc(nx, ny, L, dt, q, m, e0, B, nchunks, precision, tile, flags, slack,
                         device, ppt)
        h = C.c_void_p()
        check(lib().cpic_dev_create(C.byref(desc), C.byref(h)))
        self._init([desc], h)

    def _init(self, descs, h):
        self._h = h
        self.descs = list(descs)
        self.desc = descs[0]
        self.P = len(descs)
        self.nx, self.ny = int(descs[0].nx), int(descs[0].ny)
        self.nspecies = int(descs[0].nspecies)
        self.flags = int(descs[0].flags)

    def close(self):
        if getattr(self, "_h", None):
            lib().cpic_dev_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def upload(self, species, ids, x, y, ux, uy, plasma=0):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        x, y, ux, uy = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, ux, uy))
        check(lib().cpic_batch_upload_particles(self._h, plasma, species, len(x), _ip(ids), _dp(x),
                                                _dp(y), _dp(ux), _dp(uy)))

    def count(self, species, plasma=0):
        n = C.c_int64()
        check(lib().cpic_batch_count(self._h, plasma, species, C.byref(n)))
        return n.value

    def particles(self, species, sort=True, plasma=0):
        n = self.count(species, plasma)
        out = {k: np.empty(n) for k in ("x", "y", "ux", "uy")}
        out["cell"] = np.empty(n, np.int64)
        if self.flags & F_IDS:
            out["id"] = np.empty(n, np.int64)
        if self.flags & F_PARTICLE_E:
            out["Ex"] = np.empty(n)
            out["Ey"] = np.empty(n)
        p = Particles()
        p.n = n
        for k in ("x", "y", "ux", "uy", "Ex", "Ey"):
            if k in out:
                setattr(p, k, _dp(out[k]))
        p.cell = _ip(out["cell"])
        if "id" in out:
            p.id = _ip(out["id"])
        check(lib().cpic_batch_download_particles(self._h, plasma, species, C.byref(p)))
        assert p.n == n, (p.n, n)
        if sort and "id" in out:
            order = np.argsort(out["id"], kind="stable")
            out = {k: v[order] for k, v in out.items()}
        return out

    def pre_step(self):
        check(lib().cpic_dev_pre_step(self._h))

    def stage(self, name, it):
        check(getattr(lib(), "cpic_dev_" + name)(self._h, int(it)))

    def step(self, it):
        check(lib().cpic_dev_step(self._h, int(it)))

    def run(self, it0, n):
        check(lib().cpic_dev_run(self._h, int(it0), int(n)))

    def sync(self):
        check(lib().cpic_dev_sync(self._h))

    def solve_xy(self):
        check(lib().cpic_dev_solve_xy(self._h))

    def shape(self, plasma=0):
        d = self.descs[plasma]
        return int(d.ny), int(d.nx)

    def field(self, which, plasma=0):
        out = np.empty(self.shape(plasma))
        check(lib().cpic_batch_download_field(self._h, plasma, FIELDS[which], _dp(out)))
        return out

    def set_field(self, which, arr, plasma=0):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(plasma))
        check(lib().cpic_batch_upload_field(self._h, plasma, FIELDS[which], _dp(arr)))

    def energy(self, maxn=1 << 16, plasma=0):
        n = C.c_int64()
        f, k, px, py = (np.empty(maxn) for _ in range(4))
        check(lib().cpic_batch_energy(self._h, plasma, maxn, _dp(f), _dp(k), _dp(px), _dp(py),
                                      C.byref(n)))
        m = min(n.value, maxn)
        return {"field": f[:m], "kinetic": k[:m], "px": px[:m], "py": py[:m]}

    def migration(self):
        n = C.c_int64(0)
        ret = lib().cpic_dev_migration(self._h, C.byref(n), None, None, None, None)
        if ret not in (0, -1):
            check(ret)
        k = n.value
        ids = np.empty(k, np.int64)
        sp, src, dst = (np.empty(k, np.int32) for _ in range(3))
        i32 = C.POINTER(C.c_int32)
        check(lib().cpic_dev_migration(self._h, C.byref(n), _ip(ids) if self.flags & F_IDS else None,
                                       sp.ctypes.data_as(i32), src.ctypes.data_as(i32),
                                       dst.ctypes.data_as(i32)))
        return {"id": ids, "species": sp, "src": src, "dst": dst}

    def cell_counts(self, species, plasma=0):
        ny, nx = self.shape(plasma)
        out = np.empty(nx * ny, np.int64)
        check(lib().cpic_batch_cell_counts(self._h, plasma, species, _ip(out)))
        return out.reshape(ny, nx)

    def rebuild(self):
        check(lib().cpic_dev_rebuild(self._h))

    def stats(self):
        st = Stats()
        check(lib().cpic_dev_stats(self._h, C.byref(st)))
        return {k: (list(getattr(st, k)) if k == "tile" else getattr(st, k))
                for k, _ in Stats._fields_}

    def kernel_times(self):
        n = C.c_int32(32)
        names = (C.c_char_p * 32)()
        launches = np.zeros(32, np.int64)
        total = np.zeros(32)
        check(lib().cpic_dev_kernel_times(self._h, C.byref(n), names, _ip(launches), _dp(total)))
        return {names[i].decode(): (int(launches[i]), float(total[i])) for i in range(n.value)}

    def reset_kernel_times(self):
        check(lib().cpic_dev_reset_kernel_times(self._h))

    def set_profiling(self, on):
        check(lib().cpic_dev_set_profiling(self._h, 1 if on else 0))


class Batch(Device):
    """P independent plasmas in one handle (cpic_batch_create): `descs` is a
    list of Desc (make_desc) sharing precision, flags, device and slack."""

    def __init__(self, descs):
        descs = list(descs)
        arr = (Desc * len(descs))(*descs)
        h = C.c_void_p()
        check(lib().cpic_batch_create(arr, len(descs), C.byref(h)))
        self._init(descs, h)


def solver_rho_size(nx, ny):
    cnx, cny = C.c_int64(), C.c_int64()
    n = lib().cpic_solver_rho_size(nx, ny, C.byref(cnx), C.byref(cny))
    return n, cnx.value, cny.value
