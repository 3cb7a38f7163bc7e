"""Allocation-free MFT solve (cpic_solve_xy_noalloc) against the handle's
solve (cpic_dev_solve_xy) on the same rho.

The entry point keeps its spectrum and tables in one static workspace per GPU
that every call reuses, so the test makes two calls in a row, the second on a
smaller grid of another shape than the first, and checks each against the
handle's phi (FP64; bar 1e-12 relative L2, the oracle bar of the FP64 path).
The device buffers for rho and phi are the caller's: the test allocates them
with the HIP runtime the library is linked against."""
import ctypes as C

import numpy as np
import pytest

import cpic_amd as ca
from common import rel_l2

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    ca.lib()   # loads the HIP runtime the library was built against
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipSetDevice.argtypes = [C.c_int]
    return h


def ok(ret):
    assert ret == 0, ret


def handle_phi(rho):
    ny, nx = rho.shape
    d = ca.Device(nx, ny, (float(nx), float(ny)), 0.1, [-1.0], [1.0])
    try:
        d.set_field("rho", rho)
        d.solve_xy()
        return d.field("phi")
    finally:
        d.close()


class DeviceGrid:
    """rho and phi buffers of one grid in device memory (the caller's)."""

    def __init__(self, h, rho):
        self.h, self.shape = h, rho.shape
        self.nbytes = rho.nbytes
        self.rho, self.phi = C.c_void_p(), C.c_void_p()
        ok(h.hipMalloc(C.byref(self.rho), self.nbytes))
        ok(h.hipMalloc(C.byref(self.phi), self.nbytes))
        ok(h.hipMemcpy(self.rho, rho.ctypes.data, self.nbytes, H2D))

    def read(self, which):
        out = np.empty(self.shape)
        ok(self.h.hipMemcpy(out.ctypes.data, getattr(self, which), self.nbytes, D2H))
        return out

    def free(self):
        self.h.hipFree(self.rho)
        self.h.hipFree(self.phi)


def test_noalloc_solve_matches_handle_solve_twice_then_smaller():
    h = hip_runtime()
    ok(h.hipSetDevice(0))
    lib = ca.lib()
    rng = np.random.default_rng(7)
    big = rng.normal(size=(256, 512))      # ny x nx
    small = rng.normal(size=(128, 64))     # fewer nodes, other aspect
    gb, gs = DeviceGrid(h, big), DeviceGrid(h, small)
    stream = C.c_void_p()
    ok(h.hipStreamCreate(C.byref(stream)))
    try:
        print("noalloc solve 512x256 on the null stream, then 64x128 on a stream", flush=True)
        ok(lib.cpic_solve_xy_noalloc(512, 256, gb.rho, gb.phi, None))
        ok(h.hipStreamSynchronize(None))
        phi_big = gb.read("phi")
        ok(lib.cpic_solve_xy_noalloc(64, 128, gs.rho, gs.phi, stream))
        ok(h.hipStreamSynchronize(stream))
        phi_small = gs.read("phi")
        # the inputs are read, not written
        assert np.array_equal(gb.read("rho"), big)
        assert np.array_equal(gs.read("rho"), small)
    finally:
        h.hipStreamDestroy(stream)
        gb.free()
        gs.free()
    for rho, phi in ((big, phi_big), (small, phi_small)):
        ref = handle_phi(rho)
        err = rel_l2(phi, ref)
        assert err <= 1e-12, (rho.shape, err)
        # zero mode removed (G(0,0) = 0), as in the handle's solve
        assert abs(phi.mean()) <= 1e-12 * np.abs(phi).max()


def test_noalloc_rejects_grids_beyond_its_workspace():
    h = hip_runtime()
    ok(h.hipSetDevice(0))
    lib = ca.lib()
    g = DeviceGrid(h, np.zeros((2, 2)))
    try:
        # 4096 x 4096 needs a 134 MB spectrum: refused before any launch
        assert lib.cpic_solve_xy_noalloc(4096, 4096, g.rho, g.phi, None) == -8
        assert lib.cpic_solve_xy_noalloc(48, 64, g.rho, g.phi, None) == -1
        assert lib.cpic_solve_xy_noalloc(64, 64, None, g.phi, None) == -1
        # the largest grid it takes, and a call after the refusals still works
        rng = np.random.default_rng(3)
        rho = rng.normal(size=(2048, 2048))
        gl = DeviceGrid(h, rho)
        try:
            ok(lib.cpic_solve_xy_noalloc(2048, 2048, gl.rho, gl.phi, None))
            ok(h.hipStreamSynchronize(None))
            phi = gl.read("phi")
        finally:
            gl.free()
        assert rel_l2(phi, handle_phi(rho)) <= 1e-12
    finally:
        g.free()


def test_noalloc_solve_replayed_from_a_captured_graph():
    """The allocation-free solve captured into a HIP graph (its callers'
    case) and replayed four times, each on a new rho written between the
    replays, matches the handle's solve of that rho every time."""
    h = hip_runtime()
    ok(h.hipSetDevice(0))
    h.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
    h.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    h.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_size_t]
    h.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
    h.hipGraphExecDestroy.argtypes = [C.c_void_p]
    h.hipGraphDestroy.argtypes = [C.c_void_p]
    lib = ca.lib()
    rng = np.random.default_rng(11)
    nx, ny = 256, 128
    g = DeviceGrid(h, np.zeros((ny, nx)))
    stream, graph, exe = C.c_void_p(), C.c_void_p(), C.c_void_p()
    ok(h.hipStreamCreate(C.byref(stream)))
    try:
        print("capturing cpic_solve_xy_noalloc 256x128", flush=True)
        ok(h.hipStreamBeginCapture(stream, 0))          # hipStreamCaptureModeGlobal
        ret = lib.cpic_solve_xy_noalloc(nx, ny, g.rho, g.phi, stream)
        ok(h.hipStreamEndCapture(stream, C.byref(graph)))
        ok(ret)
        ok(h.hipGraphInstantiate(C.byref(exe), graph, None, None, 0))
        for rep in range(4):
            rho = rng.normal(size=(ny, nx))
            ok(h.hipMemcpy(g.rho, rho.ctypes.data, rho.nbytes, H2D))
            ok(h.hipGraphLaunch(exe, stream))
            ok(h.hipStreamSynchronize(stream))
            phi = g.read("phi")
            err = rel_l2(phi, handle_phi(rho))
            assert err <= 1e-12, (rep, err)
    finally:
        if exe.value:
            h.hipGraphExecDestroy(exe)
        if graph.value:
            h.hipGraphDestroy(graph)
        h.hipStreamDestroy(stream)
        g.free()
