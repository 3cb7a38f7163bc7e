"""The C-ABI library loads and exports every symbol include/cpic_hip.h declares
(no compute calls: these run without a GPU)."""
import os
import re

import pytest

import cpic_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "cpic_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cpic_\w+)\s*\(", src)))


def test_header_declares_known_entry_points():
    assert declared() == sorted(ca.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ca.lib()
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert missing == []


def test_probe_library_exports_the_bench_ceilings():
    """bench.py's streaming ceilings come from lib/libcpic_probe.so (not part
    of the ABI): plain stream, the push's pattern, the same per tile."""
    import ctypes
    probe = ctypes.CDLL(os.path.join(ROOT, "cpic_amd", "lib", "libcpic_probe.so"))
    for s in ("cpic_probe_stream", "cpic_probe_stream_groups", "cpic_probe_stream_tiles"):
        assert hasattr(probe, s), s
    ms = ctypes.c_double()
    # argument checks run before any device call
    assert probe.cpic_probe_stream_tiles(ctypes.c_int64(1 << 20), 1000, 1, ctypes.byref(ms)) == -1
    assert probe.cpic_probe_stream_tiles(ctypes.c_int64(1 << 20), 1024, 0, ctypes.byref(ms)) == -1
    assert probe.cpic_probe_stream(ctypes.c_int64(1 << 20), 3, 1, ctypes.byref(ms)) == -1


def test_strerror_and_rho_size():
    lib = ca.lib()
    for code in ca.ERRORS:
        assert lib.cpic_dev_strerror(code)
    n, cnx, cny = ca.solver_rho_size(64, 32)
    assert (cnx, cny, n) == (66, 32, 66 * 32)


def test_invalid_descriptions_rejected_before_any_device_call():
    bad = [dict(nx=48), dict(B=(0.0, 1.0, 0.0)), dict(precision=16), dict(nchunks=3),
           dict(tile=(3, 4))]
    for kw in bad:
        args = dict(nx=16, ny=16, L=(16.0, 16.0), dt=0.1, q=[-1.0], m=[1.0])
        args.update(kw)
        with pytest.raises(ca.CpicError) as ei:
            ca.Device(**args)
        assert ei.value.code == -1, kw


def test_invalid_batches_rejected_before_any_device_call():
    good = dict(nx=16, ny=16, L=(16.0, 16.0), dt=0.1, q=[-1.0], m=[1.0])
    mixed = [dict(precision=32), dict(flags=ca.F_IDS), dict(slack=0.5), dict(nx=48), dict(ppt=4),
             dict(ppt=3)]
    for kw in mixed:
        other = dict(good)
        other.update(kw)
        with pytest.raises(ca.CpicError) as ei:
            ca.Batch([ca.make_desc(**good), ca.make_desc(**other)])
        assert ei.value.code == -1, kw
    with pytest.raises(ca.CpicError) as ei:
        ca.Batch([])
    assert ei.value.code == -1


def test_adjoint_calls_reject_a_null_handle_before_any_device_call():
    """The differentiable-run entry points check their handle first."""
    lib = ca.lib()
    assert lib.cpic_adj_set_field_history(None, 0) == -1
    assert lib.cpic_adj_backward(None, 1.0, None, None, None, None) == -1
    assert lib.cpic_adj_state(None, None, None, None, None) == -1
    assert lib.cpic_adj_device_ms(None, None, None) == -1
