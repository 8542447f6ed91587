"""The C-ABI library loads and exports every symbol include/aleph_hip.h
declares; host-side logic that needs no device."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd.circuit import Circuit, Gate
from aleph_amd import workloads as W

HEADER = Path(__file__).resolve().parent.parent / "include" / "aleph_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(aleph_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 30
    for core in ("aleph_hip_ctx_create", "aleph_hip_state_alloc", "aleph_hip_apply_gate", "aleph_hip_plan_create",
                 "aleph_hip_sample", "aleph_hip_expectation", "aleph_hip_probabilities", "aleph_hip_measure"):
        assert core in syms


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(L.SIGNATURES)


def test_struct_layouts():
    assert C.sizeof(L.Gate) == 4 * 4 + 8 + 8 + 24 + 8
    assert C.sizeof(L.PlanOpts) == 8 * 4
    assert C.sizeof(L.Error) == 4 * 5 + 4 + 8 + 48 + 256
    assert C.sizeof(L.BatchOpts) == 8 * 4
    assert C.sizeof(L.BatchStats) == 8 * 3 + 14 * 4
    assert C.sizeof(L.BatchTune) == 4 * 4 + 8
    assert C.sizeof(L.PlanStats) == 8 + 6 * 4 + 8 + 4 * 4 + 8 + 2 * 4 + 8 + 2 * 4
    assert C.sizeof(L.Channel) == 6 * 4 + 3 * 8


def test_no_device_paths_fail_cleanly():
    """Without a device, contexts are refused with NoDevice (aleph-cuda Error::NoDevice)."""
    if L.device_count() > 0:
        pytest.skip("a device is present")
    lib = L.load()
    ctx = C.c_void_p()
    assert lib.aleph_hip_ctx_create(0, 1, C.byref(ctx)) == L.ERR_NO_DEVICE
    assert not ctx.value
    from aleph_amd.backend import HipSvBackend, NoDevice
    with pytest.raises(NoDevice):
        HipSvBackend()


def test_program_encoding_roundtrip():
    c = W.brickwall(5, 3, seed=1)
    words, reals = c.encode()
    # 5 U3 per layer (3 reals each) and CZ pairs per layer
    assert reals.size == 3 * 5 * 3
    n_cz = 2 + 2 + 2
    assert words.size == 15 * (4 + 1) + n_cz * (4 + 2)
    c2 = Circuit(3).gate(Gate.unitary_kq(np.eye(8)), [0, 1, 2]).cp(0.5, 2, 0)
    w, r = c2.encode()
    assert list(w[:7]) == [25, 3, 0, 3, 0, 1, 2]
    assert r.size == 2 * 64 + 1


@pytest.mark.gpu
def test_state_may_outlive_its_context():
    """Freeing a state after its context is destroyed must not touch the
    context (teardown order at interpreter exit is arbitrary)."""
    from aleph_amd.backend import HipSvBackend
    if L.device_count() == 0:
        pytest.skip("no HIP device")
    b = HipSvBackend(seed=1)
    st = b.allocate(6)
    b.close()
    st.free()


def test_rust_binding_declares_every_entry_point():
    """crates/aleph-hip/src/ffi.rs declares every function of the header with
    the same parameter count (the crate cannot be compiled in this image, so
    this is its check against the C ABI)."""
    import re
    from pathlib import Path
    root = Path(__file__).resolve().parent.parent
    header = re.sub(r"/\*.*?\*/", "", (root / "include" / "aleph_hip.h").read_text(), flags=re.S)
    protos = dict((m.group(1), m.group(2)) for m in re.finditer(
        r"(?:aleph_hip_status|void)\s+(aleph_hip_\w+)\s*\(([^;]*)\)\s*;", header))
    ffi = (root / "crates" / "aleph-hip" / "src" / "ffi.rs").read_text()
    decls = dict((m.group(1), m.group(2)) for m in re.finditer(r"pub fn (aleph_hip_\w+)\(([^)]*)\)", ffi))
    assert set(protos) == set(decls)
    for name, args in protos.items():
        n_c = 0 if args.strip() in ("", "void") else args.count(",") + 1
        n_rs = 0 if not decls[name].strip() else decls[name].count(",") + 1
        assert n_c == n_rs, name
