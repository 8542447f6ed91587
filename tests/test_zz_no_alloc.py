"""aleph_hip_sample_with_uniforms_noalloc: the allocation-free sampler, through
the C ABI, returns exactly what aleph_hip_sample_with_uniforms returns for the
same state and uniforms, on two calls in a row (the second with a smaller
state and fewer shots, reusing the device array), and rejects sizes beyond
its limits."""

import ctypes as C

import numpy as np
import pytest

from aleph_amd import _lib as L
from aleph_amd import workloads as W
from aleph_amd.backend import HipSvBackend

pytestmark = pytest.mark.gpu


def noalloc_sample(b, st, u):
    out = np.zeros(u.size, dtype=np.uint64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    code = L.load().aleph_hip_sample_with_uniforms_noalloc(b._ctx, st.handle, u.ctypes.data_as(L.PF64), u.size,
                                                          out.ctypes.data_as(L.PU64))
    return code, out


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_sampler_equals_allocating_sampler(precision):
    b = HipSvBackend(seed=5, precision=precision)
    for n, shots in ((22, 5000), (14, 700)):
        st = b.allocate(n)
        b.run_program(st, W.brickwall(n, 6, seed=n))
        u = np.random.default_rng(n).random(shots)
        want = b.sample(st, shots, uniforms=u)
        code, got = noalloc_sample(b, st, u)
        print(f"noalloc sampler precision={precision} n={n} shots={shots}: status {code}", flush=True)
        assert code == L.OK
        assert np.array_equal(got, want)
        st.free()


def test_noalloc_sampler_limits():
    b = HipSvBackend(seed=5)
    st = b.allocate(10)
    code, _ = noalloc_sample(b, st, np.full((1 << 16) + 1, 0.5))
    assert code == L.ERR_INVALID_ARGUMENT
    code, out = noalloc_sample(b, st, np.zeros(0))
    assert code == L.OK and out.size == 0
    code, out = noalloc_sample(b, st, np.array([0.25, 0.75]))
    assert code == L.OK and np.array_equal(out, [0, 0])
    st.free()
