"""shp_mlp_forward_noalloc (the MLP forward that allocates nothing, its
staging in the library's static device scratch) gives the bits of
shp_mlp_forward on the same inputs, on two calls in a row (the second
smaller), through the C ABI; it refuses sizes beyond the scratch."""
import ctypes as C
import os

import numpy as np
import pytest

from shapes_amd import engine

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hip():
    """The HIP runtime the library links (device buffers for the test)."""
    engine.load_library()
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return L


def _check(rc):
    assert rc == 0, f"HIP error {rc}"


def test_mlp_forward_noalloc_matches_forward():
    hip = _hip()
    params = np.load(os.path.join(ROOT, "shapes_amd", "warmstart_mlp.npy"))
    m = engine.Mlp(params)
    rng = np.random.default_rng(7)
    for n in (5000, 777):
        print(f"no-alloc forward: {n} rows", flush=True)
        feats = rng.normal(size=(n, 14)).astype(np.float32)
        want = m.forward(feats)
        d_in, d_out = C.c_void_p(), C.c_void_p()
        _check(hip.hipMalloc(C.byref(d_in), feats.nbytes))
        _check(hip.hipMalloc(C.byref(d_out), n * 2 * 4))
        _check(hip.hipMemcpy(d_in, feats.ctypes.data, feats.nbytes, 1))  # host to device
        _check(hip.hipDeviceSynchronize())
        free0, free1, total = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _check(hip.hipMemGetInfo(C.byref(free0), C.byref(total)))
        m.forward_noalloc(d_in.value, n, d_out.value)
        _check(hip.hipDeviceSynchronize())
        _check(hip.hipMemGetInfo(C.byref(free1), C.byref(total)))
        assert free1.value == free0.value  # nothing allocated by the call
        got = np.empty((n, 2), dtype=np.float32)
        _check(hip.hipMemcpy(got.ctypes.data, d_out, got.nbytes, 2))  # device to host
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        if n == 777:
            with pytest.raises(engine.ShapesError) as ei:
                m.forward_noalloc(d_in.value, 65537, d_out.value)
            assert ei.value.status == 1 and "65536" in str(ei.value)
        _check(hip.hipFree(d_in))
        _check(hip.hipFree(d_out))
    m.close()
