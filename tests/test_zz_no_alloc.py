"""cosenu_mlp_forward_noalloc: the surrogate network's inference on device
arrays with no allocation (its BF16 input copy and hidden activations in the
library's per-GPU scratch array), against cosenu_mlp_forward on the same
network and inputs: bit for bit, on two calls in a row, the second with fewer
rows; more rows than the scratch holds are refused with a message.

Device arrays come from the HIP runtime the library itself is linked to
(found through the process's loaded libraries), called through ctypes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip_runtime():
    """The libamdhip64 the library loaded (the same runtime instance)."""
    from cosenu_amd import lib
    lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 is not loaded"
    H = ctypes.CDLL(path)
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    H.hipDeviceSynchronize.argtypes = []
    return H


H2D, D2H = 1, 2


def test_forward_noalloc_equals_forward(gpu):
    from cosenu_amd import surrogate as S
    H = _hip_runtime()
    L = S._mlib()
    L.cosenu_mlp_forward_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_void_p]
    net = S.MLP(S.DIMS, max_batch=512)
    net.set_params(S.init_params(S.DIMS, seed=3))
    rng = np.random.default_rng(8)
    big = 3000
    dX, dY = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dX), big * S.DIMS[0] * 4) == 0
    assert H.hipMalloc(ctypes.byref(dY), big * S.DIMS[-1] * 4) == 0
    try:
        for M in (big, 37):  # the second call has fewer rows
            X = rng.standard_normal((M, S.DIMS[0])).astype(np.float32)
            ref = net.forward(X)
            assert H.hipMemcpy(dX, X.ctypes.data, X.nbytes, H2D) == 0
            assert H.hipMemset(dY, 0xFF, big * S.DIMS[-1] * 4) == 0  # NaN until written
            assert H.hipDeviceSynchronize() == 0
            rc = L.cosenu_mlp_forward_noalloc(net._h, dX, M, dY, None)
            assert rc == 0, L.cosenu_mlp_last_error()
            assert L.cosenu_mlp_sync(net._h) == 0
            out = np.empty_like(ref)
            assert H.hipMemcpy(out.ctypes.data, dY, out.nbytes, D2H) == 0
            print(f"M={M}: equal to cosenu_mlp_forward: {np.array_equal(out, ref)}", flush=True)
            assert np.array_equal(out, ref)
        # beyond the scratch array: refused before any launch
        rc = L.cosenu_mlp_forward_noalloc(net._h, dX, 1 << 20, dY, None)
        assert rc == -1
        assert b"scratch" in L.cosenu_mlp_last_error()
        rc = L.cosenu_mlp_forward_noalloc(net._h, dX, 0, dY, None)
        assert rc == -1
    finally:
        H.hipFree(dX)
        H.hipFree(dY)
        net.close()
