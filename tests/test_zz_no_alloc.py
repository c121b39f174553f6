"""hpg_mlp_loss_grad_noalloc (device buffers, static scratch, no allocation) against
hpg_mlp_loss_grad on the same samples: loss and gradient bitwise equal, on calls in a
row with shrinking sample counts (the scratch keeps the previous call's contents), and
a sample count above the scratch capacity rejected with nothing enqueued."""
import ctypes as C

import numpy as np
import pytest

from hpmp import _lib
from hpmp.learned import MLP

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip_runtime():
    """The HIP runtime libhpgpu.so itself mapped (device buffers of one runtime only)."""
    _lib.load()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 not mapped"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class _Dev:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 4)) == 0

    def put(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def test_loss_grad_noalloc_matches_loss_grad():
    hip = _hip_runtime()
    net = MLP(7)
    L = net.L
    nmax = _lib.HPG_MLP_NOALLOC_MAX_SAMPLES
    rng = np.random.default_rng(5)
    X = rng.standard_normal((nmax, 11)).astype(np.float32)
    T = rng.standard_normal((nmax, 9)).astype(np.float32)
    dX, dT = _Dev(hip, X.nbytes), _Dev(hip, T.nbytes)
    dL, dG = _Dev(hip, 4), _Dev(hip, net.nparam * 4)
    try:
        dX.put(X)
        dT.put(T)
        for n in (nmax, 20000, 777):
            print(f"noalloc loss_grad n={n}", flush=True)
            loss_ref, grad_ref = net.loss_grad(X[:n], T[:n])
            # device rows 0..n-1 are the first n samples of the uploaded arrays
            rc = L.hpg_mlp_loss_grad_noalloc(net.h, dX.p, dT.p, n, dL.p, dG.p)
            assert rc == 0
            loss = dL.get(np.zeros(1, np.float32))[0]
            grad = dG.get(np.zeros(net.nparam, np.float32))
            assert np.isfinite(grad).all()
            assert loss.tobytes() == np.float32(loss_ref).tobytes()
            assert np.array_equal(grad.view(np.uint32), grad_ref.view(np.uint32))
        # above the capacity: rejected before any work, outputs untouched
        dG.put(np.full(net.nparam, 7.0, np.float32))
        assert L.hpg_mlp_loss_grad_noalloc(net.h, dX.p, dT.p, nmax + 1, dL.p, dG.p) == -2
        assert (dG.get(np.zeros(net.nparam, np.float32)) == 7.0).all()
        assert L.hpg_mlp_loss_grad_noalloc(net.h, dX.p, dT.p, 0, dL.p, dG.p) == -2
        assert L.hpg_mlp_loss_grad_noalloc(net.h, None, dT.p, 10, dL.p, dG.p) == -2
    finally:
        for d in (dX, dT, dL, dG):
            d.free()
        net.close()
