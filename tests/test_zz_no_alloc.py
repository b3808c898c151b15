"""GPU test of bh_compute_accelerations_static, the plain force evaluation
that allocates nothing (static device buffer): through the C ABI, on device
buffers the test allocates itself through the HIP runtime the library uses,
against bh_compute_accelerations (PLAIN) on the same bodies, bit for bit,
for two calls in a row, the second with fewer bodies."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libbh.so is linked to."""

    def __init__(self, bh):
        bh.load()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        self.live.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.zeros(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def reference_accelerations(bh, pos, vel, m, theta, G_, soft2):
    """bh_compute_accelerations (PLAIN) on a context, back in input order."""
    with bh.Engine(theta=theta) as eng:
        eng.set_bodies(pos, vel, m)
        acc = eng.compute_accelerations(form=bh.FORM_PLAIN, theta=theta, G_=G_, soft2=soft2)
        ids = eng.get_bodies()["ids"].astype(np.int64)
    out = np.empty_like(acc)
    out[ids] = acc
    return out


@pytest.mark.parametrize("theta,G_,soft2", [(1.05, None, 0.0), (0.5, 6.674e-11, 1e-6)])
def test_static_matches_context_evaluation(bh, theta, G_, soft2):
    G_ = bh.G if G_ is None else G_
    lib = bh.load()
    hip = Hip(bh)
    try:
        for n, seed in ((19054, 1), (4000, 2)):   # the second call has fewer bodies
            say("static evaluation", n, "bodies, theta", theta)
            pos, vel, m = bh.generate(bh.GEN_BELT, seed, n)
            dpos, dm, dacc = hip.put(pos), hip.put(m), hip.alloc(24 * n)
            rc = lib.bh_compute_accelerations_static(n, dpos, dm, theta, G_, soft2, dacc, None)
            assert rc == 0, lib.bh_static_last_error()
            hip.sync()
            got = hip.get(dacc, (n, 3))
            want = reference_accelerations(bh, pos, vel, m, theta, G_, soft2)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
            # the inputs are left as they were
            assert np.array_equal(hip.get(dpos, (n, 3)).view(np.uint64), pos.view(np.uint64))
    finally:
        hip.free()


def test_static_limits(bh):
    lib = bh.load()
    assert lib.bh_compute_accelerations_static(0, None, None, 1.05, bh.G, 0.0, None, None) == 0
    big = 65537
    rc = lib.bh_compute_accelerations_static(big, C.c_void_p(8), C.c_void_p(8), 1.05, bh.G, 0.0, C.c_void_p(8), None)
    assert rc == bh.ERR_ARG and b"65536" in lib.bh_static_last_error()
    assert lib.bh_compute_accelerations_static(10, None, None, 1.05, bh.G, 0.0, None, None) == bh.ERR_ARG
    assert lib.bh_compute_accelerations_static(10, C.c_void_p(8), C.c_void_p(8), -1.0, bh.G, 0.0,
                                               C.c_void_p(8), None) == bh.ERR_ARG
