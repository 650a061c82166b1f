"""opalsc_beam_parameters_noalloc (device arrays, scratch in a file-scope
device array) against opalsc_beam_parameters (host arrays, the context's
buffers): the same kernels over the same data, so the statistics agree bit for
bit, on two calls in a row, the second with a smaller bunch."""
import ctypes as C

import numpy as np
import pytest

from conftest import gaussian_bunch

pytestmark = pytest.mark.gpu

sc = pytest.importorskip("src_amd")


class _Hip:
    """hipMalloc/hipMemcpy of the HIP runtime the library links, via ctypes (test-side allocation)."""

    def __init__(self):
        path = None
        for line in open("/proc/self/maps"):
            if "libamdhip64" in line:
                path = line.split()[-1]
                break
        self.hip = C.CDLL(path or "libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p

    def get_bytes(self, p, nbytes):
        out = (C.c_char * nbytes)()
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out, p, C.c_size_t(nbytes), 2) == 0
        return bytes(out)

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(L, H, R, P, mass):
    N = len(R)
    dR, dP = H.put(R), H.put(P)
    dS = H.alloc(C.sizeof(sc.opalsc.Stats))
    # a stale value in the output: every byte must be written by the call
    H.hip.hipMemset(dS, 0x7f, C.c_size_t(C.sizeof(sc.opalsc.Stats)))
    rc = L.opalsc_beam_parameters_noalloc(N, dR, dP, mass, dS, None)
    assert rc == 0, L.opalsc_last_error(None)
    return H.get_bytes(dS, C.sizeof(sc.opalsc.Stats))


def test_noalloc_beam_parameters_equal_the_allocating_entry_point():
    L = sc.lib()
    H = _Hip()
    try:
        with sc.SpaceChargeContext((8, 8, 8), max_particles=300000) as c:
            for N, seed in ((300000, 41), (1000, 42), (0, 43)):
                R, P, _ = gaussian_bunch(N, seed=seed, p0=0.7771) if N else (np.zeros((0, 3)), np.zeros((0, 3)), None)
                s = sc.opalsc.Stats()
                c._c(L.opalsc_beam_parameters(c._h, N, sc.opalsc._ptr(R), sc.opalsc._ptr(P), sc.M_P_EV,
                                              C.byref(s)))
                got = _noalloc(L, H, R, P, sc.M_P_EV)
                assert got == bytes(s), N
                if N:
                    d = sc.opalsc.stats_to_dict(sc.opalsc.Stats.from_buffer_copy(got))
                    assert d["n"] == N and np.all(np.asarray(d["rrms"]) > 0)
                H.free()
    finally:
        H.free()


def test_noalloc_rejects_host_and_short_arrays():
    L = sc.lib()
    H = _Hip()
    try:
        R = np.zeros((100, 3))
        dR = H.put(R)
        dS = H.alloc(C.sizeof(sc.opalsc.Stats))
        # host memory for the device arrays
        assert L.opalsc_beam_parameters_noalloc(100, R.ctypes.data, dR, 1.0, dS, None) == -1
        assert b"device memory" in L.opalsc_last_error(None)
        # arrays shorter than N x 3 doubles
        assert L.opalsc_beam_parameters_noalloc(101, dR, dR, 1.0, dS, None) == -1
        # an output that is not device memory
        host_out = sc.opalsc.Stats()
        assert L.opalsc_beam_parameters_noalloc(100, dR, dR, 1.0, C.addressof(host_out), None) == \
            -1
        assert L.opalsc_beam_parameters_noalloc(-1, dR, dR, 1.0, dS, None) == -1
    finally:
        H.free()
