"""The allocation-free FFT entry point (cpsd_fft3d_batch_noalloc) against the
allocating one (cpsd_fft3d_batch) through the C-ABI: two calls in a row on
the caller's device arrays, the second on a smaller input (fewer, smaller
transforms), both directions, plus the rejection of a request larger than
the library's scratch array.  Bar: the FFT parity bar of the GPU tests
(1e-13 relative to the transform's scale, times log2 N)."""
import ctypes as C
import sys

import numpy as np
import pytest

import cpsd_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """just enough of the HIP runtime (through ctypes) to own device memory
    and a stream on the test's side"""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]

    def ok(self, rc, what):
        assert rc == 0, "%s failed: %d" % (what, rc)

    def upload(self, a):
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, a):
        self.ok(self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, 2), "hipMemcpy D2H")
        return a


def run_noalloc(hip, lib, direction, n, a, stream):
    d = hip.upload(a)
    try:
        rc = lib.cpsd_fft3d_batch_noalloc(direction, n, a.shape[0], d, stream)
        assert rc == 0, lib.cpsd_last_error()
        hip.ok(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
        return hip.download(d, np.empty_like(a))
    finally:
        hip.L.hipFree(d)


def check(direction, n, nb, seed, hip, lib, stream):
    say("noalloc fft dir=%+d N=%d nbatch=%d" % (direction, n, nb))
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nb, (n + 2) * n * n))
    ref = cpsd_amd.fft3d(direction, n, a)            # the allocating entry point
    got = run_noalloc(hip, lib, direction, n, a, stream)
    for b in range(nb):
        scale = np.sqrt(n ** 3) * np.max(np.abs(a[b]))
        assert np.max(np.abs(got[b] - ref[b])) < 1e-13 * scale * np.log2(n) * 4
    if direction > 0:
        assert np.all(got.reshape(nb, n, n, n + 2)[:, :, :, n:] == 0.0)


def test_noalloc_fft_matches_allocating_entry_point(gpu):
    lib = cpsd_amd.hip_lib()
    hip = Hip()
    s = C.c_void_p()
    hip.ok(hip.L.hipStreamCreate(C.byref(s)), "hipStreamCreate")
    try:
        # first call: 3 transforms of 64^3; second: 2 of 32^3 on the same
        # scratch (the first call's tables and intermediate are still there)
        for direction in (+1, -1):
            check(direction, 64, 3, 11 + direction, hip, lib, s)
            check(direction, 32, 2, 21 + direction, hip, lib, s)
        # the null stream too
        check(+1, 48, 2, 31, hip, lib, None)
    finally:
        hip.L.hipStreamDestroy(s)


def test_noalloc_fft_rejects_oversize(gpu):
    lib = cpsd_amd.hip_lib()
    say("noalloc fft rejects oversize")
    assert lib.cpsd_fft3d_batch_noalloc_max(128) == 3
    assert lib.cpsd_fft3d_batch_noalloc_max(96) == 9
    assert lib.cpsd_fft3d_batch_noalloc_max(7) == 0
    # refused before anything is enqueued, so a dummy non-null pointer is
    # never dereferenced
    rc = lib.cpsd_fft3d_batch_noalloc(+1, 128, 4, C.c_void_p(256), None)
    assert rc != 0
    assert b"scratch" in lib.cpsd_last_error()
