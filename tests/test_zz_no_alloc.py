"""lkc_An_ZtSn_ww_noalloc: the device-buffer, no-allocation variant of lkc_An_ZtSn_ww.

Its output must equal the allocating entry point's bit for bit, on two calls in a row
with the second one smaller (the fixed scratch array keeps the first call's contents),
and a problem whose intermediates do not fit the fixed array is rejected before any
device work.  The test's own device buffers come from the HIP runtime liblkc links."""
import ctypes as C

import numpy as np
import pytest


def _problem(rng, nr_total, n_total, na, nn, ns):
    from lakcluster_c_amd import lkc
    rb = lkc.bsize(n_total)
    bits = rng.random((nr_total, n_total)) < 0.45
    A = np.zeros((nr_total, rb), np.uint8)
    A[:, : (n_total + 7) // 8] = np.packbits(bits, axis=1, bitorder="big")
    alist = np.sort(rng.permutation(nr_total)[:na]).astype(np.int32)
    nlist = np.sort(rng.permutation(n_total)[:nn]).astype(np.int32)
    m = rng.random(n_total) < 0.85
    amask = np.zeros(rb, np.uint8)
    amask[: (n_total + 7) // 8] = np.packbits(m, bitorder="big")
    L = rng.normal(size=(nn, ns)) * 5
    return A, alist, amask, L, nlist


class _Hip:
    """Device buffers through the HIP runtime liblkc itself links (one runtime per process)."""

    def __init__(self):
        self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.bufs = []

    def _ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self._ok(self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 1))))
        self.bufs.append(p)
        self._ok(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1))   # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        self._ok(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2))   # device to host
        return out

    def sync(self):
        self._ok(self.h.hipDeviceSynchronize())

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point():
    from lakcluster_c_amd import lkc
    lib = lkc.lib()
    rng = np.random.default_rng(7)
    hip = _Hip()
    try:
        for nr_total, n_total, na, nn, ns, B in [(300, 5000, 250, 4100, 3, 32), (90, 1500, 60, 1000, 2, 24)]:
            print(f"noalloc An_ZtSn_ww: rows {nr_total} n {n_total} na {na} nn {nn} ns {ns} B {B}", flush=True)
            A, alist, amask, L, nlist = _problem(rng, nr_total, n_total, na, nn, ns)
            want = lkc.an_ztsn_ww(A, alist, amask, L, nlist, n_total, B)
            dA, dal, dam, dL, dnl = hip.put(A), hip.put(alist), hip.put(amask), hip.put(L), hip.put(nlist)
            dout = hip.put(np.full((nr_total, ns), np.nan))
            rc = lib.lkc_An_ZtSn_ww_noalloc(dA, nr_total, dal, na, dam, dL, dnl, nn, n_total, ns, B, 1 << B, dout, None)
            assert rc == 0, lib.lkc_last_error()
            hip.sync()
            got = hip.get(dout, (nr_total, ns), np.float64)
            assert got.tobytes() == want.tobytes()
    finally:
        hip.free()


def test_noalloc_rejects_problems_beyond_its_scratch():
    """The size check comes before any device work (no pointer is touched)."""
    from lakcluster_c_amd import lkc
    lib = lkc.lib()
    fake = C.c_void_p(16)
    # 3 covariates x 4M n x 4 bytes of integers alone exceed the 32 MiB array
    rc = lib.lkc_An_ZtSn_ww_noalloc(fake, 10, fake, 10, fake, fake, fake, 10, 1 << 22, 3, 32, 1 << 32, fake, None)
    assert rc == -5
    assert b"scratch" in lib.lkc_last_error()
