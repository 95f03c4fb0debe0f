"""ssa_ic_apply_noalloc: the preconditioner apply for callers that cannot
allocate device memory.  Through the C ABI, against ssa_ic_apply on the same
input (same kernels, same order: bit for bit), on two calls in a row with the
second on a smaller system (the shared scratch array keeps the first call's
contents), for one-system, mixed-precision, level-path-with-scaling and
batched handles; the call leaves the device's free memory unchanged, and a
system larger than the scratch array is rejected.

Device buffers come from the HIP runtime the library itself links
(libamdhip64.so.7), called through ctypes, so the test runs in one HIP
runtime whatever else the process has loaded."""
import ctypes as C

import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class Hip:
    """The few HIP runtime calls the test needs."""

    def __init__(self):
        _lib.load()   # the library is loaded first: its runtime is the one found
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.L.hipDeviceSynchronize.argtypes = []
        self.L.hipSetDevice.argtypes = [C.c_int]

    def ok(self, rc):
        assert rc == 0, f"HIP call failed: {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)))
        return p.value

    def free(self, p):
        self.ok(self.L.hipFree(C.c_void_p(p)))

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.malloc(a.nbytes)
        self.ok(self.L.hipMemcpy(C.c_void_p(p), a.ctypes.data, a.nbytes, _H2D))
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.ok(self.L.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, _D2H))
        return out

    def sync(self):
        self.ok(self.L.hipDeviceSynchronize())

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        self.ok(self.L.hipMemGetInfo(C.byref(f), C.byref(t)))
        return f.value


@pytest.fixture(scope="module")
def hip():
    _lib.context(0)    # device 0 current, the library's scratch resolved
    return Hip()


def noalloc(h, x_ptr, y_ptr):
    return h.lib.ssa_ic_apply_noalloc(h.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr))


def varcoef(m, seed):
    return P.laplace7_varcoef(m, 10.0 ** P.uniform(seed, m ** 3, -2.0, 2.0))


def rhs(n, complex_, seed):
    x = P.uniform(seed, n, -1.0, 1.0)
    if complex_:
        x = x + 1j * P.uniform(seed + 1, n, -1.0, 1.0)
    return x


def case_handles(case):
    """(big, small) handles of one kind."""
    if case == "f64-walk":
        cfg = _lib.default_config(use_abmc=1)
        return [_lib.IcHandle(P.poisson(m).A, cfg) for m in (48, 12)]
    if case == "c128-mixed":
        cfg = _lib.default_config(use_abmc=1, abmc_num_colors=8, precision=_lib.PRECISION_MIXED)
        return [_lib.IcHandle(P.eddy(m, cond_frac=1.0, sigma=0.1).A, cfg) for m in (12, 6)]
    if case == "level-scaled":
        cfg = _lib.default_config(use_abmc=0, diagonal_scaling=1)
        return [_lib.IcHandle(varcoef(m, 3), cfg) for m in (16, 8)]
    cfg = _lib.default_config(use_abmc=1)
    big = _lib.BatchHandle([varcoef(m, 5 + m) for m in (14, 9, 11)], cfg)
    return [big, _lib.BatchHandle([varcoef(m, 7 + m) for m in (6, 5)], cfg)]


@pytest.mark.parametrize("case", ["f64-walk", "c128-mixed", "level-scaled", "batch"])
def test_noalloc_apply_matches_apply(hip, case):
    for k, h in enumerate(case_handles(case)):
        x = rhs(h.n, h.dtype == np.complex128, 11 + k)
        xd = hip.upload(x)
        got = hip.upload(np.full_like(x, np.nan))
        want = hip.upload(np.zeros_like(x))
        try:
            _lib.check(noalloc(h, xd, got))
            h.apply_device(xd, want)
            hip.sync()
            w = hip.download(want, x)
            assert np.array_equal(hip.download(got, x), w), (case, k)
            # again, after the regular apply has run on this handle
            _lib.check(noalloc(h, xd, got))
            hip.sync()
            assert np.array_equal(hip.download(got, x), w), (case, k)
            # and against the host-buffer apply
            assert np.array_equal(h.apply(x), w), (case, k)
        finally:
            for p in (xd, got, want):
                hip.free(p)


def test_noalloc_apply_allocates_nothing(hip):
    cfg = _lib.default_config(use_abmc=1)
    # the kernels' code is loaded by a regular apply on another handle first
    warm = _lib.IcHandle(P.poisson(8).A, cfg)
    warm.apply(np.ones(warm.n))
    h = _lib.IcHandle(P.poisson(40).A, cfg)   # its work vectors would take ~5.6 MB
    x = rhs(h.n, False, 3)
    xd = hip.upload(x)
    yd = hip.upload(np.zeros_like(x))
    try:
        hip.sync()
        free0 = hip.free_bytes()
        for _ in range(3):
            _lib.check(noalloc(h, xd, yd))
        hip.sync()
        assert hip.free_bytes() == free0
        assert np.array_equal(hip.download(yd, x), h.apply(x))
    finally:
        hip.free(xd)
        hip.free(yd)


def test_noalloc_apply_rejects_a_system_beyond_the_scratch(hip):
    import scipy.sparse as sp
    from sparsesolv_amd.problems import CSR
    # complex<double>, FP64: 48 bytes a row -> 3,000,000 rows need ~144 MB
    n = 3_000_000
    A = CSR.from_scipy(sp.diags(np.full(n, 2.0 + 0.5j)).tocsr())
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=1))
    x = np.ones(n, dtype=np.complex128)
    xd = hip.upload(x)
    yd = hip.upload(np.zeros_like(x))
    try:
        rc = noalloc(h, xd, yd)
        assert rc == _lib.SSA_ERR_ARGUMENT
        msg = h.lib.ssa_last_error().decode()
        assert "scratch" in msg and str(128 << 20) in msg
        hip.sync()
        assert not np.any(hip.download(yd, x) != 0)
    finally:
        hip.free(xd)
        hip.free(yd)
