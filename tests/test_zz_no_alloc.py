"""rwpt_kde_device, the reconstruction that allocates nothing, against
rwpt_kde through the C ABI: bit-identical density and histogram on two
calls in a row on one GPU, the second with fewer particles and a smaller
grid (the shared scratch array keeps the first call's values, so a
missing clear would show), plus its argument checks."""
import ctypes as C

import numpy as np
import pytest

from rwpt import Engine, RwptError, engine, problems

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """The HIP runtime librwpt.so is linked against: symbols looked up
    through the library's own handle resolve in its dependencies, so the
    test's buffers come from the runtime the engine uses even when another
    HIP runtime (torch's own copy) is loaded in the process."""
    global _hip
    if _hip is None:
        _hip = engine.lib()
        for f in ("hipMalloc", "hipMemset", "hipMemcpy", "hipFree"):
            getattr(_hip, f).restype = C.c_int
    return _hip


class DeviceBuffer:
    """A device array of doubles allocated by the test (the entry point
    under test allocates nothing itself)."""

    def __init__(self, n):
        self.n = n
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0
        # poison, so that values the call does not write would show
        assert hip().hipMemset(self.ptr, 0x7f, C.c_size_t(8 * n)) == 0

    def host(self):
        out = np.empty(self.n)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                               C.c_size_t(8 * self.n), 2) == 0   # device to host
        return out

    def free(self):
        hip().hipFree(self.ptr)


def _compare(E, origin, bins, nb, **kw):
    want_d, want_h = E.kde(origin, bins, nb, **kw)
    info_host = E.kde_info()
    n = int(np.prod(nb))
    d, h = DeviceBuffer(n), DeviceBuffer(n)
    try:
        E.kde_device(d.ptr.value, h.ptr.value, origin, bins, nb, **kw)
        got_d, got_h = d.host(), h.host()
    finally:
        d.free()
        h.free()
    print("kde_device", nb, kw, "sum", float(got_h.sum()), flush=True)
    assert np.array_equal(got_h, want_h)
    assert np.array_equal(got_d, want_d)
    assert E.kde_info() == info_host
    return got_h


def test_kde_device_equals_kde_on_two_calls():
    P = problems.heterogeneous_flow(40, 30, 8, corr_len=4.0, seed=7)
    P.aL, P.aT, P.dmeff = np.array([0.5]), np.array([0.05]), np.array([1e-4])
    P.ts_params = (0.5, 0.0)
    origin = (0.0, 0.0, float(P.bot.min()))
    E = Engine(P, precision=1)
    try:
        rel = problems.release_in_block(P, 200_000, (2, 8), (8, 22), (1, 8), seed=3)
        E.load_particles(*rel)
        E.advance(4.0)
        h1 = _compare(E, origin, (1.0, 1.0, 1.0), (P.ncol, P.nrow, P.nlay), mode=1)
        assert h1.sum() > 100_000
        # second call: fewer particles on a smaller, coarser 2-D grid
        rel = problems.release_in_block(P, 30_000, (2, 5), (10, 14), (2, 6), seed=4)
        E.load_particles(*rel)
        E.set_time(0.0)
        E.advance(2.0)
        h2 = _compare(E, origin, (2.0, 2.0, 0.0), (P.ncol // 2, P.nrow // 2, 1), mode=1,
                      unit_volume=True, scaling_factor=0.5)
        assert 0 < h2.sum() < h1.sum()
        # histogram only, and the bin-distance initial smoothing
        _compare(E, origin, (1.0, 1.0, 1.0), (P.ncol, P.nrow, P.nlay), mode=0)
        _compare(E, origin, (1.0, 1.0, 1.0), (P.ncol, P.nrow, P.nlay), mode=1,
                 initial_smoothing=1, smoothing_factor=2.0)
    finally:
        E.close()


def test_kde_device_rejects_what_needs_allocation():
    P = problems.config0()
    E = Engine(P)
    try:
        rel = problems.release_in_cell(P, 1000, 1, 25, 10, lz=0.5)
        E.load_particles(*rel)
        d = DeviceBuffer(8)
        try:
            for kw in ({"n_opt_loops": 2}, {"weighted": True}):
                with pytest.raises(RwptError, match="-5"):
                    E.kde_device(d.ptr.value, d.ptr.value, (0, 0, 0), (1, 1, 1),
                                 (2, 2, 2), **kw)
            with pytest.raises(RwptError, match="-5"):
                E.kde_device(d.ptr.value, d.ptr.value, (0, 0, 0), (1, 1, 1),
                             (2048, 2048, 2))   # above the scratch grid
            with pytest.raises(RwptError, match="-1"):
                E.kde_device(0, d.ptr.value, (0, 0, 0), (1, 1, 1), (2, 2, 2))
        finally:
            d.free()
    finally:
        E.close()
