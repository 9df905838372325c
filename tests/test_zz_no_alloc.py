"""fh_step_device_noalloc, the entry point that allocates nothing: the
caller's device records stepped in place with the library's own scratch
array, against fh_step on a handle with the same tunables (canonical order),
bit for bit, on two calls in a row, the second with a smaller flock."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CANON = 1
H2D, D2H = 1, 2


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.c_void_p, C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def on_device(h, words):
    """Device copy of the records (test plumbing: the caller's own buffer)."""
    w = np.ascontiguousarray(words, np.float32)
    ptr = C.c_void_p()
    assert h.hipMalloc(C.byref(ptr), w.nbytes) == 0
    assert h.hipMemcpy(ptr, w.ctypes.data, w.nbytes, H2D) == 0
    return ptr, w.nbytes


def from_device(h, ptr, nbytes, n):
    out = np.zeros((n, 15), np.float32)
    assert h.hipMemcpy(out.ctypes.data, ptr, nbytes, D2H) == 0
    return out


def cube(fh, n, seed):
    return fh.gen_cube(n, fh.cube_side_for(n), seed)


@pytest.mark.parametrize("prm", [{}, {"decide_group": 4, "cell": 6.0, "flags": 128 | 512}])
def test_noalloc_step_equals_the_handle(fh, prm):
    h = hip()
    for n, steps, seed, t0 in ((20_000, 7, 3, 0.0), (3000, 5, 4, 1.25)):
        w = cube(fh, n, seed)
        ptr, nbytes = on_device(h, w)
        try:
            t = fh.step_device_noalloc(ptr.value, n, t0, steps, **prm)
            assert h.hipDeviceSynchronize() == 0
            got = from_device(h, ptr, nbytes, n)
        finally:
            h.hipFree(ptr)
        flags = prm.get("flags", 0) | CANON
        with fh.Flock.from_words(w, t0, **{**prm, "flags": flags}) as f:
            f.step(steps)
            want, tw = f.download()
        assert t == tw, (n, t, tw)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), n


def test_noalloc_refuses_what_does_not_fit(fh):
    with pytest.raises(fh.FlockHipError) as e:
        fh.step_device_noalloc(0x1000, fh.NOALLOC_MAX_BIRDS + 1, 0.0, 1)
    assert e.value.status == fh.ERR_INVALID
    with pytest.raises(fh.FlockHipError) as e:
        fh.step_device_noalloc(0x1000, 100, 0.0, 1, flags=64)  # the resident layout needs a handle
    assert e.value.status == fh.ERR_INVALID
    with pytest.raises(fh.FlockHipError) as e:
        fh.step_device_noalloc(0x1000, 100, 0.0, 1, buckets=1 << 18)
    assert e.value.status == fh.ERR_INVALID
