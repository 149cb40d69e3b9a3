"""rsd_dot_f64_noalloc, the allocation-free dot product, against rsd_dot_f64
through the C-ABI: the same bits on two calls in a row (the second with a
smaller input, after the first left its partials in the shared array), a
zero-length input, and the size limit refused with a status code."""
import ctypes as C

import numpy as np
import pytest


@pytest.mark.gpu
def test_dot_noalloc_matches_dot(ctx):
    import residuum_hip as rh
    L = rh.lib()
    rng = np.random.default_rng(31)
    out = ctx.zeros(2)
    for n in (5_000_003, 70_001, 0):
        x = ctx.from_numpy(rng.standard_normal(n))
        y = ctx.from_numpy(rng.standard_normal(n))
        want = C.c_double()
        rh.check(L.rsd_dot_f64(ctx.h, n, rh.dptr(x), rh.dptr(y), None, C.byref(want)))
        rh.check(L.rsd_dot_f64_noalloc(ctx.h, n, rh.dptr(x), rh.dptr(y), C.c_void_p(out.ptr)))
        ctx.sync()
        got = out.numpy()[0]
        print("n", n, got, want.value, flush=True)
        assert got == want.value
        x.free()
        y.free()
    # longer than the scratch array serves: refused on the host before any launch
    big = (1 << 24) + 1
    assert L.rsd_dot_f64_noalloc(ctx.h, big, C.c_void_p(out.ptr), C.c_void_p(out.ptr), C.c_void_p(out.ptr)) == -1
    assert b"2^24" in L.rsd_last_error_string()
