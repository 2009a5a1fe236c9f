"""fr_t0_reduce_noalloc (the allocation-free reduction) against fr_t0_reduce:
the same kernels on the same partition, so every result is bit-identical.
Two calls in a row, the second on a smaller field, reuse the library's fixed
device scratch; bad arguments are refused on the host."""
import ctypes as C

import numpy as np
import pytest

import first_repo_amd as fr
from gpu_util import dbuf, dptr

pytestmark = pytest.mark.gpu


def test_reduce_noalloc_matches_reduce():
    L = fr.lib()
    rng = np.random.default_rng(21)
    cases = []
    for n in [(40, 24, 36), (9, 7, 5)]:  # the second call takes a smaller field
        g = fr.make_geom(*n, L=(1.0, 2.0, 3.0))
        a = rng.standard_normal((g.Nz, g.Ny, g.Nx))
        da = dbuf(a)
        want = []
        for op in range(6):
            r = C.c_double()
            fr.check(L.fr_t0_reduce(g, dptr(da), op, C.byref(r), None), "fr_t0_reduce")
            want.append(r.value)
        cases.append((g, a, da, want))
    for g, a, da, want in cases:
        for op in range(6):
            r = C.c_double()
            fr.check(L.fr_t0_reduce_noalloc(g, dptr(da), op, C.byref(r), None),
                     "fr_t0_reduce_noalloc")
            assert r.value == want[op], (g.Nx, op, r.value, want[op])
        inner = a[1:-1, 1:-1, 1:-1]
        assert abs(want[0] - inner.sum()) <= 1e-12 * np.abs(inner).sum()
        assert want[2] == inner.max() and want[3] == inner.min()
    # host-side refusals: no kernel runs
    g, _, da, _ = cases[0]
    r = C.c_double()
    assert L.fr_t0_reduce_noalloc(g, dptr(da), 6, C.byref(r), None) == 1001
    bad = fr.make_geom(4, 4, 4)
    bad.bs = 2
    assert L.fr_t0_reduce_noalloc(bad, dptr(da), 0, C.byref(r), None) == 1001
    assert L.fr_t0_reduce_noalloc(g, None, 0, C.byref(r), None) == 1001
