"""bt_sample_noalloc (the sampler with its buffers in a static device array)
against bt_sample through the C ABI: identical indices and boundary counts on
two calls in a row, the second with a smaller state and fewer shots; a
request beyond the array's size is an error that leaves the library usable."""
import ctypes

import numpy as np
import pytest

from refcirc import random_spec, to_bt

pytestmark = pytest.mark.gpu


def _sample(L, fn, state, u):
    out = np.zeros(u.size, dtype=np.uint64)
    nb = ctypes.c_int64(-1)
    L.check(getattr(L.lib(), fn)(state.handle, u.size, L.dptr(u), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                 ctypes.byref(nb)))
    return out, nb.value


def _state(bt, N, prec, seed):
    rng = np.random.default_rng(seed)
    st = bt.zero_state(N, prec)
    return bt.apply(st, to_bt(random_spec(N, 8 * N, rng), bt))


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_noalloc_matches_sample(bt, prec):
    from bluetangle_jl_amd import _lib as L
    rng = np.random.default_rng(31)
    # the larger call first, then a smaller state and fewer shots: leftovers of
    # the first call in the shared array must not leak into the second
    for N, shots in ((22, 300_000), (11, 1_000)):
        st = _state(bt, N, prec, N)
        u = rng.random(shots)
        a, na = _sample(L, "bt_sample", st, u)
        b, nb = _sample(L, "bt_sample_noalloc", st, u)
        assert np.array_equal(a, b), (N, shots)
        assert na == nb, (N, shots)
        assert a.max() < 2 ** N
    # beyond the static array: an error, then the entry point still works
    big = _state(bt, 12, prec, 5)
    u = rng.random(2_000_000)
    out = np.zeros(u.size, dtype=np.uint64)
    nb = ctypes.c_int64()
    rc = L.lib().bt_sample_noalloc(big.handle, u.size, L.dptr(u), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                   ctypes.byref(nb))
    assert rc == -1
    u = rng.random(512)
    a, na = _sample(L, "bt_sample", big, u)
    b, nb2 = _sample(L, "bt_sample_noalloc", big, u)
    assert np.array_equal(a, b) and na == nb2
