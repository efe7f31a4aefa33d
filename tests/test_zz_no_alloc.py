"""a5g_box_summaries_noalloc, the entry point that allocates nothing,
against a5g_box_summaries on the same trees, through the C ABI: two calls in
a row on a larger tree and then on a smaller one (the fixed scratch keeps
the first call's values beyond the second's boxes), bit for bit. Also its
refusal of a tree that changed since its last upload."""
import ctypes as C

import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu


def _tree(dim, rules, seed):
    import a5_streamer_amd as A
    nc = 8
    T = A.Tree(dim, nc, 9, 2, 1.0 / nc, coarsen_to=2)
    S.build_tree(T, dim, nc, rules, n_base=(2, 2, 2))
    rng = np.random.default_rng(seed)
    b = T.boxes()
    n = len(b["lvl"])
    bl = T.box_len(nc)
    for iv in (1, 6, 7):
        T.set_cc_all(iv, rng.standard_normal((n, bl)) * (1e18 if iv == 1 else 1e6))
    T.streamer_setup()
    return T


def _call_abi(name, T):
    from a5_streamer_amd import lib
    L = lib.load()
    n = T.info()["max_id"]
    out = np.zeros((n, 3))
    st = getattr(L, name)(T.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert st == 0, L.a5g_last_error().decode()
    return out


def _big(lvl, c):
    return 1 if lvl < 3 or (lvl < 4 and np.linalg.norm(c - 0.5) < 0.3) else 0


def _small(lvl, c):
    return 1 if lvl < 2 else 0


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_summaries_match(dim):
    print("no-alloc summaries, dim", dim, flush=True)
    big = _tree(dim, _big, 1)
    small = _tree(dim, _small, 2)
    assert small.info()["max_id"] < big.info()["max_id"]
    for T in (big, small):
        ref = _call_abi("a5g_box_summaries", T)
        for _ in range(2):
            np.testing.assert_array_equal(_call_abi("a5g_box_summaries_noalloc", T), ref)
    big.close()
    small.close()


def test_noalloc_summaries_refuse_a_changed_tree():
    from a5_streamer_amd import lib
    T = _tree(2, _small, 3)
    T.box_summaries()
    b = T.boxes()
    flags = np.where(b["lvl"] >= 1, 1, 0).astype(np.int32)
    T.adjust(flags)
    with pytest.raises(lib.A5GError, match="changed since its last upload"):
        T.box_summaries_noalloc()
    # after a call that may allocate, it works again
    ref = T.box_summaries()
    np.testing.assert_array_equal(T.box_summaries_noalloc(), ref)
    T.close()
