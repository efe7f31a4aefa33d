"""The streamer diagnostics on the GPU (a5g_reduction_loc, a5g_cc_axis
through the C ABI) against the oracle, bit for bit, on random data, on data
with many ties and with thresholds no cell passes."""
import numpy as np
import pytest

import oracle as O
import scenarios as S
from test_oracle_diag import NC, _fill, _rule

pytestmark = pytest.mark.gpu


def _pair(dim):
    import a5_streamer_amd as A
    ts = [M(dim, NC, 3, 2, 1.0 / NC, coarsen_to=2, n_boxes=50) for M in (O.OTree, A.Tree)]
    for T in ts:
        S.build_tree(T, dim, NC, _rule(dim))
    return ts


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("ties", [False, True])
def test_reduction_loc_bit_exact(oracle_lib, dim, ties):
    print("reduction_loc dim", dim, "ties", ties, flush=True)
    ts = _pair(dim)
    _fill(ts, dim, 17 + dim, ties)
    for d in range(1, dim + 1):
        for kind, var, thr, init in ((1, 1, 0.0, -1e99), (1, 2, 0.0, -1e99), (1, 1, 0.0, 1e9),
                                     (2, 2, 0.5, 0.0), (2, 3, 3.5, 0.0), (2, 1, 1e9, 0.0),
                                     (2, 1, 1e9, -1.0)):
            a = ts[0].reduction_loc(kind, var, d, thr, init)
            b = ts[1].reduction_loc(kind, var, d, thr, init)
            assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]), \
                (kind, var, d, thr, init, a, b)


def test_cc_axis_bit_exact(oracle_lib):
    print("cc_axis 2d", flush=True)
    ts = _pair(2)
    _fill(ts, 2, 33)
    a = ts[0].cc_axis([1, 3], [2, 1])
    b = ts[1].cc_axis([1, 3], [2, 1])
    assert a.shape == b.shape and a.shape[1] > 0
    np.testing.assert_array_equal(a, b)
    # no variables: only the coordinates
    np.testing.assert_array_equal(ts[1].cc_axis([], []), a[:1])
