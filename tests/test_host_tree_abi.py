"""The library's host tree core against the reference and the oracle (runs
without a GPU: topology calls work in host-only mode), and the C ABI: the
library loads and exports every function include/a5g.h declares."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import scenarios as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rules_a(lvl, c):
    if lvl < 2:
        return 1
    return 1 if (lvl < 4 and np.linalg.norm(c - 0.35) < 0.25) else -1


def _rules_b(lvl, c):
    if lvl < 2:
        return 1
    return 1 if (lvl < 5 and np.linalg.norm(c - 0.7) < 0.15) else -1


@pytest.mark.parametrize("dim,nc,base", [(2, 8, (2, 1, 1)), (3, 8, (1, 1, 1)),
                                         (2, 4, (1, 1, 1)), (3, 4, (1, 2, 1))])
def test_host_tree_matches_oracle(oracle_lib, dim, nc, base):
    import a5_streamer_amd as A
    ts = [M(dim, nc, 3, 0, 1.0 / nc, coarsen_to=2) for M in (O.OTree, A.Tree)]
    for T in ts:
        S.build_tree(T, dim, nc, _rules_a, n_base=base)
    for rules in (_rules_b, _rules_a):
        for _ in range(3):
            res = []
            for T in ts:
                b = T.boxes()
                c = S.box_centres(b, nc)
                fl = np.array([rules(int(b["lvl"][i]), c[i]) if b["in_use"][i] else 0
                               for i in range(len(b["lvl"]))], np.int32)
                res.append(T.adjust(fl))
            assert res[0] == res[1]
            a, b = ts[0].boxes(), ts[1].boxes()
            for k in a:
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
            for lvl in range(1, ts[0].info()["max_lvl"] + 2):
                for x, y in zip(ts[0].ref_info_lvl(lvl), ts[1].ref_info_lvl(lvl)):
                    np.testing.assert_array_equal(x, y)
        for reorder in (False, True):
            for T in ts:
                T.tidy_up(0.9, 0.25, 3, reorder)
            a, b = ts[0].boxes(), ts[1].boxes()
            for k in a:
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
            assert ts[0].info() == ts[1].info()
            for lvl in range(ts[0].info()["min_lvl"], ts[0].info()["max_lvl"] + 1):
                for x, y in zip(ts[0].level(lvl), ts[1].level(lvl)):
                    np.testing.assert_array_equal(x, y)


def test_host_tree_matches_reference(ref):
    import a5_streamer_amd as A
    ts = [M(3, 8, 3, 0, 1.0 / 8, coarsen_to=2) for M in (ref.RefTree, A.Tree)]
    for T in ts:
        S.build_tree(T, 3, 8, _rules_a, n_base=(2, 1, 1))
    a, b = ts[0].boxes(), ts[1].boxes()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_top_level_leaf_list(ref, oracle_lib):
    """When refinement reaches lvls_max, a$D_adjust_refinement leaves the loop
    without its exit, so max_lvl + 1 is past the top level and the leaf and
    parent lists of level lvls_max are not rebuilt (m_aX_core.f90:497-531).
    The leaf loops of the streamer then skip those boxes until a tidy_up
    renumbers them. The oracle and the library keep this behaviour."""
    import a5_streamer_amd as A
    ts = [M(2, 8, 3, 1, 1.0 / 8, coarsen_to=2, lvls_max=3)
          for M in (ref.RefTree, O.OTree, A.Tree)]
    for T in ts:
        S.build_tree(T, 2, 8, lambda lvl, c: 1, n_base=(2, 1, 1))
    for T in ts:
        assert T.info()["max_lvl"] == 3
        ids, leaves, parents = T.level(3)
        assert len(ids) == 2 * 16 and len(leaves) == 0 and len(parents) == 0
    for T in ts:
        T.tidy_up(0.9, 0.25, 3, True)
        ids, leaves, parents = T.level(3)
        assert len(leaves) == len(ids) == 32


def test_errors_are_reported():
    import a5_streamer_amd as A
    with pytest.raises(A.A5GError, match="n_cell should be even"):
        A.Tree(2, 7, 3, 0, 0.1)
    T = A.Tree(2, 8, 3, 0, 0.1)
    with pytest.raises(A.A5GError, match="unresolved neighbors"):
        T.set_base(np.array([[1, 1]]), np.array([[0, -1, -1, -1]]))
    with pytest.raises(A.A5GError, match="check_mg|GPU"):
        T.fas_fmg(False, True)


def test_abi_exports_every_declared_symbol():
    import a5_streamer_amd as A
    hdr = open(os.path.join(ROOT, "include", "a5g.h")).read()
    names = sorted(set(re.findall(r"\b(a5g_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) > 20
    lib = ctypes.CDLL(A.lib_path())
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_shape_arguments_are_checked():
    """a5g_set_shape / a5g_get_shape work on the host tree and reject an
    unknown kernel, box size, size class or shape."""
    import a5_streamer_amd as A
    T = A.Tree(2, 8, 3, 0, 1.0 / 8)
    for args in ((5, 8, 0, 0), (-1, 8, 0, 0), (0, 6, 0, 0), (0, 8, 3, 0), (0, 8, 0, 4)):
        with pytest.raises(Exception):
            T.set_shape(*args)
    assert T.get_shape(0, 8, 2) == 0
    T.set_shape(0, 8, 2, 3)
    assert T.get_shape(0, 8, 2) == 3
    assert T.get_shape(0, 16, 2) == 0
    T.set_shape(4, 8, 1, 2)
    assert T.get_shape(4, 8, 1) == 2 and T.get_shape(3, 8, 1) == 0
    T.close()
