"""Multigrid parity: the HIP library against the oracle (the C restatement,
itself pinned bit-for-bit to the reference in test_oracle_ref.py), on the
same seeded trees and fields. The FP64 path is bit-exact: every comparison is
exact equality over whole boxes, ghost layers included."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

pytestmark = pytest.mark.gpu


def _rules(lvl, c):
    if lvl < 2:
        return 1
    if lvl < 4 and np.linalg.norm(c - 0.3) < 0.25:
        return 1
    return -1


CASES = [
    # dim, n_cell, coord, base
    (2, 8, 0, (2, 1, 1)),
    (2, 8, 1, (1, 2, 1)),
    (3, 8, 0, (1, 1, 1)),
    (2, 4, 0, (1, 1, 1)),
    (3, 4, 0, (2, 1, 1)),
]
# the box sizes of the benchmarks, where a lane handles several cells
LARGE = [
    (3, 16, 0, (1, 1, 1)),
    (2, 16, 1, (1, 2, 1)),
    (2, 32, 0, (1, 1, 1)),
]


def _pair(oracle_lib, dim, nc, coord, base, n_cc=4, seed=1):
    import a5_streamer_amd as A
    ts = []
    for M in (O.OTree, A.Tree):
        T = M(dim, nc, n_cc, 2, 1.0 / nc, coarsen_to=2, coord=coord)
        S.build_tree(T, dim, nc, _rules, n_base=base)
        ts.append(T)
    data = S.fill_random(ts[0], dim, list(range(1, n_cc + 1)), seed=seed)
    for (id_, iv), v in data.items():
        ts[1].set_cc(id_, iv, v)
    return ts


def _same(ts, dim, ivs):
    for iv in ivs:
        a = S.all_cc(ts[0], dim, iv)
        b = S.all_cc(ts[1], dim, iv)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg="box %d var %d" % (k, iv))


@pytest.mark.parametrize("dim,nc,coord,base", CASES + LARGE)
def test_fmg_vcycle_bit_exact(oracle_lib, dim, nc, coord, base):
    ts = _pair(oracle_lib, dim, nc, coord, base)
    kinds = np.array([0, 1, 2, 0, 1, 2][:2 * dim], np.int32)
    vals = np.arange(2 * dim, dtype=float) + 0.5
    for T in ts:
        T.mg_setup(1, 2, 3, n_base=4, bc_kind=kinds, bc_val=vals, auto_ops=True)
    for it in range(2):
        for T in ts:
            T.fas_fmg(True, it == 0)
        _same(ts, dim, (1, 2, 3))
    max_lvl = ts[0].info()["max_lvl"]
    for T in ts:
        T.fas_vcycle(max_lvl, True)
    _same(ts, dim, (1, 2, 3, 4))


@pytest.mark.parametrize("dim,nc,coord,base", CASES[:3])
@pytest.mark.parametrize("rb,bc", [(0, 0), (1, 1), (2, 1), (2, 2)])
def test_ghost_rules_bit_exact(oracle_lib, dim, nc, coord, base, rb, bc):
    ts = _pair(oracle_lib, dim, nc, coord, base)
    kinds = np.array([2, 1, 0, 2, 1, 0][:2 * dim], np.int32)
    for T in ts:
        T.mg_setup(1, 2, 3, bc_kind=kinds, bc_val=np.ones(2 * dim) * 3.0)
        T.gc_tree(4, rb, bc)
    _same(ts, dim, (4,))


@pytest.mark.parametrize("dim,nc,coord,base", CASES)
def test_restrict_copy_minmax(oracle_lib, dim, nc, coord, base):
    ts = _pair(oracle_lib, dim, nc, coord, base)
    for T in ts:
        T.restrict_tree(2)
        T.tree_copy_cc(2, 3)
    _same(ts, dim, (2, 3))
    assert ts[0].tree_minmax(1) == ts[1].tree_minmax(1)


def test_default_ops_cylindrical_is_cartesian(oracle_lib):
    """Without the automatic operators the library default is the Cartesian
    Laplacian even in cylindrical coordinates (m_aX_mg.f90:148-151)."""
    ts = _pair(oracle_lib, 2, 8, 1, (1, 1, 1))
    for T in ts:
        T.mg_setup(1, 2, 3, bc_kind=np.zeros(4, np.int32))
        T.fas_fmg(True, True)
    _same(ts, 2, (1, 2, 3))


@pytest.mark.parametrize("dim,nc,coord,base", CASES + LARGE[:1])
def test_fmg_vcycle_variable_eps_bit_exact(oracle_lib, dim, nc, coord, base):
    """Normal, constant-eps and variable-eps boxes in one tree: the operator
    tags and the lpld/clpld operators, Gauss-Seidel and correction."""
    from test_oracle_ref import eps_field
    ts = _pair(oracle_lib, dim, nc, coord, base, n_cc=5, seed=5)
    e = eps_field(ts[0], dim, 5, 1)
    for i, v in e.items():
        ts[1].set_cc(i, 5, v)
    kinds = np.array([0, 0, 1, 2, 0, 2][:2 * dim], np.int32)
    for T in ts:
        T.mg_setup(1, 2, 3, i_eps=5, n_down=2, n_up=2, n_base=4, bc_kind=kinds,
                   bc_val=np.linspace(-1, 2, 2 * dim), auto_ops=True)
    for it in range(2):
        for T in ts:
            T.fas_fmg(True, it == 0)
        _same(ts, dim, (1, 2, 3))
    for T in ts:
        T.fas_vcycle(ts[0].info()["max_lvl"], True)
    _same(ts, dim, (1, 2, 3, 4, 5))
    # refine further: new boxes start untagged and are tagged at their first
    # multigrid call, with the eps they hold by then
    mx = ts[0].info()["max_lvl"]
    for T in ts:
        b = T.boxes()
        c = S.box_centres(b, nc)
        fl = np.array([1 if (b["in_use"][i] and b["lvl"][i] == mx
                             and np.linalg.norm(c[i] - 0.3) < 0.12) else 0
                       for i in range(len(b["lvl"]))], np.int32)
        T.adjust(fl)
    data = S.fill_random(ts[0], dim, [1, 2, 3], seed=6)
    for (i, iv), v in data.items():
        ts[1].set_cc(i, iv, v)
    e = eps_field(ts[0], dim, 5, 1)
    for i, v in e.items():
        ts[1].set_cc(i, 5, v)
    for T in ts:
        T.fas_fmg(False, False)
    _same(ts, dim, (1, 2, 3))


def test_graph_replays_match_plain_launches(oracle_lib, monkeypatch):
    """Multigrid calls are captured as HIP graphs on first use and replayed
    after: every replay, with other work and a topology change in between,
    must equal the plain launches (A5G_GRAPHS=0) and the oracle."""
    import a5_streamer_amd as A
    dim, nc, coord, base = 3, 8, 0, (1, 1, 1)
    ts = []
    for M, graphs in ((O.OTree, None), (A.Tree, "1"), (A.Tree, "0")):
        if graphs is not None:
            monkeypatch.setenv("A5G_GRAPHS", graphs)
        T = M(dim, nc, 5, 2, 1.0 / nc, coarsen_to=2, coord=coord)
        S.build_tree(T, dim, nc, _rules, n_base=base)
        ts.append(T)
    data = S.fill_random(ts[0], dim, [1, 2, 3, 4, 5], seed=9)
    for (i, iv), v in data.items():
        for T in ts[1:]:
            T.set_cc(i, iv, v)
    kinds = np.array([0, 1, 2, 0, 1, 2], np.int32)
    for T in ts:
        T.mg_setup(1, 2, 3, n_base=4, bc_kind=kinds, bc_val=np.arange(6) + 0.5, auto_ops=True)
    mx = ts[0].info()["max_lvl"]

    def same(ivs):
        for iv in ivs:
            a = S.all_cc(ts[0], dim, iv)
            for T in ts[1:]:
                b = S.all_cc(T, dim, iv)
                for k in a:
                    np.testing.assert_array_equal(a[k], b[k], err_msg="box %d var %d" % (k, iv))

    for T in ts:
        T.fas_fmg(False, True)
    same((1, 2, 3))
    for rep in range(4):
        for T in ts:
            T.fas_vcycle(mx, rep % 2 == 1)
            T.gc_tree(4, 1, 1)  # other work on the same stream between replays
        same((1, 2, 3, 4))
        for T in ts:
            T.fas_fmg(False, False)
        same((1, 2, 3))
    # a topology change invalidates the captured graphs
    for T in ts:
        b = T.boxes()
        c = S.box_centres(b, nc)
        fl = np.array([1 if (b["in_use"][i] and b["lvl"][i] == mx and
                             np.linalg.norm(c[i] - 0.3) < 0.1) else 0
                       for i in range(len(b["lvl"]))], np.int32)
        T.adjust(fl)
    data = S.fill_random(ts[0], dim, [1, 2, 3], seed=10)
    for (i, iv), v in data.items():
        for T in ts[1:]:
            T.set_cc(i, iv, v)
    mx = ts[0].info()["max_lvl"]
    for rep in range(3):
        for T in ts:
            T.fas_vcycle(mx, False)
        same((1, 2, 3))
