"""Pins the oracle (oracle/a5_oracle.c, the C restatement) to the reference
itself: the reference's Afivo/fosito/streamer Fortran compiled from source by
oracle/Makefile into oracle/_ref and driven through oracle/ref_harness.f90.in.
Every comparison is exact: topology, ids and FP64 field values bit for bit."""
import numpy as np
import pytest

import oracle as O
import scenarios as S


def _rules(lvl, c):
    if lvl < 2:
        return 1
    if lvl < 4 and np.linalg.norm(c - 0.3) < 0.25:
        return 1
    return -1


def _rules_b(lvl, c):
    # a second pattern that derefines part of the first
    if lvl < 2:
        return 1
    if lvl < 4 and np.linalg.norm(c - 0.6) < 0.2:
        return 1
    return -1


TREES = [(2, 8, 0, (1, 1, 1)), (2, 8, 1, (1, 2, 1)), (3, 8, 0, (1, 1, 1)),
         (2, 4, 0, (2, 1, 1)), (3, 4, 0, (2, 1, 1)), (2, 6, 0, (3, 1, 1))]


def _both(ref, dim, nc, coord, base, rules=_rules, n_cc=4, n_fc=2):
    ts = []
    for M in (ref.RefTree, O.OTree):
        T = M(dim, nc, n_cc, n_fc, 1.0 / nc, coarsen_to=2 if nc % 4 == 0 else -1,
              coord=coord)
        S.build_tree(T, dim, nc, rules, n_base=base)
        ts.append(T)
    return ts


def _same_topology(ts, lvl_range=None):
    a, b = ts[0].boxes(), ts[1].boxes()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    ia, ib = ts[0].info(), ts[1].info()
    for k in ("max_id", "max_lvl", "min_lvl", "n_storage"):
        assert ia[k] == ib[k], k
    for lvl in range(ia["min_lvl"], ia["max_lvl"] + 2):
        for x, y in zip(ts[0].level(lvl), ts[1].level(lvl)):
            np.testing.assert_array_equal(x, y)


def _same_cc(ts, dim, ivs):
    for iv in ivs:
        a = S.all_cc(ts[0], dim, iv)
        b = S.all_cc(ts[1], dim, iv)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg="box %d var %d" % (k, iv))


@pytest.mark.parametrize("dim,nc,coord,base", TREES)
def test_topology_refine_derefine_tidy(ref, oracle_lib, dim, nc, coord, base):
    ts = _both(ref, dim, nc, coord, base)
    _same_topology(ts)
    # move the refined region: derefinement + refinement in one call
    for _ in range(4):
        out = []
        for T in ts:
            b = T.boxes()
            c = S.box_centres(b, nc)
            fl = np.array([_rules_b(int(b["lvl"][i]), c[i]) if b["in_use"][i] else 0
                           for i in range(len(b["lvl"]))], np.int32)
            out.append(T.adjust(fl))
        assert out[0] == out[1]
        _same_topology(ts)
        for lvl in range(1, ts[0].info()["max_lvl"] + 2):
            for x, y in zip(ts[0].ref_info_lvl(lvl), ts[1].ref_info_lvl(lvl)):
                np.testing.assert_array_equal(x, y)
    for T in ts:
        T.tidy_up(0.9, 0.25, 5, False)
    _same_topology(ts)
    for T in ts:
        T.tidy_up(0.9, 0.25, 5, True)
    _same_topology(ts)


@pytest.mark.parametrize("dim,nc,coord,base", TREES)
def test_fmg_vcycle(ref, oracle_lib, dim, nc, coord, base):
    ts = _both(ref, dim, nc, coord, base)
    data = S.fill_random(ts[0], dim, [1, 2, 3, 4], seed=7)
    for (i, iv), v in data.items():
        ts[1].set_cc(i, iv, v)
    kinds = np.array([2, 0, 1, 2, 0, 1][:2 * dim], np.int32)
    for T in ts:
        T.mg_setup(1, 2, 3, n_down=2, n_up=3, n_base=5, bc_kind=kinds,
                   bc_val=np.linspace(-1, 2, 2 * dim), auto_ops=True)
    for it in range(3):
        for T in ts:
            T.fas_fmg(it != 1, it == 0)
        _same_cc(ts, dim, (1, 2, 3))
    for T in ts:
        T.fas_vcycle(ts[0].info()["max_lvl"] - 1, True)
    _same_cc(ts, dim, (1, 2, 3, 4))


@pytest.mark.parametrize("dim,nc,coord,base", TREES[:4])
@pytest.mark.parametrize("rb,bc", [(0, 0), (0, 2), (1, 1), (2, 1), (1, 2)])
def test_ghost_rules(ref, oracle_lib, dim, nc, coord, base, rb, bc):
    ts = _both(ref, dim, nc, coord, base)
    data = S.fill_random(ts[0], dim, [1, 2, 3, 4], seed=3)
    for (i, iv), v in data.items():
        ts[1].set_cc(i, iv, v)
    for T in ts:
        T.mg_setup(1, 2, 3, bc_kind=np.array([1, 2, 0, 1, 2, 0][:2 * dim], np.int32),
                   bc_val=np.full(2 * dim, -0.75))
        T.gc_tree(2, rb, bc)
    _same_cc(ts, dim, (2,))


@pytest.mark.parametrize("dim,nc,coord,base", TREES)
def test_restrict_prolong_gc2_minmax(ref, oracle_lib, dim, nc, coord, base):
    ts = _both(ref, dim, nc, coord, base)
    data = S.fill_random(ts[0], dim, [1, 2, 3, 4], seed=11)
    for (i, iv), v in data.items():
        ts[1].set_cc(i, iv, v)
    for T in ts:
        T.restrict_tree(1)
        T.tree_copy_cc(1, 2)
    _same_cc(ts, dim, (1, 2))
    b = ts[0].boxes()
    ids = [i + 1 for i in range(len(b["lvl"])) if b["in_use"][i] and b["lvl"][i] >= 2]
    for i in ids[:12]:
        for T in ts:
            T.prolong1_to(i, 3, add=(i % 2 == 0))
            T.prolong0_to(i, 4)
    _same_cc(ts, dim, (3, 4))
    for i in ids[:12]:
        np.testing.assert_array_equal(ts[0].gc2_box(i, 1), ts[1].gc2_box(i, 1))
    assert ts[0].tree_minmax(1) == ts[1].tree_minmax(1)
    assert ts[0].min_dr() == ts[1].min_dr()


@pytest.mark.parametrize("dim,nc,coord,base", TREES[:3] + TREES[4:5])
def test_consistent_fluxes(ref, oracle_lib, dim, nc, coord, base):
    ts = _both(ref, dim, nc, coord, base, n_fc=2)
    rng = np.random.default_rng(5)
    b = ts[0].boxes()
    n = (nc + 1) * nc ** (dim - 1)
    # face arrays of the coarse sub-levels (lvl <= 0) are smaller; fluxes live
    # on levels >= 1
    for i in range(len(b["lvl"])):
        if not b["in_use"][i] or b["lvl"][i] < 1:
            continue
        for d in range(1, dim + 1):
            for f in (1, 2):
                v = rng.standard_normal(n)
                for T in ts:
                    T.set_fc(i + 1, d, f, v)
    for T in ts:
        T.consistent_fluxes([1])
    for i in range(len(b["lvl"])):
        if b["in_use"][i] and b["lvl"][i] >= 1:
            for d in range(1, dim + 1):
                for f in (1, 2):
                    np.testing.assert_array_equal(ts[0].get_fc(i + 1, d, f),
                                                  ts[1].get_fc(i + 1, d, f))


def test_koren_limiter(ref, oracle_lib):
    rng = np.random.default_rng(2)
    vals = list(rng.standard_normal((400, 2))) + [(0, 0), (1, 0), (0, 1), (1, 4),
                                                  (1, 0.4), (2.5, 1), (-1, -4), (1, 2.5)]
    for a, b in vals:
        assert O.koren_mlim(float(a), float(b)) == ref.koren_mlim(float(a), float(b))


def eps_field(T, dim, iv, kind_seed):
    """Permittivity by box position: a dielectric slab (eps 3) whose face
    cuts through boxes (variable-eps boxes), a region of constant eps 2
    (constant-eps boxes) and eps 1 elsewhere; ghosts included, as the
    drivers' set_box_eps writes them."""
    b = T.boxes()
    out = {}
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        nc = int(b["n_cell"][i])
        dr = b["dr"][i]
        ax = [b["r_min"][i][d] + (np.arange(nc + 2) - 0.5) * dr for d in range(dim)]
        g = np.meshgrid(*ax, indexing="ij")
        e = np.ones(g[0].shape)
        e[g[1] < 0.31 + 0.02 * kind_seed] = 3.0
        e[(g[0] > 0.75) & (g[1] > 0.5)] = 2.0
        v = e.ravel(order="F")
        out[i + 1] = v
        T.set_cc(i + 1, iv, v)
    return out


@pytest.mark.parametrize("dim,nc,coord,base", TREES[:4])
def test_fmg_vcycle_variable_eps(ref, oracle_lib, dim, nc, coord, base):
    """mg$D_auto_* with an eps variable: normal, constant-eps and
    variable-eps boxes (m_aX_mg.f90:568-712, lpld :895-1071, clpld
    :1467-1531) in one tree."""
    ts = _both(ref, dim, nc, coord, base, n_cc=5)
    data = S.fill_random(ts[0], dim, [1, 2, 3], seed=17)
    for (i, iv), v in data.items():
        ts[1].set_cc(i, iv, v)
    e = eps_field(ts[0], dim, 5, 0)
    for i, v in e.items():
        ts[1].set_cc(i, 5, v)
    kinds = np.array([0, 0, 1, 2, 0, 2][:2 * dim], np.int32)
    for T in ts:
        T.mg_setup(1, 2, 3, i_eps=5, n_down=2, n_up=2, n_base=4, bc_kind=kinds,
                   bc_val=np.linspace(-1, 2, 2 * dim), auto_ops=True)
    for it in range(3):
        for T in ts:
            T.fas_fmg(it != 1, it == 0)
        _same_cc(ts, dim, (1, 2, 3))
    for T in ts:
        T.fas_vcycle(ts[0].info()["max_lvl"], True)
    _same_cc(ts, dim, (1, 2, 3, 5))
