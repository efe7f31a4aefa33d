"""Streamer physics parity: the HIP kernels against the oracle (bit-exact),
on seeded 2D, cylindrical and 3D trees with refinement boundaries."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

pytestmark = pytest.mark.gpu

EPS0, QE = 8.8541878176e-12, 1.6022e-19


def _rules(lvl, c):
    if lvl < 2:
        return 1
    return 1 if (lvl < 4 and np.linalg.norm(c - 0.4) < 0.22) else -1


def _table_data():
    """Transport data shaped like the configs' air model: mobility 0.03,
    diffusion 0.2, alpha = 4.3e5 exp(-2e7/E), eta = 0, on 0..3e7 V/m."""
    x = np.linspace(0, 3.2e7, 433)
    alpha = 4.3e5 * np.exp(-2e7 / np.maximum(x, 1.0))
    return x, [np.full_like(x, 0.03), np.full_like(x, 0.2), alpha, np.zeros_like(x)]


def _setup(dim, nc, coord, base, seed):
    import a5_streamer_amd as A
    L = 1e-3
    ts = []
    for M in (O.OTree, A.Tree):
        T = M(dim, nc, 9, 2, L / nc, coarsen_to=2, coord=coord)
        S.build_tree(T, dim, nc, lambda l, c: _rules(l, c / L), n_base=base)
        ts.append(T)
    rng = np.random.default_rng(seed)
    b = ts[0].boxes()
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        n = (b["n_cell"][i] + 2) ** dim
        vals = {1: 1e18 * (1 + rng.random(n)), 2: 1e18 * (1 + rng.random(n)),
                3: 1e18 * rng.random(n), 4: 1e18 * rng.random(n),
                5: 1e3 * rng.standard_normal(n), 6: 1e7 * (0.5 + rng.random(n)),
                7: rng.standard_normal(n), 8: np.zeros(n),
                9: np.where(rng.random(n) < 0.1, 1.5, 1.0)}
        for iv, v in vals.items():
            for T in ts:
                T.set_cc(i + 1, iv, v)
        if b["lvl"][i] >= 1:
            nf = (b["n_cell"][i] + 1) * b["n_cell"][i] ** (dim - 1)
            for d in range(1, dim + 1):
                for f in (1, 2):
                    v = rng.standard_normal(nf) * (1e7 if f == 2 else 1e20)
                    for T in ts:
                        T.set_fc(i + 1, d, f, v)
    x, cols = _table_data()
    tbl = O.OTable(0.0, 3e7, 1000, 4)
    ts[1].table_create(0.0, 3e7, 1000, 4)
    for c, y in enumerate(cols, start=1):
        tbl.set_col(c, x, y)
        ts[1].table_set_col(c, x, y)
    v = O.streamer_vars(dielectric_rule=(coord == 0))
    ts[1].streamer_setup(dielectric_rule=(coord == 0))
    return ts, tbl, v


def _same_cc(ts, dim, ivs, lvl_min=-10):
    b = ts[0].boxes()
    for iv in ivs:
        for i in range(len(b["lvl"])):
            if b["in_use"][i] and b["lvl"][i] >= lvl_min:
                nc = int(b["n_cell"][i])
                np.testing.assert_array_equal(ts[0].get_cc(i + 1, iv, nc), ts[1].get_cc(i + 1, iv, nc),
                                              err_msg="box %d var %d" % (i + 1, iv))


def _same_fc(ts, dim, f, lvl_min=1):
    b = ts[0].boxes()
    for i in range(len(b["lvl"])):
        if b["in_use"][i] and b["lvl"][i] >= lvl_min:
            for d in range(1, dim + 1):
                np.testing.assert_array_equal(ts[0].get_fc(i + 1, d, f), ts[1].get_fc(i + 1, d, f),
                                              err_msg="box %d dim %d face var %d" % (i + 1, d, f))


CASES = [(2, 8, 0, (2, 1, 1)), (2, 8, 1, (1, 2, 1)), (3, 8, 0, (1, 1, 1))]


def test_table_bit_exact(oracle_lib):
    import a5_streamer_amd as A
    T = A.Tree(2, 8, 9, 2, 1.0)
    x, cols = _table_data()
    tbl = O.OTable(0.0, 3e7, 1000, 4)
    T.table_create(0.0, 3e7, 1000, 4)
    for c, y in enumerate(cols, start=1):
        tbl.set_col(c, x, y)
        T.table_set_col(c, x, y)
    a, b = tbl.data(), T.table_get()
    np.testing.assert_array_equal(a[0], b[0])
    assert a[1:] == b[1:]


@pytest.mark.parametrize("dim,nc,coord,base", CASES)
def test_field_and_fluxes(oracle_lib, dim, nc, coord, base):
    ts, tbl, v = _setup(dim, nc, coord, base, 1)
    O.space_charge(ts[0], v, QE / EPS0)
    ts[1].space_charge(QE / EPS0)
    _same_cc(ts, dim, (7,))
    O.fld_from_pot(ts[0], v)
    ts[1].fld_from_pot()
    _same_cc(ts, dim, (6,))
    _same_fc(ts, dim, 2)
    for T in ts:
        T.gc_tree(6, 1, 1)
        T.gc_tree(1, 2, 1)
    _same_cc(ts, dim, (1, 6))
    dt = 3e-12
    O.fluxes_koren(ts[0], v, tbl, dt)
    ts[1].fluxes_koren(dt)
    _same_fc(ts, dim, 1)
    for T in ts:
        T.consistent_fluxes([1])
    _same_fc(ts, dim, 1)


@pytest.mark.parametrize("dim,nc,coord,base", CASES)
def test_update_average_summaries(oracle_lib, dim, nc, coord, base):
    ts, tbl, v = _setup(dim, nc, coord, base, 2)
    dt = 2e-12
    O.update_solution(ts[0], v, tbl, dt)
    ts[1].update_solution(dt)
    _same_cc(ts, dim, (1, 2))
    O.average_dens(ts[0], v)
    ts[1].average_dens()
    _same_cc(ts, dim, (1, 2))
    np.testing.assert_array_equal(O.box_summaries(ts[0], v), ts[1].box_summaries())


@pytest.mark.parametrize("dim,nc,coord,base", CASES)
def test_prolong_new_boxes(oracle_lib, dim, nc, coord, base):
    ts, tbl, v = _setup(dim, nc, coord, base, 3)
    b = ts[0].boxes()
    ids = [i + 1 for i in range(len(b["lvl"])) if b["in_use"][i] and b["lvl"][i] >= 2]
    for i in ids:
        ts[0].prolong1_to(i, 1)
    ts[1].prolong1_to(ids, 1)
    _same_cc(ts, dim, (1,))
    for T in ts:
        T.mg_setup(5, 6, 7, bc_kind=np.array([0, 0, 1, 2, 1, 2][:2 * dim], np.int32),
                   bc_val=np.full(2 * dim, 7.0))
    # level by level, as prolong_to_new_boxes does (src/streamer_2d.f90:598-616)
    for i in sorted(ids, key=lambda i: (b["lvl"][i - 1], i)):
        ts[0].gc_box(i, 1, 2, 1)
        ts[0].gc_box(i, 5, 0, 0)
    ts[1].gc_boxes(ids, 1, 2, 1)
    ts[1].gc_boxes(ids, 5, 0, 0)
    _same_cc(ts, dim, (1, 5))
