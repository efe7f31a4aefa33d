"""Host side of the batched runs (a5_streamer_amd.streamer.BatchRun): the
array forms of norm2, dist_line, the table lookup and the refinement flags
give the same doubles and flags as the per-box forms the single run uses."""
import numpy as np
import pytest

import oracle as O
from a5_streamer_amd import streamer as S


def test_norm2_rows_is_norm2():
    rng = np.random.default_rng(1)
    v = rng.standard_normal((4000, 3)) * 10.0 ** rng.integers(-20, 20, (4000, 3))
    v[rng.random(v.shape) < 0.15] = 0.0
    v[:50, 1] = v[:50, 0]           # ties
    v[50:60] = 0.0                  # all zero
    got = S.norm2_rows(v)
    for i in range(len(v)):
        assert got[i] == S.norm2(list(v[i]))


def test_dist_line_rows_is_dist_line():
    rng = np.random.default_rng(2)
    r = rng.uniform(-1, 1, (3000, 2))
    for r0, r1 in (([0.1, 0.2], [0.1, 0.2]), ([0.0, -0.5], [0.3, 0.7]), ([0.5, 0.5], [-0.5, 0.5])):
        got = S.dist_line_rows(r, r0, r1)
        for i in range(len(r)):
            assert got[i] == S.dist_line(list(r[i]), r0, r1)


def test_table_col_rows_is_col(oracle_lib):
    cfg = S.Config()
    x, cols = cfg.transport_columns()
    t = O.OTable(0.0, 3e7, 1000, 4)
    for k, y in enumerate(cols, start=1):
        t.set_col(k, x, y)
    tb = S.Table(*t.data())
    xs = np.concatenate([np.random.default_rng(3).uniform(-1e6, 3.2e7, 5000), [0.0, 3e7, 4e7]])
    got = tb.col_rows(S.T_ALPHA, xs)
    for i, xv in enumerate(xs):
        assert got[i] == tb.col(S.T_ALPHA, xv)


def test_batch_flags_are_the_box_flags(oracle_lib):
    """BatchRun.flags on synthetic boxes of three runs equals box_flag of
    each box in its own run; inactive runs and boxes below level 1 keep."""
    cfgs = [S.Config(variant="cyl", domain_len=16e-3, n_base=(1, 2), applied_fld=e,
                     seed_width=[w], seed_r0=[[0.0, z]], seed_r1=[[0.0, z + dz]])
            for e, w, z, dz in ((2e6, 0.4e-3, 24e-3, 0.0), (2.2e6, 0.7e-3, 20e-3, 2e-3),
                                (1.7e6, 0.25e-3, 26e-3, 0.0))]
    B = S.BatchRun.__new__(S.BatchRun)
    B.cfgs, B.v, B.dim, B.K = cfgs, S.VARIANTS["cyl"], 2, 3
    t = O.OTable(0.0, 3e7, 1000, 4)
    x, cols = cfgs[0].transport_columns()
    for k, y in enumerate(cols, start=1):
        t.set_col(k, x, y)
    tb = S.Table(*t.data())
    B.runs = []
    for i, c in enumerate(cfgs):
        r = S.StreamerRun.__new__(S.StreamerRun)
        r.c, r.v, r.dim, r.tbl = c, B.v, 2, tb
        r.time = (1e-9, 6e-9, 2e-9)[i]
        B.runs.append(r)
    B.active = np.array([True, True, False])
    rng = np.random.default_rng(4)
    n = 3000
    lvl = rng.integers(-1, 8, n)
    dr = 2e-3 * 0.5 ** (lvl - 1)
    b = {"run": rng.integers(0, 3, n), "in_use": (rng.random(n) < 0.9).astype(np.int32),
         "lvl": lvl, "dr": dr, "n_cell": np.full(n, 8),
         "r_loc": np.stack([rng.uniform(0, 16e-3, n), rng.uniform(0, 32e-3, n)], axis=1)}
    s = np.stack([rng.uniform(0, 8, n), 10 ** rng.uniform(5, 7.5, n), np.zeros(n)], axis=1)
    got = B.flags(b, s)
    for i in range(n):
        g = b["run"][i]
        want = 0
        if b["in_use"][i] and lvl[i] >= 1 and B.active[g]:
            want = B.runs[g].box_flag(B.runs[g].time, s[i], dr[i], 8, b["r_loc"][i])
        assert got[i] == want, i
    assert set(got.tolist()) == {-1, 0, 1}
    # refine regions (two runs share theirs, one has none): each run's
    # flags_rows on its own boxes
    before = got
    regions = [(5, [0.0, 10e-3], [8e-3, 25e-3]), (7, [0.0, 18e-3], [4e-3, 25e-3])]
    cfgs[0].refine_regions = cfgs[1].refine_regions = regions
    got = B.flags(b, s)
    ok = (b["in_use"] != 0) & (lvl >= 1)
    for g in range(3):
        mine = b["run"] == g
        want = B.runs[g].flags_rows(B.runs[g].time, s[mine], dr[mine], b["n_cell"][mine],
                                    b["r_loc"][mine], ok[mine] & B.active[g], lvl[mine])
        np.testing.assert_array_equal(got[mine], want, err_msg="run %d" % g)
    assert (got[b["run"] == 0] != before[b["run"] == 0]).any()
    np.testing.assert_array_equal(got[b["run"] == 2], before[b["run"] == 2])


def test_dist_line_rows_per_is_dist_line():
    rng = np.random.default_rng(6)
    n = 2000
    r = rng.uniform(-1, 1, (n, 2))
    r0 = rng.uniform(-1, 1, (n, 2))
    r1 = np.where(rng.random((n, 1)) < 0.3, r0, rng.uniform(-1, 1, (n, 2)))
    got = S.dist_line_rows_per(r, r0, r1)
    for i in range(n):
        assert got[i] == S.dist_line(list(r[i]), list(r0[i]), list(r1[i]))
