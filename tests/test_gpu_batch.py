"""Independent streamer runs advanced together in one tree (instances,
a5_streamer_amd.streamer.BatchRun over a5g_set_instances) against each run
stepped alone through the oracle: every box of every run, ghosts included,
bit for bit, with the run's own time, time step and mesh.

The level limit is above the levels the runs reach: at the limit the
reference leaves the leaf lists of the top level stale until a tidy_up
(afivo/src/m_aX_core.f90:497-531), and those lists belong to the whole tree,
so a run in a batch and a run alone refresh them at different times."""
import numpy as np
import pytest

import oracle as O
from a5_streamer_amd import streamer as S

pytestmark = pytest.mark.gpu


def _cfgs():
    """Three runs: one whose mesh stays put, one that settles its initial
    refinement an iteration earlier, and two whose meshes grow every block."""
    base = dict(variant="cyl", domain_len=16e-3, n_base=(1, 2), lvls_max=9)
    return [S.Config(applied_fld=e, seed_width=[w], bg_dens=bg, seed_dens=[sd],
                     seed_r0=[[0.0, z]], seed_r1=[[0.0, z]], **base)
            for e, w, bg, z, sd in ((2.0e6, 0.4e-3, 1e15, 24e-3, 1e20),
                                    (2.0e7, 1.2e-3, 1e13, 20e-3, 5e19),
                                    (1.0e7, 0.25e-3, 1e12, 22e-3, 5e19))]


def _state(T, boxes, sel, r_min):
    out = {}
    for i in np.nonzero(sel)[0]:
        key = (int(boxes["lvl"][i]), tuple(r_min[i]))
        nc = int(boxes["n_cell"][i])
        out[key] = np.concatenate([T.get_cc(int(i) + 1, iv, nc) for iv in
                                   (S.I_ELEC, S.I_PION, S.I_PHI, S.I_FLD, S.I_RHS)])
    return out


def test_batch_matches_separate_runs(oracle_lib):
    import a5_streamer_amd as A
    print("batch of 3 cylindrical runs", flush=True)
    cfgs = _cfgs()
    batch = S.BatchRun(A.Tree, cfgs)
    solo = [S.StreamerRun(O.OStreamTree, c) for c in cfgs]
    batch.initialise()
    for r in solo:
        r.initialise()
    for blk in range(4):
        if blk:
            res = batch.step_block()
            for g, r in enumerate(solo):
                a, rm = r.step_block()
                assert (res[0][g], res[1][g]) == (a, rm), (blk, g)
        b = batch._boxes()
        for g, r in enumerate(solo):
            assert r.T.info()["max_lvl"] < 9
            assert batch.runs[g].time == r.time and batch.runs[g].dt == r.dt, (blk, g)
            ob = r.T.boxes()
            mine = _state(batch.T, b, (b["run"] == g) & (b["in_use"] != 0), b["r_loc"])
            ref = _state(r.T, ob, ob["in_use"] != 0, ob["r_min"])
            assert mine.keys() == ref.keys(), (blk, g)
            for k in ref:
                np.testing.assert_array_equal(mine[k], ref[k], err_msg="block %d run %d box %s"
                                              % (blk, g, k))
    assert list(batch.leaf_cells()) == [r.leaf_cells() for r in solo]


def test_groups_side_by_side_match_groups_in_turn():
    """The field solves of run groups side by side (a5g_fas_fmg_groups) give
    every box, ghosts included, the doubles that the groups solved one after
    the other with only that group active give. The runs reach different
    finest levels, so the grouped path is taken."""
    import a5_streamer_amd as A
    print("batch groups side by side vs in turn", flush=True)
    runs = []
    for side in (True, False):
        r = S.BatchRun(A.Tree, _cfgs())
        r.groups_side_by_side = side
        r.initialise()
        for _ in range(2):
            r.step_block()
        runs.append(r)
    assert len(set(runs[0].inst_lvl.tolist())) > 1
    b = [r.T.boxes() for r in runs]
    assert np.array_equal(b[0]["in_use"], b[1]["in_use"])
    for i in np.nonzero(b[0]["in_use"])[0]:
        for iv in (S.I_ELEC, S.I_PION, S.I_PHI, S.I_FLD, S.I_RHS):
            np.testing.assert_array_equal(runs[0].T.get_cc(int(i) + 1, iv),
                                          runs[1].T.get_cc(int(i) + 1, iv),
                                          err_msg="box %d var %d" % (i + 1, iv))
