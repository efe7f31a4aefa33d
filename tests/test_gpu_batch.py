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


def _batch_state(T, boxes, run):
    """_state of one run of a batch tree, read with one get_cc_all per
    variable."""
    allv = {iv: T.get_cc_all(iv) for iv in (S.I_ELEC, S.I_PION, S.I_PHI, S.I_FLD, S.I_RHS)}
    out = {}
    for i in np.nonzero((boxes["run"] == run) & (boxes["in_use"] != 0))[0]:
        m = (int(boxes["n_cell"][i]) + 2) ** 2
        out[(int(boxes["lvl"][i]), tuple(boxes["r_loc"][i]))] = np.concatenate(
            [allv[iv][i, :m] for iv in (S.I_ELEC, S.I_PION, S.I_PHI, S.I_FLD, S.I_RHS)])
    return out


@pytest.mark.parametrize("developed", [False, True])
def test_batch_draws_match_separate_runs(oracle_lib, developed):
    """16 seeded draws of the batched workload (field 1.5-2.5 MV/m, seed
    width 0.2-0.8 mm, background 1e14-1e16 m^-3), batched, against each run
    stepped alone through the oracle, bit for bit on every box and ghost.
    developed: every run refined to the developed streamer mesh of the
    benchmark (about 1,100 boxes a run) before it starts."""
    import a5_streamer_amd as A
    from a5_streamer_amd.workloads import BATCH_REGIONS, batch_configs
    print("batch of 16 seeded draws, developed=%s" % developed, flush=True)
    lvls_max = 12
    cfgs = batch_configs(16, lvls_max, BATCH_REGIONS if developed else None)
    batch = S.BatchRun(A.Tree, cfgs)
    solo = [S.StreamerRun(O.OStreamTree, c) for c in cfgs]
    batch.initialise()
    for r in solo:
        r.initialise()
    n_blocks = 2 if developed else 5
    for blk in range(n_blocks + 1):
        if blk:
            res = batch.step_block()
            for g, r in enumerate(solo):
                a, rm = r.step_block()
                assert (res[0][g], res[1][g]) == (a, rm), (blk, g)
        if blk not in (0, n_blocks):
            continue
        b = batch._boxes()
        for g, r in enumerate(solo):
            assert r.T.info()["max_lvl"] < lvls_max
            assert batch.runs[g].time == r.time and batch.runs[g].dt == r.dt, (blk, g)
            ob = r.T.boxes()
            mine = _batch_state(batch.T, b, g)
            ref = _state(r.T, ob, ob["in_use"] != 0, ob["r_min"])
            assert mine.keys() == ref.keys(), (blk, g)
            for k in ref:
                np.testing.assert_array_equal(mine[k], ref[k], err_msg="block %d run %d box %s"
                                              % (blk, g, k))
    assert list(batch.leaf_cells()) == [r.leaf_cells() for r in solo]
    if developed:
        per_run = np.bincount(b["run"][b["in_use"] != 0], minlength=16)
        assert per_run.min() >= 200 and per_run.max() <= 3000, per_run
    batch.T.close()


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


def test_batch_init_on_device_matches_host_init():
    """set_init_cond of every run on the device (a5g_set_gaussians_instances)
    gives the host's densities within rounding (the device sums the squares
    of the distance plainly, the host as the reference does: 1e-12 relative)
    and the same eps and cleared potential, on the mesh after the initial
    refinement."""
    import a5_streamer_amd as A
    print("batch device initial condition", flush=True)
    B = S.BatchRun(A.Tree, _cfgs())
    B.initialise()
    b = B._boxes()
    use = np.nonzero(b["in_use"])[0]
    ivs = (S.I_ELEC, S.I_PION, S.I_EPS)
    host = {iv: B.T.get_cc_all(iv)[use] for iv in ivs}
    for c in B.cfgs:
        c.init_on_device = True
    B.T.fill_cc((use + 1).astype(np.int32), S.I_PHI, 7.0)
    for iv in ivs:
        B.T.fill_cc((use + 1).astype(np.int32), iv, -1.0)
    B._set_active(np.ones(B.K, bool))
    B.set_init_cond(b)
    bl = B.T.box_len(B.T.n_cell)
    for iv in ivs:
        dev = B.T.get_cc_all(iv)[use]
        for r, i in enumerate(use):
            m = (int(b["n_cell"][i]) + 2) ** 2
            np.testing.assert_allclose(dev[r, :m], host[iv][r, :m], rtol=1e-12, atol=0,
                                       err_msg="var %d box %d" % (iv, i + 1))
        assert dev.shape[1] == bl
    phi = B.T.get_cc_all(S.I_PHI)[use]
    for r, i in enumerate(use):
        m = (int(b["n_cell"][i]) + 2) ** 2
        assert (phi[r, :m] == 0.0).all()
