"""The refinement rule and the time-step maxima on the device
(a5g_ref_flags, a5g_instance_maxima) against the host's array forms of the
same rule (StreamerRun.flags_rows, BatchRun.flags and max_dt's maxima,
themselves checked against the per-box rule in tests/test_batch_host.py and,
through whole runs, against the reference drivers): every flag and every
maximum equal, on the states of single runs (2D, cylindrical, 3D, a run whose
mesh grows and shrinks, the production 3D set-up) and of batches of the
seeded workload, inside and after the seed rule's time window."""
import numpy as np
import pytest

from a5_streamer_amd import streamer as S

pytestmark = pytest.mark.gpu

SINGLE = {
    "2d": dict(variant="2d", lvls_max=7),
    "2d_refining": dict(variant="2d", lvls_max=8, applied_fld=1.0e7, seed_width=[0.5e-3],
                        seed_dens=[5e19], bg_dens=1e12),
    "cyl": dict(variant="cyl", domain_len=16e-3, n_base=(1, 2), lvls_max=7,
                seed_r0=[[0.0, 24e-3]], seed_r1=[[0.0, 24e-3]], applied_fld=2.0e6),
    "3d_line": dict(variant="3d", domain_len=8e-3, n_base=2, lvls_max=4,
                    seed_r0=[[4e-3, 4e-3, 5e-3]], seed_r1=[[4e-3, 4e-3, 6e-3]],
                    seed_width=[0.6e-3], applied_fld=2.5e6, n_boxes=4000),
    "3d_regions": dict(variant="3d", domain_len=16e-3, n_base=1, lvls_max=6,
                       seed_r0=[[8e-3, 8e-3, 12e-3]], seed_r1=[[8e-3, 8e-3, 12e-3]],
                       applied_fld=2.0e6, n_boxes=8000,
                       refine_regions=[(4, [0.0, 0.0, 0.0], [16e-3] * 3),
                                       (5, [4e-3, 4e-3, 10e-3], [12e-3, 12e-3, 16e-3])]),
}


def _host_flags(r):
    T = r.T
    b = T.boxes()
    s = T.box_summaries()
    ok = (b["in_use"] != 0) & (b["lvl"] >= 1)
    return r.flags_rows(r.time, s, b["dr"], b["n_cell"], b["r_min"], ok, b["lvl"])


@pytest.mark.parametrize("name", sorted(SINGLE))
def test_single_run_flags(name):
    import a5_streamer_amd as A
    print("device flags, single run", name, flush=True)
    r = S.StreamerRun(A.Tree, S.Config(**SINGLE[name]))
    assert r.device_flags
    r.initialise()
    n_flag = 0
    for blk in range(3):
        saved = r.time
        for t in (saved, saved + 1.0):        # inside and after the seed window
            r.time = t
            host = _host_flags(r)
            dev = r.T.ref_flags([t])
            np.testing.assert_array_equal(dev, host, err_msg="block %d time %g" % (blk, t))
            n_flag += int((dev != 0).sum())
        r.time = saved
        r.step_block()
    assert n_flag > 0
    r.T.close()


@pytest.mark.parametrize("developed", [False, True])
def test_batch_flags_and_maxima(developed):
    import a5_streamer_amd as A
    from a5_streamer_amd.workloads import BATCH_REGIONS, batch_configs
    print("device flags, batch developed=%s" % developed, flush=True)
    B = S.BatchRun(A.Tree, batch_configs(16, 12, BATCH_REGIONS if developed else None))
    assert B.device_flags
    B.initialise()
    for blk in range(3):
        b = B._boxes()
        # some runs inside the seed window, some after it
        saved = [r.time for r in B.runs]
        for g, r in enumerate(B.runs):
            if g % 3 == 0:
                r.time = r.time + 1.0
        host = B.flags(b, B.T.box_summaries())
        dev = B.T.ref_flags([r.time for r in B.runs])
        for r, t in zip(B.runs, saved):
            r.time = t
        np.testing.assert_array_equal(dev, host, err_msg="block %d" % blk)
        # the time-step maxima: device reduction against the host's over the
        # per-box maxima
        fld_d, lvl_d = B.T.instance_maxima(S.I_FLD, B.K)
        s = B.T.box_summaries()
        sel = (b["in_use"] != 0) & (b["lvl"] >= 1)
        leaf = sel & (b["children"][:, 0] == 0)
        fld_h = S.group_max(b["run"][leaf], s[leaf, 1], B.K, -np.inf)
        lvl_h = S.group_max(b["run"][sel], b["lvl"][sel].astype(np.int64), B.K, 0)
        np.testing.assert_array_equal(fld_d, fld_h)
        np.testing.assert_array_equal(lvl_d, lvl_h)
        B.step_block()
    # a subset of runs active: the others get no flags
    act = np.arange(B.K) % 2 == 0
    B._set_active(act)
    b = B._boxes()
    host = B.flags(b, B.T.box_summaries())
    dev = B.T.ref_flags([r.time for r in B.runs])
    np.testing.assert_array_equal(dev, host)
    assert not dev[~act[b["run"]]].any()
    B._set_active(np.ones(B.K, bool))
    B.T.close()
