"""The Helmholtz multigrid operator and the Helmholtz photoionization source
on the GPU against the oracle, bit for bit (the oracle's versions are pinned
by the mathematics in test_oracle_helmholtz.py: the reference has no
counterpart)."""
import numpy as np
import pytest

import oracle as O
import scenarios as S
from a5_streamer_amd import streamer as ST

pytestmark = pytest.mark.gpu

from test_gpu_mg import CASES, LARGE, _pair, _same  # noqa: E402


@pytest.mark.parametrize("dim,nc,coord,base", CASES + LARGE[:2])
def test_helmholtz_fmg_vcycle_bit_exact(oracle_lib, dim, nc, coord, base):
    print("helmholtz case", dim, nc, coord, base, flush=True)
    ts = _pair(oracle_lib, dim, nc, coord, base)
    kinds = np.ones(2 * dim, np.int32)
    for T in ts:
        T.mg_setup(1, 2, 3, n_base=4, bc_kind=kinds, auto_ops=True)
        T.mg_set_helmholtz(3.7e2 * nc * nc)
    for it in range(2):
        for T in ts:
            T.fas_fmg(True, it == 0)
        _same(ts, dim, (1, 2, 3))
    max_lvl = ts[0].info()["max_lvl"]
    for T in ts:
        T.fas_vcycle(max_lvl, True)
    _same(ts, dim, (1, 2, 3, 4))
    # back to the Laplacian
    for T in ts:
        T.mg_set_helmholtz(0.0)
        T.fas_vcycle(max_lvl, False)
    _same(ts, dim, (1, 2, 3))


@pytest.mark.parametrize("variant,kw", [
    ("2d", dict(lvls_max=6)),
    ("cyl", dict(domain_len=16e-3, n_base=(1, 2), lvls_max=6,
                 seed_r0=[[0.0, 24e-3]], seed_r1=[[0.0, 24e-3]], applied_fld=2.0e6)),
    ("3d", dict(domain_len=8e-3, n_base=2, lvls_max=3, seed_r0=[[4e-3, 4e-3, 5e-3]],
                seed_r1=[[4e-3, 4e-3, 6e-3]], seed_width=[0.6e-3], n_boxes=2000)),
])
def test_photoi_source_bit_exact(oracle_lib, variant, kw):
    import a5_streamer_amd as A
    print("photoi case", variant, flush=True)
    runs = [ST.StreamerRun(M, ST.Config(variant=variant, photoi=True, **kw))
            for M in (O.OStreamTree, A.Tree)]
    for r in runs:
        r.initialise()
    cyc = [r.set_photoionization() for r in runs]
    assert cyc[0] == cyc[1]
    b = runs[0].T.boxes()
    for i in range(len(b["lvl"])):
        if not b["in_use"][i] or b["children"][i][0] != 0:
            continue
        nc = int(b["n_cell"][i])
        a = runs[0].T.get_cc(i + 1, ST.I_PHO, nc)
        g = runs[1].T.get_cc(i + 1, ST.I_PHO, nc)
        m = np.zeros([nc + 2] * runs[0].dim, bool)
        m[(slice(1, -1),) * runs[0].dim] = True
        m = m.ravel(order="F")
        np.testing.assert_array_equal(a[m], g[m], err_msg="box %d" % (i + 1))
    assert max(np.abs(runs[1].T.get_cc(i + 1, ST.I_PHO)).max()
               for i in range(len(b["lvl"])) if b["in_use"][i]) > 0


def _leaf_interiors(T, dim, iv):
    b = T.boxes()
    out = []
    for i in range(len(b["lvl"])):
        if b["in_use"][i] and b["children"][i][0] == 0:
            nc = int(b["n_cell"][i])
            a = T.get_cc(i + 1, iv, nc).reshape([nc + 2] * dim, order="F")
            out.append(a[(slice(1, -1),) * dim].ravel())
    return np.concatenate(out)


@pytest.mark.parametrize("variant,kw", [
    ("2d", dict(lvls_max=6)),
    ("3d", dict(domain_len=8e-3, n_base=2, lvls_max=3, seed_r0=[[4e-3, 4e-3, 5e-3]],
                seed_r1=[[4e-3, 4e-3, 6e-3]], seed_width=[0.6e-3], n_boxes=2000)),
])
def test_photoi_terms_warm_and_cold(oracle_lib, variant, kw):
    """The terms solved side by side. A warm call on an unchanged source
    needs at most one V-cycle per term (restricting the stored solution to
    the parents moves the refinement-boundary ghosts a little) and matches
    the oracle's repeated call bit for bit. After a step block, a warm
    solve, a cold one (FMG from zero) and a cold one to 1e-12 each match the
    oracle's bit for bit, and the warm start is as accurate as the cold one
    (check_warm_as_accurate)."""
    import a5_streamer_amd as A
    from test_oracle_helmholtz import _warm_cold_tight, check_warm_as_accurate
    print("photoi terms case", variant, flush=True)
    runs = [ST.StreamerRun(M, ST.Config(variant=variant, photoi=True, **kw))
            for M in (O.OStreamTree, A.Tree)]
    dim = runs[0].dim
    for r in runs:
        r.initialise()
    cyc = [r.set_photoionization() for r in runs]
    assert cyc[0] == cyc[1] <= 3, cyc
    np.testing.assert_array_equal(_leaf_interiors(runs[0].T, dim, ST.I_PHO),
                                  _leaf_interiors(runs[1].T, dim, ST.I_PHO))
    for r in runs:
        r.step_block()
    o, g = [_warm_cold_tight(r, dim) for r in runs]
    for k in range(3):
        for j in range(3):
            np.testing.assert_array_equal(o[k][j], g[k][j], err_msg="solve %d term %d" % (k, j))
    check_warm_as_accurate(*g)


def test_photoi_terms_argument_errors():
    import a5_streamer_amd as A
    r = ST.StreamerRun(A.Tree, ST.Config(variant="2d", photoi=True, lvls_max=4))
    r.initialise()
    coeff, amp, kappa = r.c.photoi_model()
    amp, kappa = np.array(amp), np.array(kappa)
    with pytest.raises(A.A5GError, match="must all differ"):
        r.T.photoi_helmholtz_terms(coeff, amp, kappa, (10, 11, 12), (13, 14, 13),
                                   (16, 17, 18), False, 1e-6, 40)
    with pytest.raises(A.A5GError, match="must all differ"):
        r.T.photoi_helmholtz_terms(coeff, amp, kappa, (10, 11, ST.I_PHO), (13, 14, 15),
                                   (16, 17, 18), False, 1e-6, 40)
    with pytest.raises(A.A5GError, match="1 to 4 terms"):
        r.T.photoi_helmholtz_terms(coeff, np.ones(5), np.ones(5), range(10, 15),
                                   range(15, 20), range(20, 25), False, 1e-6, 40)
