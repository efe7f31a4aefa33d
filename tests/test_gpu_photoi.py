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
