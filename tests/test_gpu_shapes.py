"""Launch shapes of the multigrid kernels (a5g_set_shape) and the launch
autotuner (a5g_autotune). Every shape (default; persistent grid; half the
lanes per box and twice the boxes per workgroup; both; for the ghost fill,
residual + restriction, coarse rhs and correction the default and half the
lanes) computes the same doubles: an FMG and a V-cycle with one shape forced everywhere leave every
variable of every box, ghosts included, equal to the default's. The first
and third trees hold levels of more workgroup slots (32,768 boxes of 8^2,
4,096 boxes of 8^3) than a persistent grid has workgroups (2,048 on 256
compute units), so persistent workgroups walk several slots. The autotuner leaves the state
as it found it and reports one chosen shape per (box size, size class)."""
import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu

CASES = [
    # dim, nc, coord, base, uniform level (every box below it refined)
    (2, 8, 1, (1, 2, 1), 8),
    (2, 8, 0, (2, 1, 1), 6),
    (3, 8, 0, (1, 1, 1), 5),
    (2, 16, 1, (1, 2, 1), 6),
    (3, 16, 0, (1, 1, 1), 3),
    (2, 4, 0, (1, 1, 1), 5),
]


def _tree(dim, nc, coord, base, lvl_u, eps):
    import a5_streamer_amd as A
    T = A.Tree(dim, nc, 5, 1, 1.0 / nc, coarsen_to=2, coord=coord)

    def rules(lvl, c):
        if lvl < lvl_u:
            return 1
        return 1 if (lvl < lvl_u + 1 and np.linalg.norm(c - 0.3) < 0.15) else 0
    S.build_tree(T, dim, nc, rules, n_base=base)
    rng = np.random.default_rng(3)
    n, bl = T.info()["max_id"], T.box_len(nc)
    T.set_cc_all(3, rng.standard_normal((n, bl)))
    T.set_cc_all(1, np.zeros((n, bl)))
    T.set_cc_all(5, np.where(rng.random((n, bl)) < 0.05, 2.0, 1.0) if eps else np.ones((n, bl)))
    kinds = np.array([0, 1, 2, 0, 1, 2][:2 * dim], np.int32)
    T.mg_setup(1, 2, 3, i_eps=5 if eps else -1, n_base=4, bc_kind=kinds,
               bc_val=np.arange(2 * dim, dtype=float) + 0.5, auto_ops=True)
    return T


def _force(T, shape):
    for kernel in range(5):
        for nc in (2, 4, 8, 16, 32):
            for cls in range(3):
                T.set_shape(kernel, nc, cls, shape)


def _solve(T, start):
    for iv, v in start.items():
        T.set_cc_all(iv, v)
    T.fas_fmg(True, True)
    T.fas_vcycle(T.info()["max_lvl"], True)
    return {iv: T.get_cc_all(iv) for iv in (1, 2, 3, 4)}


@pytest.mark.parametrize("eps", [False, True])
@pytest.mark.parametrize("dim,nc,coord,base,lvl_u", CASES)
def test_every_shape_gives_the_same_doubles(dim, nc, coord, base, lvl_u, eps):
    print("shapes", dim, nc, coord, base, lvl_u, eps, flush=True)
    T = _tree(dim, nc, coord, base, lvl_u, eps)
    start = {iv: T.get_cc_all(iv) for iv in (1, 2, 3, 4)}
    _force(T, 0)
    ref = _solve(T, start)
    for shape in (1, 2, 3):
        _force(T, shape)
        got = _solve(T, start)
        for iv in ref:
            np.testing.assert_array_equal(got[iv], ref[iv], err_msg="shape %d var %d" % (shape, iv))
    T.close()


def test_autotune_keeps_the_state_and_reports_its_choice():
    print("autotune", flush=True)
    T = _tree(2, 8, 1, (1, 2, 1), 8, False)
    T.fas_fmg(True, True)
    before = {iv: T.get_cc_all(iv) for iv in range(1, 6)}
    rec = T.autotune(reps=3)
    for iv, v in before.items():
        np.testing.assert_array_equal(T.get_cc_all(iv), v, err_msg="var %d" % iv)
    keys = {(r["kernel"], r["nc"], r["size_class"]) for r in rec}
    # the finest level (32,768 boxes) is of class 2; its parents (8,192) of 1
    assert {(0, 8, 2), (1, 8, 2), (2, 8, 2), (3, 8, 1), (4, 8, 2)} <= keys
    for k in keys:
        rows = [r for r in rec if (r["kernel"], r["nc"], r["size_class"]) == k]
        assert sorted(r["shape"] for r in rows) == ([0, 1, 2, 3] if k[0] == 0 else [0, 2])
        chosen = [r for r in rows if r["chosen"]]
        assert len(chosen) == 1 and all(r["us"] > 0 for r in rows)
        assert T.get_shape(*k) == chosen[0]["shape"]
        best = min(r["us"] for r in rows)
        default = [r for r in rows if r["shape"] == 0][0]["us"]
        assert chosen[0]["us"] == best or chosen[0]["shape"] == 0
        assert chosen[0]["shape"] != 0 or best >= 0.97 * default
    # the solve after tuning gives the default shapes' doubles
    start = {iv: before[iv] for iv in (1, 2, 3, 4)}
    tuned = _solve(T, start)
    _force(T, 0)
    ref = _solve(T, start)
    for iv in ref:
        np.testing.assert_array_equal(tuned[iv], ref[iv])
    T.close()
