"""The reverse mode of the multigrid solve (a5g_fas_fmg_adjoint) against its
forward map. mg$D_fas_fmg with first_call false is linear in the potential,
tmp and rhs it starts from and in the boundary values V (every operation of
the FMG and its V-cycles is), so its transpose satisfies the dot-product
identity <w, F(x)> = <F^T w, x> for every x and w. The forward map is the
library's FMG, bit for bit the oracle's and the reference's
(tests/test_gpu_mg.py); x and w are seeded random values on every cell,
ghosts included, and random V. A transpose that drops, doubles or misplaces
any term of any operation breaks the identity far above the bound
(1e-12 of the sum of the absolute products). With a permittivity variable
(eps = 1, a constant 2 or random values in [1, 3], box by box) the forward
solve takes the permittivity forms of the operator, the smoother and the
correction on the tagged boxes (mg$D_box_lpld, gsrb_lpld, corr_lpld), and
so does the transpose; with a Helmholtz term (the photoionization solves'
operator lpl(u) - kappa^2 u) the smoother divides by 2D + kappa^2 h^2."""
import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu

CASES = [
    # dim, nc, coord, base, face rules
    (2, 8, 0, (2, 1, 1), (0, 1, 2, 0)),
    (2, 8, 1, (1, 2, 1), (0, 0, 1, 3)),
    (2, 4, 0, (1, 1, 1), (1, 2, 0, 2)),
    (3, 8, 0, (1, 1, 1), (0, 0, 1, 2, 0, 2)),
    (3, 4, 0, (2, 1, 1), (2, 1, 0, 0, 1, 2)),
]
LP, LT, LR, LW = 5, 6, 7, 8


def _rules(lvl, c):
    if lvl < 2:
        return 1
    return 1 if (lvl < 4 and np.linalg.norm(c - 0.35) < 0.22) else -1


def _masks(T, dim):
    """True on the cells of each box slot that the box has ((nc + 2)^dim of
    box_len(n_cell); coarser boxes fill the start of their slot)."""
    b = T.boxes()
    bl = T.box_len(T.n_cell)
    m = np.zeros((len(b["lvl"]), bl), bool)
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            m[i, :(int(b["n_cell"][i]) + 2) ** dim] = True
    return m


I_EPS = 4


def set_eps(T, dim, rng):
    """eps box by box in turn: 1 (no tag), a constant 2, random in [1, 3]."""
    b = T.boxes()
    n, bl = T.info()["max_id"], T.box_len(T.n_cell)
    e = np.ones((n, bl))
    for i in range(n):
        if b["in_use"][i]:
            m = (int(b["n_cell"][i]) + 2) ** dim
            if i % 3 == 1:
                e[i, :m] = 2.0
            elif i % 3 == 2:
                e[i, :m] = 1.0 + 2.0 * rng.random(m)
    T.set_cc_all(I_EPS, e)


@pytest.mark.parametrize("op", ["lpl", "eps", "helmholtz"])
@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("dim,nc,coord,base,kinds", CASES)
def test_fmg_adjoint_dot_product(dim, nc, coord, base, kinds, seed, op):
    import a5_streamer_amd as A
    print("fmg adjoint", dim, nc, coord, base, kinds, seed, op, flush=True)
    T = A.Tree(dim, nc, 8, 1, 1.0 / nc, coarsen_to=2, coord=coord)
    S.build_tree(T, dim, nc, _rules, n_base=base)
    rng = np.random.default_rng(seed)
    V = rng.standard_normal(2 * dim)
    if op == "eps":
        set_eps(T, dim, rng)
    T.mg_setup(1, 2, 3, i_eps=I_EPS if op == "eps" else -1, n_base=4,
               bc_kind=np.array(kinds, np.int32), bc_val=V, auto_ops=True)
    if op == "helmholtz":
        # lpl(u) - kappa^2 u with kappa^2 h^2 of order one on the finest level
        T.mg_set_helmholtz(float(nc * 8) ** 2)
    n, bl = T.info()["max_id"], T.box_len(nc)
    mask = _masks(T, dim)
    x = [rng.standard_normal((n, bl)) * mask for _ in range(3)]
    w = [rng.standard_normal((n, bl)) * mask for _ in range(3)]
    # forward: F(x) from the state x and the values V
    for iv, v in zip((1, 2, 3), x):
        T.set_cc_all(iv, v)
    T.fas_fmg(False, False)
    y = [T.get_cc_all(iv) * mask for iv in (1, 2, 3)]
    lhs = sum(float(np.sum(a * b)) for a, b in zip(w, y))
    scale = sum(float(np.sum(np.abs(a * b))) for a, b in zip(w, y))
    # reverse: F^T w
    for iv, v in zip((LP, LT, LR), w):
        T.set_cc_all(iv, v)
    T.set_cc_all(LW, np.zeros((n, bl)))
    vbar = T.fas_fmg_adjoint([LP, LT, LR, LW])
    xb = [T.get_cc_all(iv) * mask for iv in (LP, LT, LR)]
    rhs = sum(float(np.sum(a * b)) for a, b in zip(xb, x)) + float(np.dot(vbar, V))
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)
    # V enters where a face has the rule 2V - c, and only there
    for nb in range(2 * dim):
        if kinds[nb] < 2:
            assert vbar[nb] == 0.0
    assert any(vbar[nb] != 0.0 for nb in range(2 * dim) if kinds[nb] >= 2)
    T.close()


def test_fmg_adjoint_refuses_bad_variables():
    import a5_streamer_amd as A
    T = A.Tree(2, 8, 8, 1, 1.0 / 8, coarsen_to=2)
    S.build_tree(T, 2, 8, _rules)
    T.mg_setup(1, 2, 3, bc_kind=np.zeros(4, np.int32), auto_ops=True)
    with pytest.raises(Exception):
        T.fas_fmg_adjoint([LP, LT, LR, 99])
    T.close()


def _xw(T, dim, rng, mask):
    n, bl = T.info()["max_id"], T.box_len(T.n_cell)
    return [rng.standard_normal((n, bl)) * mask for _ in range(3)]


@pytest.mark.parametrize("op", ["lpl", "helmholtz"])
@pytest.mark.parametrize("below", [0, 1])
@pytest.mark.parametrize("dim,nc,coord,base,kinds", CASES)
def test_vcycle_adjoint_dot_product(dim, nc, coord, base, kinds, below, op):
    """a5g_fas_vcycle_adjoint against the forward V-cycle up to the finest
    level or the one below (the same affine map and identity as the FMG's),
    and the reverse of a whole solve (FMG + 2 V-cycles): the V-cycles'
    reverses, last first, then the FMG's."""
    import a5_streamer_amd as A
    print("vcycle adjoint", dim, nc, coord, base, kinds, below, op, flush=True)
    T = A.Tree(dim, nc, 8, 1, 1.0 / nc, coarsen_to=2, coord=coord)
    S.build_tree(T, dim, nc, _rules, n_base=base)
    rng = np.random.default_rng(5 + below)
    V = rng.standard_normal(2 * dim)
    T.mg_setup(1, 2, 3, n_base=4, bc_kind=np.array(kinds, np.int32), bc_val=V, auto_ops=True)
    if op == "helmholtz":
        T.mg_set_helmholtz(float(nc * 8) ** 2)
    n, bl = T.info()["max_id"], T.box_len(nc)
    mask = _masks(T, dim)
    top = T.info()["max_lvl"] - below
    for whole in (False, True):
        x, w = _xw(T, dim, rng, mask), _xw(T, dim, rng, mask)
        for iv, v in zip((1, 2, 3), x):
            T.set_cc_all(iv, v)
        if whole:
            T.fas_fmg(False, False)
            T.fas_vcycle(top, False)
            T.fas_vcycle(top, False)
        else:
            T.fas_vcycle(top, False)
        y = [T.get_cc_all(iv) * mask for iv in (1, 2, 3)]
        lhs = sum(float(np.sum(a * b)) for a, b in zip(w, y))
        scale = sum(float(np.sum(np.abs(a * b))) for a, b in zip(w, y))
        for iv, v in zip((LP, LT, LR), w):
            T.set_cc_all(iv, v)
        T.set_cc_all(LW, np.zeros((n, bl)))
        vbar = T.fas_vcycle_adjoint(top, [LP, LT, LR, LW])
        if whole:
            vbar = vbar + T.fas_vcycle_adjoint(top, [LP, LT, LR, LW])
            vbar = vbar + T.fas_fmg_adjoint([LP, LT, LR, LW])
        xb = [T.get_cc_all(iv) * mask for iv in (LP, LT, LR)]
        rhs = sum(float(np.sum(a * b)) for a, b in zip(xb, x)) + float(np.dot(vbar, V))
        assert abs(lhs - rhs) <= 1e-12 * scale, (whole, lhs, rhs, scale)
    with pytest.raises(Exception):
        T.fas_vcycle_adjoint(T.info()["max_lvl"] + 1, [LP, LT, LR, LW])
    T.close()
