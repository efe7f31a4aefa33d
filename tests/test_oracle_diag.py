"""Point location, located reductions and the axis walk of the streamer
diagnostics (get_streamer_properties / get_cc_axis, src/streamer_cyl.f90:
682-956): the C restatement against the reference's own a$D_get_loc and
a$D_reduction_loc (afivo/src/m_aX_utils.f90:244-306,499-555) built from
source, and against brute force over the leaf cells."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

NC = 8


def _rule(dim):
    # refine around an off-centre point, twice, plus one level everywhere
    def rule(lvl, c):
        if lvl < 2:
            return 1
        if lvl < 4 and np.linalg.norm(c - 0.3) < 0.25 / lvl:
            return 1
        return 0
    return rule


def _trees(dim, ref):
    out = []
    for T in (O.OTree(dim, NC, 3, 2, 1.0 / NC, coarsen_to=2, n_boxes=50),
              ref.RefTree(dim, NC, 3, 2, 1.0 / NC, coarsen_to=2, n_boxes=50) if ref else None):
        if T is not None:
            S.build_tree(T, dim, NC, _rule(dim))
        out.append(T)
    return out


def _fill(trees, dim, seed, ties=False):
    rng = np.random.default_rng(seed)
    b = trees[0].boxes()
    nfa = (NC + 1) * NC ** (dim - 1)
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        cc = [rng.standard_normal((NC + 2) ** dim) for _ in range(3)]
        fc = [[rng.standard_normal(nfa) for _ in range(dim)] for _ in range(2)]
        if ties:
            cc = [np.round(c) for c in cc]
            fc = [[np.round(f) for f in g] for g in fc]
        for T in trees:
            if T is None:
                continue
            for iv in range(3):
                T.set_cc(i + 1, iv + 1, cc[iv])
            for ifc in range(2):
                for d in range(dim):
                    T.set_fc(i + 1, d + 1, ifc + 1, fc[ifc][d])


def _leaf_cells(T, dim):
    """(id, ix, centre) of every leaf cell in the reference's order: levels
    upwards, leaves in list order, cells x fastest."""
    b = T.boxes()
    out = []
    for lvl in range(1, T.info()["max_lvl"] + 1):
        for i in T.level(lvl)[1]:
            dr, rmin = b["dr"][i - 1], b["r_min"][i - 1]
            for p in np.ndindex(*([NC] * dim)):
                ix = np.array(p[::-1]) + 1
                out.append((int(i), ix, rmin + (ix - 0.5) * dr))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_get_loc_matches_reference(oracle_lib, ref, dim):
    T, R = _trees(dim, ref)
    rng = np.random.default_rng(dim)
    pts = list(rng.uniform(-0.05, 1.05, (400, dim)))
    # points on box faces and corners: the first base box in list order that
    # holds the point wins, and a point on a child split goes to the low child
    pts += [np.full(dim, v) for v in (0.0, 0.25, 0.5, 1.0)]
    pts += [np.array([0.5, 0.125, 0.75][:dim]), np.array([1.0, 0.0, 0.5][:dim])]
    for r in pts:
        for m in (0, 1, 2, 3):
            a = T.get_loc(r, m)
            b = R.get_loc(r, m)
            assert a[0] == b[0] and np.array_equal(a[1], b[1]), (r, m, a, b)
    # the cell found contains the point (a$D_cc_ix, afivo/src/m_aX_utils.f90:301-306)
    bx = T.boxes()
    for r in rng.uniform(0.01, 0.99, (100, dim)):
        i, ix = T.get_loc(r)
        lo = bx["r_min"][i - 1] + (ix - 1) * bx["dr"][i - 1]
        assert np.all(lo <= r + 1e-15) and np.all(r <= lo + bx["dr"][i - 1] + 1e-15)
        assert all(bx["children"][i - 1] == 0)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("ties", [False, True])
def test_reduction_loc_matches_reference(oracle_lib, ref, dim, ties):
    """Both per-box rules along every dimension, on random data and on data
    rounded to integers (many ties: the first leaf in the reference's
    single-thread order and the first cell in array order win)."""
    trees = _trees(dim, ref)
    _fill(trees, dim, 5 + dim, ties)
    T, R = trees
    ref.set_threads(1)
    try:
        for d in range(1, dim + 1):
            for kind, var, thr, init in ((1, 1, 0.0, -1e99), (1, 2, 0.0, -1e99),
                                         (2, 2, 0.5, 0.0), (2, 3, 3.5, 0.0),
                                         (2, 1, 1e9, 0.0)):
                a = T.reduction_loc(kind, var, d, thr, init)
                b = R.reduction_loc(kind, var, d, thr, init)
                assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]), \
                    (kind, var, d, thr, a, b)
    finally:
        ref.set_threads(8)


@pytest.mark.parametrize("dim", [2, 3])
def test_reduction_loc_brute_force(oracle_lib, dim):
    T, _ = _trees(dim, None)
    _fill([T], dim, 21)
    cells = _leaf_cells(T, dim)
    fcs = {}
    for d in range(1, dim + 1):
        # kind 1: the largest 0.5 |f(ix) + f(ix + e_d)| over leaf cells
        best, loc = -np.inf, None
        for i, ix, _ in cells:
            key = (i, d)
            if key not in fcs:
                shape = [NC] * dim
                shape[d - 1] += 1
                fcs[key] = T.get_fc(i, d, 1).reshape(shape, order="F")
            f = fcs[key]
            hi = ix.copy()
            hi[d - 1] += 1
            v = 0.5 * abs(f[tuple(ix - 1)] + f[tuple(hi - 1)])
            if v > best:
                best, loc = v, (i, ix)
        val, i, ix = T.reduction_loc(1, 1, d)
        assert val == best and i == loc[0] and np.array_equal(ix, loc[1])
        # kind 2: the largest centre coordinate d among cells above 1.0
        best, loc = 0.0, (-1, None)
        for i, ix, rc in cells:
            v = T.get_cc(i, 2).reshape([NC + 2] * dim, order="F")[tuple(ix)]
            if v > 1.0 and rc[d - 1] > best:
                best, loc = rc[d - 1], (i, ix)
        val, i, ix = T.reduction_loc(2, 2, d, 1.0, 0.0)
        assert val == pytest.approx(best, abs=1e-15) and i == loc[0]
        assert np.array_equal(ix, loc[1])
    # nothing above the threshold: init value, no location
    val, i, ix = T.reduction_loc(2, 2, 1, 1e9, 0.0)
    assert val == 0.0 and i == -1 and np.all(ix == -1)


def test_cc_axis_walk(oracle_lib, ref):
    """The axis holds the first column of every leaf box touching x = 0,
    bottom to top; each point is in the box the reference's a2_get_loc finds
    there, whose x index is ceiling(0 / dr) = 0 (get_cc_axis reads column 1
    itself)."""
    T, R = _trees(2, ref)
    _fill([T, R], 2, 9)
    ax = T.cc_axis([1, 2], [1, 2])
    assert ax.shape[0] == 5
    z = ax[0]
    assert np.all(np.diff(z) > 0) and 0 < z[0] < 0.5 / NC and z[-1] < 1.0
    b = T.boxes()
    on_axis = [i + 1 for i in range(len(b["lvl"])) if b["in_use"][i]
               and b["children"][i][0] == 0 and b["r_min"][i][0] == 0.0]
    assert len(z) == NC * len(on_axis)
    for p in range(len(z)):
        i, ix = R.get_loc([0.0, z[p]])
        assert ix[0] == 0
        j, jx = T.get_loc([0.0, z[p]])
        assert j == i and np.array_equal(jx, ix)
        cc = [T.get_cc(i, v).reshape(NC + 2, NC + 2, order="F")[1, ix[1]] for v in (1, 2)]
        fy = [T.get_fc(i, 2, v).reshape(NC, NC + 1, order="F") for v in (1, 2)]
        fa = [0.5 * (f[0, ix[1] - 1] + f[0, ix[1]]) for f in fy]
        assert list(ax[1:, p]) == cc + fa
        assert z[p] == b["r_min"][i - 1][1] + (ix[1] - 0.5) * b["dr"][i - 1]


@pytest.mark.parametrize("dim", [2, 3])
def test_library_get_loc_matches_oracle(oracle_lib, dim):
    """The library's host tree locates points like the oracle (no GPU call)."""
    import a5_streamer_amd as A
    T, _ = _trees(dim, None)
    L = A.Tree(dim, NC, 3, 2, 1.0 / NC, coarsen_to=2, n_boxes=50)
    S.build_tree(L, dim, NC, _rule(dim))
    rng = np.random.default_rng(40 + dim)
    pts = list(rng.uniform(-0.05, 1.05, (300, dim))) + [np.full(dim, v) for v in
                                                        (0.0, 0.25, 0.5, 1.0)]
    for r in pts:
        for m in (0, 2):
            a, b = T.get_loc(r, m), L.get_loc(r, m)
            assert a[0] == b[0] and np.array_equal(a[1], b[1]), (r, m)
