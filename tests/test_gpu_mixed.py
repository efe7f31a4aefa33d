"""Mixed precision (FP32 multigrid inside FP64 defect correction): solved to
a relative residual of 1e-10 it must agree with the oracle's FP64 solve to
the same residual within 1e-8 relative potential error (max norm over leaf
cells), in 2D, cylindrical and 3D, with refinement boundaries, all three
boundary rules and, in the second test, a dielectric. The GPU's own FP64
solve is checked bit for bit against the oracle's on the way."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

pytestmark = pytest.mark.gpu

TOL_SOLVE = 1e-10
TOL_PHI = 1e-8


def _rules(lvl, c):
    if lvl < 3:
        return 1
    return 1 if (lvl < 5 and np.linalg.norm(c - 0.45) < 0.2) else 0


def _leaf_phi(T, dim, iv=1):
    out = []
    info = T.info()
    for lvl in range(1, info["max_lvl"] + 1):
        for i in T.level(lvl)[1]:
            v = T.get_cc(int(i), iv).reshape([T.n_cell + 2] * dim, order="F")
            out.append(v[(slice(1, -1),) * dim].ravel())
    return np.concatenate(out)


def _fp64_solve(T, max_cycles):
    """FMG, then V-cycles until max|r| <= TOL_SOLVE max|rhs| over the leaves."""
    T.fas_fmg(False, True)
    mx = T.info()["max_lvl"]
    rhs_max = np.abs(_leaf_phi(T, T.dim, 3)).max()
    for _ in range(max_cycles):
        r = T.max_residual() if isinstance(T, O.OTree) else T.residual_norms()[0]
        if r <= TOL_SOLVE * rhs_max:
            return
        T.fas_vcycle(mx, False)
    raise AssertionError("FP64 solve did not reach the tolerance")


def _case(dim, nc, coord, eps, max_cycles, n_inner, max_outer):
    import a5_streamer_amd as A
    from test_oracle_ref import eps_field
    kinds = np.array([0, 0, 1, 2, 1, 2][:2 * dim], np.int32)
    vals = np.array([0, 0, 0, 3.0, 0, 3.0][:2 * dim])
    n_cc = 4 if eps else 3
    blobs = [[0.4, 0.6, 0.5][:dim] + [0.08, 40.0], [0.7, 0.3, 0.4][:dim] + [0.05, -60.0]]
    trees = {}
    for mode in ("oracle", "fp64", "mixed"):
        M = O.OTree if mode == "oracle" else A.Tree
        T = M(dim, nc, n_cc, 0, 1.0 / (2 * nc), coarsen_to=2, coord=coord)
        T.dim = dim
        S.build_tree(T, dim, nc, _rules, n_base=(2, 2, 2))
        if eps:
            eps_field(T, dim, 4, 0)
        T.mg_setup(1, 2, 3, i_eps=4 if eps else -1, n_base=8, bc_kind=kinds, bc_val=vals,
                   auto_ops=True)
        trees[mode] = T
    # the rhs the GPU evaluates, in the oracle too
    trees["fp64"].set_gaussians(3, blobs, 0.0)
    trees["mixed"].set_gaussians(3, blobs, 0.0)
    rhs = trees["fp64"].get_cc_all(3)
    b = trees["oracle"].boxes()
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            trees["oracle"].set_cc(i + 1, 3, rhs[i][: (int(b["n_cell"][i]) + 2) ** dim])
    _fp64_solve(trees["oracle"], max_cycles)
    _fp64_solve(trees["fp64"], max_cycles)
    ref = _leaf_phi(trees["oracle"], dim)
    np.testing.assert_array_equal(_leaf_phi(trees["fp64"], dim), ref)
    it, rel = trees["mixed"].solve_mixed(TOL_SOLVE, max_outer, n_inner, True)
    assert rel <= TOL_SOLVE, (it, rel)
    r, bmax = trees["mixed"].residual_norms()
    assert r <= 1.01 * TOL_SOLVE * bmax
    err = np.abs(_leaf_phi(trees["mixed"], dim) - ref).max() / np.abs(ref).max()
    assert err <= TOL_PHI, err


@pytest.mark.parametrize("dim,nc,coord", [(2, 8, 0), (2, 8, 1), (3, 8, 0)])
def test_mixed_matches_oracle_fp64(oracle_lib, dim, nc, coord):
    _case(dim, nc, coord, False, 60, 1, 40)


@pytest.mark.parametrize("dim,nc,coord", [(2, 8, 0), (2, 8, 1), (3, 8, 0)])
def test_mixed_matches_oracle_fp64_variable_eps(oracle_lib, dim, nc, coord):
    """With a dielectric: constant- and variable-eps boxes take the
    permittivity operators in FP32 as in FP64."""
    _case(dim, nc, coord, True, 80, 2, 60)
