"""Mixed precision (FP32 multigrid inside FP64 defect correction): solved to
a relative residual of 1e-10 it must agree with the FP64 solve to <= 1e-8
relative potential error (max norm over leaf cells), in 2D, cylindrical and
3D, with refinement boundaries and all three boundary rules."""
import numpy as np
import pytest

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


@pytest.mark.parametrize("dim,nc,coord", [(2, 8, 0), (2, 8, 1), (3, 8, 0)])
def test_mixed_matches_fp64(dim, nc, coord):
    import a5_streamer_amd as A
    kinds = np.array([0, 0, 1, 2, 1, 2][:2 * dim], np.int32)
    vals = np.array([0, 0, 0, 3.0, 0, 3.0][:2 * dim])
    Ts = []
    for mode in ("fp64", "mixed"):
        T = A.Tree(dim, nc, 3, 0, 1.0 / (2 * nc), coarsen_to=2, coord=coord)
        S.build_tree(T, dim, nc, _rules, n_base=(2, 2, 2))
        T.mg_setup(1, 2, 3, n_base=8, bc_kind=kinds, bc_val=vals, auto_ops=True)
        blobs = [[0.4, 0.6, 0.5][:dim] + [0.08, 40.0], [0.7, 0.3, 0.4][:dim] + [0.05, -60.0]]
        T.set_gaussians(3, blobs, 0.0)
        if mode == "fp64":
            T.fas_fmg(False, True)
            mx = T.info()["max_lvl"]
            for _ in range(60):
                r, b = T.residual_norms()
                if r <= TOL_SOLVE * b:
                    break
                T.fas_vcycle(mx, False)
            assert r <= TOL_SOLVE * b
        else:
            it, rel = T.solve_mixed(TOL_SOLVE, 40, 1, True)
            assert rel <= TOL_SOLVE, (it, rel)
            r, b = T.residual_norms()
            assert r <= 1.01 * TOL_SOLVE * b
        Ts.append(T)
    a, m = _leaf_phi(Ts[0], dim), _leaf_phi(Ts[1], dim)
    err = np.abs(a - m).max() / np.abs(a).max()
    assert err <= TOL_PHI, err


@pytest.mark.parametrize("dim,nc,coord", [(2, 8, 0), (2, 8, 1), (3, 8, 0)])
def test_mixed_matches_fp64_variable_eps(dim, nc, coord):
    """The same with a dielectric: constant- and variable-eps boxes take the
    permittivity operators in FP32 as in FP64."""
    import a5_streamer_amd as A
    from test_oracle_ref import eps_field
    kinds = np.array([0, 0, 1, 2, 1, 2][:2 * dim], np.int32)
    vals = np.array([0, 0, 0, 3.0, 0, 3.0][:2 * dim])
    Ts = []
    for mode in ("fp64", "mixed"):
        T = A.Tree(dim, nc, 4, 0, 1.0 / (2 * nc), coarsen_to=2, coord=coord)
        S.build_tree(T, dim, nc, _rules, n_base=(2, 2, 2))
        eps_field(T, dim, 4, 0)
        T.mg_setup(1, 2, 3, i_eps=4, n_base=8, bc_kind=kinds, bc_val=vals, auto_ops=True)
        blobs = [[0.4, 0.6, 0.5][:dim] + [0.08, 40.0], [0.7, 0.3, 0.4][:dim] + [0.05, -60.0]]
        T.set_gaussians(3, blobs, 0.0)
        if mode == "fp64":
            T.fas_fmg(False, True)
            mx = T.info()["max_lvl"]
            for _ in range(80):
                r, b = T.residual_norms()
                if r <= TOL_SOLVE * b:
                    break
                T.fas_vcycle(mx, False)
            assert r <= TOL_SOLVE * b
        else:
            it, rel = T.solve_mixed(TOL_SOLVE, 60, 2, True)
            assert rel <= TOL_SOLVE, (it, rel)
        Ts.append(T)
    a, m = _leaf_phi(Ts[0], dim), _leaf_phi(Ts[1], dim)
    err = np.abs(a - m).max() / np.abs(a).max()
    assert err <= TOL_PHI, err
