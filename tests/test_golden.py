"""Golden fixtures written by tests/golden/make_golden.py from the reference
compiled from its own sources: FMG + V-cycle results on small 2D, cylindrical
and 3D AMR trees (with the tree they ran on), the transport table built from
the air-model rows, the Koren limiter and NORM2. The oracle must reproduce
them bit for bit on the CPU, and the HIP library on the GPU."""
import os

import numpy as np
import pytest

import oracle as O
import scenarios as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MG_CASES = ["mg2d", "mg2d_cyl", "mg3d"]


def _load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _run_mg(M, g):
    dim, nc, coord = (int(v) for v in g["dims"][:3])
    base = tuple(int(v) for v in g["dims"][3:])
    centre, radius, lvl_max = g["rules"]

    def rules(lvl, c):
        if lvl < 2:
            return 1
        return 1 if (lvl < lvl_max and np.linalg.norm(c - centre) < radius) else -1

    T = M(dim, nc, 4, 0, 1.0 / nc, coarsen_to=2, coord=coord, n_boxes=400)
    S.build_tree(T, dim, nc, rules, n_base=base)
    b = T.boxes()
    assert T.info()["max_id"] == int(g["max_id"])
    for k in ("lvl", "parent", "children", "neighbors", "n_cell"):
        np.testing.assert_array_equal(b[k], g[k], err_msg=k)
    T.mg_setup(1, 2, 3, n_base=8, bc_kind=g["bc_kind"], bc_val=g["bc_val"], auto_ops=True)
    for row, i in enumerate(g["ids"]):
        n = (int(b["n_cell"][i - 1]) + 2) ** dim
        T.set_cc(int(i), 1, np.zeros(n))
        T.set_cc(int(i), 2, np.zeros(n))
        T.set_cc(int(i), 3, g["rhs_in"][row][:n])
    T.fas_fmg(True, True)
    T.fas_fmg(True, False)
    T.fas_vcycle(T.info()["max_lvl"], True)
    for iv in (1, 2, 3):
        for row, i in enumerate(g["ids"]):
            nci = int(b["n_cell"][i - 1])
            n = (nci + 2) ** dim
            np.testing.assert_array_equal(T.get_cc(int(i), iv, nci), g["cc%d" % iv][row][:n],
                                          err_msg="box %d var %d" % (i, iv))


@pytest.mark.parametrize("name", MG_CASES)
def test_oracle_matches_golden_mg(oracle_lib, name):
    _run_mg(O.OTree, _load(name))


def test_oracle_matches_golden_table_koren_norm2(oracle_lib):
    g = _load("table_koren_norm2")
    t = O.OTable(0.0, 3e7, 1000, 4)
    for k in range(4):
        t.set_col(k + 1, g["x_in"], g["cols_in"][k])
    rc, _, _ = t.data()
    np.testing.assert_array_equal(rc, g["rows_cols"])
    got = np.array([[t.get_col(c, v) for c in (1, 2, 3, 4)] for v in g["probe"]])
    np.testing.assert_array_equal(got, g["probe_out"])
    kor = np.array([O.koren_mlim(a, b) for a, b in zip(g["koren_a"], g["koren_b"])])
    np.testing.assert_array_equal(kor, g["koren_out"])
    n2 = np.array([O.norm2(v) for v in g["norm2_in"]])
    np.testing.assert_array_equal(n2, g["norm2_out"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", MG_CASES)
def test_gpu_matches_golden_mg(name):
    import a5_streamer_amd as A
    print("golden", name, flush=True)
    _run_mg(A.Tree, _load(name))


@pytest.mark.gpu
def test_gpu_matches_golden_table():
    import a5_streamer_amd as A
    g = _load("table_koren_norm2")
    T = A.Tree(2, 8, 1, 0, 1.0)
    T.table_create(0.0, 3e7, 1000, 4)
    for k in range(4):
        T.table_set_col(k + 1, g["x_in"], g["cols_in"][k])
    rc, _, _ = T.table_get()
    np.testing.assert_array_equal(rc, g["rows_cols"])
