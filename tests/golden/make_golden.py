"""Writes the golden fixtures of tests/test_golden.py from the reference
itself (oracle/_ref: the reference's Fortran compiled from its sources by
oracle/Makefile). Each fixture holds the inputs and the reference's outputs,
so the tests need neither the reference nor this script.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import refharness as R  # noqa: E402
import scenarios as S  # noqa: E402

# (name, dim, nc, coord, n_base, refinement centre, radius, max refined lvl)
CASES = [("mg2d", 2, 8, 0, (1, 1, 1), 0.3, 0.25, 4),
         ("mg2d_cyl", 2, 8, 1, (1, 2, 1), 0.35, 0.2, 4),
         ("mg3d", 3, 4, 0, (2, 1, 1), 0.3, 0.3, 3)]


def rules_for(centre, radius, lvl_max):
    def rules(lvl, c):
        if lvl < 2:
            return 1
        return 1 if (lvl < lvl_max and np.linalg.norm(c - centre) < radius) else -1
    return rules


def mg_case(name, dim, nc, coord, base, centre, radius, lvl_max):
    T = R.RefTree(dim, nc, 4, 0, 1.0 / nc, coarsen_to=2, coord=coord, n_boxes=400)
    S.build_tree(T, dim, nc, lambda l, c: rules_for(centre, radius, lvl_max)(l, c / 1.0),
                 n_base=base)
    kinds = np.array([0, 0, 1, 2, 1, 2][:2 * dim], np.int32)
    vals = np.array([0, 0, 0, 2.0, 0, 2.0][:2 * dim])
    T.mg_setup(1, 2, 3, n_base=8, bc_kind=kinds, bc_val=vals, auto_ops=True)
    S.gaussian_rhs(T, dim, 3, [(0.4, 0.6, 0.5)[:dim], (0.7, 0.3, 0.4)[:dim]], [0.1, 0.07],
                   [30.0, -50.0])
    b = T.boxes()
    ids = [i + 1 for i in range(len(b["lvl"])) if b["in_use"][i]]
    # phi and tmp start at zero (the first FMG call does not clear phi on the
    # coarsest level, m_aX_mg.f90:526-549)
    for i in ids:
        z = np.zeros((int(b["n_cell"][i - 1]) + 2) ** dim)
        T.set_cc(i, 1, z)
        T.set_cc(i, 2, z)
    rhs_in = np.stack([T.get_cc(i, 3, int(b["n_cell"][i - 1])) if b["n_cell"][i - 1] == nc
                       else np.pad(T.get_cc(i, 3, int(b["n_cell"][i - 1])),
                                   (0, (nc + 2) ** dim - (int(b["n_cell"][i - 1]) + 2) ** dim))
                       for i in ids])
    T.fas_fmg(True, True)
    T.fas_fmg(True, False)
    T.fas_vcycle(T.info()["max_lvl"], True)
    out = {}
    for iv in (1, 2, 3):
        out["cc%d" % iv] = np.stack([
            np.pad(T.get_cc(i, iv, int(b["n_cell"][i - 1])),
                   (0, (nc + 2) ** dim - (int(b["n_cell"][i - 1]) + 2) ** dim)) for i in ids])
    info = T.info()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), ids=np.array(ids, np.int32),
                        rhs_in=rhs_in, bc_kind=kinds, bc_val=vals,
                        dims=np.array([dim, nc, coord, *base], np.int32),
                        rules=np.array([centre, radius, lvl_max]),
                        max_id=info["max_id"], max_lvl=info["max_lvl"],
                        lvl=b["lvl"], parent=b["parent"], children=b["children"],
                        neighbors=b["neighbors"], n_cell=b["n_cell"], **out)


def table_case():
    x = np.array([3e7 * k / 499 for k in range(500)])
    alpha = np.array([4.3e5 * np.exp(-2e7 / v) if v > 0 else 0.0 for v in x])
    cols = [np.full(500, 0.03), np.full(500, 0.2), alpha, np.zeros(500)]
    t = R.RefTable(0.0, 3e7, 1000, 4)
    for k, y in enumerate(cols, start=1):
        t.set_col(k, x, y)
    x_tab, cr = t.data()
    rc = np.ascontiguousarray(cr.T)  # rows_cols(n_rows, n_cols): column c is rc[c-1]
    rng = np.random.default_rng(11)
    probe = np.concatenate([rng.uniform(-1e6, 3.1e7, 200), [0.0, 3e7]])
    got = np.array([[t.get_col(c, v) for c in (1, 2, 3, 4)] for v in probe])
    a = rng.standard_normal(300) * 10.0 ** rng.integers(-5, 5, 300)
    bb = rng.standard_normal(300) * 10.0 ** rng.integers(-5, 5, 300)
    kor = np.array([R.koren_mlim(u, v) for u, v in zip(a, bb)])
    vecs = rng.standard_normal((300, 2)) * 10.0 ** rng.integers(-20, 20, (300, 1))
    n2 = np.array([R.norm2(v) for v in vecs])
    np.savez_compressed(os.path.join(HERE, "table_koren_norm2.npz"), x_in=x,
                        cols_in=np.stack(cols), rows_cols=rc, x_tab=x_tab,
                        probe=probe, probe_out=got, koren_a=a, koren_b=bb, koren_out=kor,
                        norm2_in=vecs, norm2_out=n2)


if __name__ == "__main__":
    if not R.available():
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    for c in CASES:
        mg_case(*c)
    table_case()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
