"""Output of the leaves on the GPU (a5g_leaf_cells, a5g_write_vtk,
a5g_dump_leaves through the C ABI): the cell data bit for bit against the
oracle, the .vtu file byte for byte against the one the reference's own
m_vtk writes from the oracle's arrays (the TIME value line compared as a
number), and the raw dump read back."""
import numpy as np
import pytest

import oracle as O
import scenarios as S
from test_oracle_diag import NC, _fill, _rule
from test_oracle_output import parse_vtu

pytestmark = pytest.mark.gpu


def _pair(dim, seed):
    import a5_streamer_amd as A
    ts = [M(dim, NC, 3, 2, 1.0 / NC, coarsen_to=2, n_boxes=50) for M in (O.OTree, A.Tree)]
    for T in ts:
        S.build_tree(T, dim, NC, _rule(dim))
    _fill(ts, dim, seed)
    return ts


@pytest.mark.parametrize("dim", [2, 3])
def test_leaf_cells_bit_exact(oracle_lib, dim):
    print("leaf_cells dim", dim, flush=True)
    ts = _pair(dim, 50 + dim)
    for ixs_cc, ixs_fc in (([1, 2, 3], []), ([2], [1]), ([3, 1], [2, 1]), ([], [2])):
        a = ts[0].leaf_cells(ixs_cc, ixs_fc)
        b = ts[1].leaf_cells(ixs_cc, ixs_fc)
        assert a.shape == b.shape
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dim", [2, 3])
def test_write_vtk_matches_reference_writer(oracle_lib, ref, dim, tmp_path):
    print("write_vtk dim", dim, flush=True)
    T, L = _pair(dim, 60 + dim)
    names = ["n_e", "phi"]
    fnames = ["f%d_%s" % (v, "xyz"[d]) for v in (1, 2) for d in range(dim)]
    L.write_vtk(tmp_path / "lib", names, 12, 2.5e-9, ixs_cc=[1, 3], fc_names=fnames,
                ixs_fc=[1, 2])
    coords, conn, offs, types = T.vtk_geometry()
    cells = T.leaf_cells([1, 3], [1, 2])
    # every face column of dimension d holds the sum over both face variables
    cols = np.concatenate([cells[:, :2]] + [cells[:, 2:]] * 2, axis=1)
    ref.vtk_write(tmp_path / "ref.vtu", coords, conn, offs, types, names + fnames, cols, 12,
                  2.5e-9)
    la, ba, raw_a = parse_vtu(tmp_path / "lib.vtu")
    lb, bb, raw_b = parse_vtu(tmp_path / "ref.vtu")
    assert len(la) == len(lb)
    for x, y in zip(la, lb):
        if x != y:
            assert float(x) == float(y) and x.startswith(" " * 10)
    assert raw_a[raw_a.index(b"<AppendedData"):] == raw_b[raw_b.index(b"<AppendedData"):]


def test_dump_leaves_round_trip(oracle_lib, tmp_path):
    import a5_streamer_amd as A
    print("dump_leaves", flush=True)
    for dim in (2, 3):
        T, L = _pair(dim, 70 + dim)
        L.dump_leaves(tmp_path / "d.bin", [3, 1], cycle=4, time=1e-9)
        d = A.read_leaf_dump(tmp_path / "d.bin")
        assert (d["dim"], d["n_cell"], d["cycle"], d["time"]) == (dim, NC, 4, 1e-9)
        ids = [int(i) for l in range(1, T.info()["max_lvl"] + 1) for i in T.level(l)[1]]
        np.testing.assert_array_equal(d["id"], ids)
        b = L.boxes()
        np.testing.assert_array_equal(d["lvl"], b["lvl"][np.array(ids) - 1])
        np.testing.assert_array_equal(d["r_min"], b["r_min"][np.array(ids) - 1])
        cells = T.leaf_cells([3, 1], [])
        for v in range(2):
            flat = d["data"][v].transpose((0,) + tuple(range(dim, 0, -1))).reshape(-1)
            np.testing.assert_array_equal(flat, cells[:, v])
