"""Output of the leaves (a$D_write_vtk, afivo/src/m_aX_io.f90:20-209, with
the VTK XML writer afivo/src/m_vtk.f90): the oracle's cell data and mesh
against brute force and geometry, and the file the reference's own m_vtk
writes from them (built from source in oracle/_ref), parsed back."""
import re

import numpy as np
import pytest

import scenarios as S
from test_oracle_diag import NC, _fill, _trees


def _leaves(T):
    return [int(i) for l in range(1, T.info()["max_lvl"] + 1) for i in T.level(l)[1]]


@pytest.mark.parametrize("dim", [2, 3])
def test_vtk_mesh_is_the_leaf_cells(oracle_lib, dim):
    T, _ = _trees(dim, None)
    coords, conn, offs, types = T.vtk_geometry()
    b = T.boxes()
    ids = _leaves(T)
    n_ch = 2 ** dim
    assert len(offs) == len(ids) * NC ** dim
    np.testing.assert_array_equal(offs, n_ch * np.arange(1, len(offs) + 1))
    assert np.all(types == (8 if dim == 2 else 11))
    corners = coords[conn.reshape(-1, n_ch)]          # (n_cells, n_ch, dim)
    c = 0
    for i in ids:
        dr, rmin = b["dr"][i - 1], b["r_min"][i - 1]
        for p in np.ndindex(*([NC] * dim)):
            ix = np.array(p[::-1])                     # x fastest
            lo, hi = rmin + ix * dr, rmin + (ix + 1) * dr
            # pixel/voxel node order: x toggles fastest, then y, then z
            want = np.array([[(hi if (q >> d) & 1 else lo)[d] for d in range(dim)]
                             for q in range(n_ch)])
            np.testing.assert_allclose(corners[c], want, rtol=0, atol=1e-15)
            c += 1


@pytest.mark.parametrize("dim", [2, 3])
def test_leaf_cell_columns(oracle_lib, dim):
    T, _ = _trees(dim, None)
    _fill([T], dim, 4)
    b = T.boxes()
    ids = _leaves(T)
    for ixs_fc in ([], [2], [1, 2]):
        out = T.leaf_cells([3, 1], ixs_fc)
        assert out.shape == (len(ids) * NC ** dim, 2 + (dim if ixs_fc else 0))
        c = 0
        for i in ids:
            cc = [T.get_cc(i, v).reshape([NC + 2] * dim, order="F") for v in (3, 1)]
            fs = {}
            for v in ixs_fc:
                for d in range(dim):
                    shape = [NC] * dim
                    shape[d] += 1
                    fs[v, d] = T.get_fc(i, d + 1, v).reshape(shape, order="F")
            for p in np.ndindex(*([NC] * dim)):
                ix = p[::-1]
                assert out[c, 0] == cc[0][tuple(np.add(ix, 1))]
                assert out[c, 1] == cc[1][tuple(np.add(ix, 1))]
                for d in range(dim):
                    if not ixs_fc:
                        break
                    hi = list(ix)
                    hi[d] += 1
                    s = 0.0
                    for v in ixs_fc:
                        s = s + fs[v, d][tuple(ix)]
                        s = s + fs[v, d][tuple(hi)]
                    assert out[c, 2 + d] == 0.5 * s
                c += 1


def parse_vtu(path):
    """The XML lines and the appended arrays of a raw-appended .vtu file
    (a name -> (type, offset) map and the decoded blocks in file order)."""
    raw = open(path, "rb").read()
    i = raw.index(b"_", raw.index(b"<AppendedData"))
    head, app = raw[:i], raw[i + 1:]
    arrays = {}
    for m in re.finditer(rb'<DataArray type="(\w+)"[^>]*?Name="([^"]+)"[^>]*offset="(\d+)"', head):
        arrays[m.group(2).decode()] = (m.group(1).decode(), int(m.group(3)))
    blocks = {}
    for name, (tp, off) in arrays.items():
        nb = int(np.frombuffer(app, "<i4", 1, off)[0])
        dt = {"Float32": "<f4", "Float64": "<f8", "Int32": "<i4"}[tp]
        blocks[name] = np.frombuffer(app, dt, nb // np.dtype(dt).itemsize, off + 4)
    return head.decode().splitlines(), blocks, raw


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_vtk_writer(oracle_lib, ref, dim, tmp_path):
    """The reference's m_vtk, called as a$D_write_vtk calls it, on the
    oracle's arrays: the file decodes back to those arrays."""
    T, _ = _trees(dim, None)
    _fill([T], dim, 8)
    coords, conn, offs, types = T.vtk_geometry()
    cells = T.leaf_cells([1, 2], [1])
    names = ["a", "b"] + ["f_" + "xyz"[d] for d in range(dim)]
    f = tmp_path / "ref.vtu"
    ref.vtk_write(f, coords, conn, offs, types, names, cells, 7, 1.5e-9)
    lines, blocks, _ = parse_vtu(f)
    assert lines[0] == '<?xml version="1.0"?>'
    assert lines[1].startswith('<VTKFile type="UnstructuredGrid"')
    assert float(lines[5]) == 1.5e-9 and int(lines[8]) == 7
    pts = blocks["Points"].reshape(-1, 3)
    np.testing.assert_array_equal(pts[:, :dim], coords.astype(np.float32))
    assert np.all(pts[:, dim:] == 0)
    np.testing.assert_array_equal(blocks["connectivity"], conn)
    np.testing.assert_array_equal(blocks["offsets"], offs)
    np.testing.assert_array_equal(blocks["types"], types)
    for k, n in enumerate(names):
        np.testing.assert_array_equal(blocks[n], cells[:, k])
