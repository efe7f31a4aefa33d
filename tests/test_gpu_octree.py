"""The octree Poisson workload of the benchmark (boxes of 16^3 cells on a
2x2x2 base, a uniform level plus a refined sub-volume, 64 seeded Gaussian
charge blobs, 2V - c / -c / Neumann faces) at a reduced size: uniform level
3 and a 5 x 5 x 8-box sub-volume at level 4 (2,208 boxes, 7.8M leaf cells).

- FP64: one FMG and two V-cycles (mg3_fas_fmg / mg3_fas_vcycle) bit-exact
  against the oracle, with the streaming half-sweeps forced on every 16^3
  level (A5G_STREAM_BYTES=0; the full-size benchmark selects them for its
  finest levels) and with the default selection;
- mixed precision: the FP32-in-FP64 solve to a 1e-10 relative residual,
  with the streaming FP32 kernels forced and without, within 1e-8 relative
  potential error (max norm over leaf cells) of the oracle's FP64 solve to
  the same residual."""
import os

import numpy as np
import pytest

import oracle as O
from a5_streamer_amd import workloads as Wm

pytestmark = pytest.mark.gpu

TOL_SOLVE = 1e-10
TOL_PHI = 1e-8


def _workload():
    return Wm.OctreePoisson(uniform_lvl=3, fine_lvl=4, refine_frac=(5, 5, 8))


def _gpu_tree(W, stream_bytes):
    import a5_streamer_amd as A
    old = os.environ.get("A5G_STREAM_BYTES")
    try:
        if stream_bytes is None:
            os.environ.pop("A5G_STREAM_BYTES", None)
        else:
            os.environ["A5G_STREAM_BYTES"] = str(stream_bytes)
        T = A.Tree(3, W.n_cell, 3, 0, W.dr, coarsen_to=2, n_boxes=1000)
    finally:
        if old is None:
            os.environ.pop("A5G_STREAM_BYTES", None)
        else:
            os.environ["A5G_STREAM_BYTES"] = old
    W.build(T)
    kinds, vals = W.bc()
    T.mg_setup(1, 2, 3, n_down=2, n_up=2, n_base=W.n_cycle_base, bc_kind=kinds, bc_val=vals,
               auto_ops=True)
    T.set_gaussians(3, W.blobs(), 0.0)
    return T


def _oracle_tree(W, Tg):
    """The same tree in the oracle, with the rhs the GPU evaluated."""
    T = O.OTree(3, W.n_cell, 3, 0, W.dr, coarsen_to=2, n_boxes=1000)
    W.build(T)
    b = T.boxes()
    np.testing.assert_array_equal(b["children"], Tg.boxes()["children"])
    rhs = Tg.get_cc_all(3)
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            nc = int(b["n_cell"][i])
            T.set_cc(i + 1, 3, rhs[i][: (nc + 2) ** 3])
    kinds, vals = W.bc()
    T.mg_setup(1, 2, 3, n_down=2, n_up=2, n_base=W.n_cycle_base, bc_kind=kinds, bc_val=vals,
               auto_ops=True)
    return T


def _leaf_phi(T):
    out = []
    info = T.info()
    for lvl in range(1, info["max_lvl"] + 1):
        for i in T.level(lvl)[1]:
            v = T.get_cc(int(i), 1).reshape([T.n_cell + 2] * 3, order="F")
            out.append(v[1:-1, 1:-1, 1:-1].ravel())
    return np.concatenate(out)


@pytest.fixture(scope="module")
def oracle_fp64(oracle_lib):
    """The oracle's FMG + 2 V-cycles, then its solve to the tolerance."""
    W = _workload()
    Tg = _gpu_tree(W, None)
    To = _oracle_tree(W, Tg)
    Tg.close()
    mx = To.info()["max_lvl"]
    To.fas_fmg(False, True)
    for _ in range(2):
        To.fas_vcycle(mx, False)
    b = To.boxes()
    after2 = {i + 1: To.get_cc(i + 1, 1, int(b["n_cell"][i]))
              for i in range(len(b["lvl"])) if b["in_use"][i]}
    rhs_max = np.abs(np.concatenate([To.get_cc(int(i), 3).reshape([18] * 3, order="F")
                                     [1:-1, 1:-1, 1:-1].ravel()
                                     for lvl in range(1, mx + 1)
                                     for i in To.level(lvl)[1]])).max()
    for _ in range(40):
        if To.max_residual() <= TOL_SOLVE * rhs_max:
            break
        To.fas_vcycle(mx, False)
    assert To.max_residual() <= TOL_SOLVE * rhs_max
    return after2, _leaf_phi(To)


@pytest.mark.parametrize("stream_bytes", [0, None])
def test_octree_fmg_vcycles_bit_exact(oracle_fp64, stream_bytes):
    print("octree FP64, A5G_STREAM_BYTES =", stream_bytes, flush=True)
    after2, _ = oracle_fp64
    W = _workload()
    Tg = _gpu_tree(W, stream_bytes)
    mx = Tg.info()["max_lvl"]
    Tg.fas_fmg(False, True)
    for _ in range(2):
        Tg.fas_vcycle(mx, False)
    b = Tg.boxes()
    for id_, ref in after2.items():
        np.testing.assert_array_equal(Tg.get_cc(id_, 1, int(b["n_cell"][id_ - 1])), ref,
                                      err_msg="box %d" % id_)
    Tg.close()


@pytest.mark.parametrize("stream_bytes", [0, None])
def test_octree_mixed_within_tolerance_of_oracle(oracle_fp64, stream_bytes):
    print("octree mixed, A5G_STREAM_BYTES =", stream_bytes, flush=True)
    _, phi_ref = oracle_fp64
    W = _workload()
    Tg = _gpu_tree(W, stream_bytes)
    it, rel = Tg.solve_mixed(TOL_SOLVE, 40, 4, True)
    assert rel <= TOL_SOLVE, (it, rel)
    phi = _leaf_phi(Tg)
    err = np.abs(phi - phi_ref).max() / np.abs(phi_ref).max()
    assert err <= TOL_PHI, err
    Tg.close()
