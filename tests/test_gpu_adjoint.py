"""The GPU's reverse-mode transport substep (a5g_update_solution_adjoint,
a5g_consistent_fluxes_adjoint, a5g_fluxes_koren_adjoint) against the
oracle's, which central differences pin (tests/test_oracle_adjoint.py), and
against central differences of the GPU's own forward pass; and the
torch.autograd.Function over it.

The reverse kernels add with atomics, so sums arrive in another order than
in the oracle's loops: adjoints agree to 1e-12 of each array's largest
magnitude. Finite differences: 20 seeded directions, bounds as in the
oracle's test (1e-5 at h = 1e-6, 1e-8 at h = 1e-8, relative to the scale
sum |g| |dx| of the directional derivative)."""
import numpy as np
import pytest

import adjoint_cases as AC
import oracle as O

pytestmark = pytest.mark.gpu

TOL_REL = 1e-12
STEPS = ((1e-6, 1e-5), (1e-8, 1e-8))
ADJ = np.array([AC.LE, AC.LI, AC.LE_FLD, AC.LPHO, AC.LF, AC.LFLD], np.int32)


def _pair(name, seed):
    import a5_streamer_amd as A
    To, diel, photoi = AC.build(O.OTree, name, seed)
    Tg, _, _ = AC.build(A.Tree, name, seed)
    x, cols = AC.table_data()
    tbl = O.OTable(0.0, 3e7, 1000, 4)
    Tg.table_create(0.0, 3e7, 1000, 4)
    for c, y in enumerate(cols, start=1):
        tbl.set_col(c, x, y)
        Tg.table_set_col(c, x, y)
    v = O.streamer_vars(dielectric_rule=diel, photoi=photoi)
    Tg.streamer_setup(dielectric_rule=diel, pho=8 if photoi else -1)
    return To, tbl, v, Tg


def _gpu_reverse(Tg, w, dt):
    Tg.fluxes_koren(dt)
    Tg.consistent_fluxes([1])
    AC.seed_adjoint(Tg, w)
    Tg.table_adjoint(clear=True)
    Tg.update_solution_adjoint(dt, ADJ)
    Tg.consistent_fluxes_adjoint(AC.LF)
    Tg.fluxes_koren_adjoint(dt, ADJ)
    return AC.gradient(Tg, Tg.table_adjoint(clear=True))


def _oracle_reverse(To, tbl, v, w, dt):
    O.fluxes_koren(To, v, tbl, dt)
    To.consistent_fluxes([1])
    AC.seed_adjoint(To, w)
    a = O.OAdj(*[int(x) for x in ADJ])
    tg = np.zeros(4 * 1000)
    O.update_solution_adj(To, v, tbl, dt, a, tg)
    O.consistent_fluxes_adj(To, AC.LF)
    O.fluxes_koren_adj(To, v, tbl, dt, a, tg)
    return AC.gradient(To, tg.reshape(4, 1000))


@pytest.mark.parametrize("dt", [3e-12, 3e-9])
@pytest.mark.parametrize("name", sorted(AC.CASES))
def test_gpu_adjoint_matches_oracle(oracle_lib, name, dt):
    print("adjoint vs oracle", name, dt, flush=True)
    To, tbl, v, Tg = _pair(name, 5)
    w = AC.weights(AC.leaf_cells(To), 6)
    go = _oracle_reverse(To, tbl, v, w, dt)
    gg = _gpu_reverse(Tg, w, dt)
    assert go.keys() == gg.keys()
    for k in go:
        scale = max(np.abs(go[k]).max(), 1e-300)
        np.testing.assert_allclose(gg[k], go[k], rtol=0, atol=TOL_REL * scale, err_msg=str(k))
    Tg.close()


@pytest.mark.parametrize("name", ["2d", "3d"])
def test_gpu_adjoint_matches_central_differences(oracle_lib, name):
    print("adjoint vs finite differences", name, flush=True)
    dt = 3e-9
    _, _, _, Tg = _pair(name, 5)
    w = AC.weights(AC.leaf_cells(Tg), 6)
    rc, x_min, inv_dx = Tg.table_get()
    x0 = AC.get_state(Tg, rc)
    set_table = lambda r: Tg.set_table(r, x_min, inv_dx)  # noqa: E731

    def L(x):
        AC.set_state(Tg, x, set_table)
        Tg.fluxes_koren(dt)
        Tg.consistent_fluxes([1])
        Tg.update_solution(dt)
        return AC.loss(Tg, w)

    AC.set_state(Tg, x0, set_table)
    g = _gpu_reverse(Tg, w, dt)
    dirs = AC.directions(x0, 7, 20)
    for h, bound in STEPS:
        errs = []
        for dx in dirs:
            fd = (L({k: x0[k] + h * dx[k] for k in x0}) -
                  L({k: x0[k] - h * dx[k] for k in x0})) / (2 * h)
            scale = sum(float(np.dot(np.abs(g[k]).ravel(), np.abs(dx[k]).ravel())) for k in g)
            errs.append(abs(fd - AC.dot(g, dx)) / scale)
        assert max(errs) <= bound, (h, errs)
    Tg.close()


def test_torch_function_gradient(oracle_lib):
    """torch.autograd through TransportSubstep: the gradient of a loss of
    n_e', n_i' equals the oracle's reverse pass seeded with that loss's
    weights, and a directional derivative by central differences."""
    import torch

    from a5_streamer_amd import adjoint as ADJM
    print("torch autograd transport substep", flush=True)
    dt = 3e-9
    To, tbl, v, Tg = _pair("2d_photoi", 5)
    lay = ADJM.AdjointLayout(*[int(x) for x in ADJ])
    leaves = AC.leaf_cells(Tg)
    w = AC.weights(leaves, 6)
    n, bl = Tg.info()["max_id"], Tg.box_len(Tg.n_cell)
    W_e, W_i = np.zeros((n, bl)), np.zeros((n, bl))
    for (i, iv), (ix, wv) in w.items():
        (W_e if iv == 1 else W_i)[i - 1, ix] = wv
    ne = torch.tensor(Tg.get_cc_all(1), requires_grad=True)
    ni = torch.tensor(Tg.get_cc_all(2), requires_grad=True)
    rc, x_min, inv_dx = Tg.table_get()
    tab = torch.tensor(rc, requires_grad=True)
    out_e, out_i = ADJM.transport_substep(Tg, dt, lay, ne, ni, tab)
    loss = (out_e * torch.from_numpy(W_e)).sum() + (out_i * torch.from_numpy(W_i)).sum()
    loss.backward()
    # the oracle's reverse pass on the same inputs
    go = _oracle_reverse(To, tbl, v, w, dt)
    b = To.boxes()
    for i in range(n):
        if not b["in_use"][i]:
            continue
        m = (int(b["n_cell"][i]) + 2) ** Tg.dim
        for grad, key in ((ne.grad, 1), (ni.grad, 2)):
            ref = go[("c", i + 1, key)]
            np.testing.assert_allclose(grad[i, :m].numpy(), ref, rtol=0,
                                       atol=TOL_REL * np.abs(ref).max())
    np.testing.assert_allclose(tab.grad.numpy(), go[("t",)], rtol=0,
                               atol=TOL_REL * np.abs(go[("t",)]).max())
    # a directional derivative through the Function's forward pass
    rng = np.random.default_rng(3)
    d_e = torch.from_numpy(Tg.get_cc_all(1) * rng.standard_normal((n, bl)))
    h = 1e-8
    with torch.no_grad():
        def Lt(s):
            oe, oi = ADJM.transport_substep(Tg, dt, lay, ne + s * d_e, ni, tab)
            return float((oe * torch.from_numpy(W_e)).sum() + (oi * torch.from_numpy(W_i)).sum())
        fd = (Lt(h) - Lt(-h)) / (2 * h)
    an = float((ne.grad * d_e).sum())
    scale = float((ne.grad.abs() * d_e.abs()).sum())
    assert abs(fd - an) <= 1e-8 * scale, (fd, an)
    Tg.close()
