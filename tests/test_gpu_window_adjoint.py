"""Gradients of whole time steps on a frozen mesh (adjoint.FrozenWindow and
the torch.autograd.Function StreamerWindow) against central differences of
the same window's forward pass.

The forward pass is the drivers' time step (src/streamer_2d.f90:91-129)
with the mesh update left out, stepped by the library, whose every
operation equals the oracle's and the reference's bit for bit
(tests/test_gpu_driver.py); here it is also compared with the oracle
stepping the same window. The loss is the adjoint workload's: electron
count plus the softmax-smoothed head position. The parameters are the
transport table (every column, so the mobility and alpha rows of the
workload are among them) and the applied voltage V; directions are
seeded and relative (each entry scaled by its own value). A central
difference differs from the derivative by the kinks it straddles (limiter
cases, flux clamp, the table's rows: an error that shrinks with h), by
rounding (which grows as h shrinks) and, for the stiff softmax, by its
truncation error. Measured on this window over h = 1e-4 .. 1e-9, the
error relative to the directional derivative's absolute scale is
1e-9 .. 2e-5, and at the best h of each direction 9e-10 .. 1.1e-6 (it
reaches the rounding floor instead of levelling off at a fixed error,
which a wrong term would leave). At h = 1e-6 it stays below 1.5e-6; the
bound is 1e-5. Each reverse operation is pinned exactly on its own
(tests/test_gpu_step_adjoint.py, tests/test_gpu_mg_adjoint.py,
tests/test_gpu_adjoint.py), and the whole gradient is compared with the
oracle's FP64 reverse mode of the same window. The electron count and
the head position are checked apart, since their sizes differ by 15
orders."""
import numpy as np
import pytest

import oracle as O
from a5_streamer_amd import streamer as S
from window_cases import WINDOWS

pytestmark = pytest.mark.gpu

CFG = dict(variant="2d", lvls_max=6, adjoint=True)


def _window(n_steps, every, dts=None):
    import a5_streamer_amd as A
    from a5_streamer_amd import adjoint as AD
    run = S.StreamerRun(A.Tree, S.Config(**CFG))
    run.initialise()
    return run, AD.FrozenWindow(run, n_steps, every=every, dts=dts)


def _losses(win, geo):
    import torch
    from a5_streamer_amd import adjoint as AD
    ne = torch.from_numpy(win.T.get_cc_all(S.I_ELEC))
    E = torch.from_numpy(win.T.get_cc_all(S.I_FLD))
    c = win.run.c
    _, count, head = AD.window_loss(ne, E, geo, c.domain_len, c.applied_fld)
    return float(count), float(head)


def _grads(win, geo, which):
    """(d table, d V) of the count (which 0) or the head position (1)."""
    import torch
    from a5_streamer_amd import adjoint as AD
    win.forward()
    ne = torch.from_numpy(win.T.get_cc_all(S.I_ELEC)).requires_grad_(True)
    E = torch.from_numpy(win.T.get_cc_all(S.I_FLD)).requires_grad_(True)
    c = win.run.c
    out = AD.window_loss(ne, E, geo, c.domain_len, c.applied_fld)
    out[1 + which].backward()
    gne = ne.grad.numpy() if ne.grad is not None else np.zeros(ne.shape)
    gE = E.grad.numpy() if E.grad is not None else np.zeros(E.shape)
    gt, gv, _, _ = win.backward(gne, gE)
    return gt, gv


def test_window_forward_matches_oracle():
    """The window's forward pass is the drivers' step on the frozen mesh:
    the oracle stepping the same steps with the same time steps gives the
    same state bit for bit."""
    print("window forward vs oracle", flush=True)
    run, win = _window(4, 2)
    win.forward()
    ref = S.StreamerRun(O.OStreamTree, S.Config(**dict(CFG, adjoint=False)))
    ref.initialise()
    for dt in win.dts:
        T = ref.T
        T.tree_copy_cc(S.I_ELEC, S.I_ELEC_OLD)
        T.tree_copy_cc(S.I_PION, S.I_PION_OLD)
        ref._euler_pair(dt, False)
        T.average_dens()
        ref.compute_fld(False)
    b = ref.T.boxes()
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        for iv in (S.I_ELEC, S.I_PION, S.I_PHI, S.I_FLD):
            np.testing.assert_array_equal(ref.T.get_cc(i + 1, iv), run.T.get_cc(i + 1, iv),
                                          err_msg="box %d var %d" % (i + 1, iv))


@pytest.mark.parametrize("which", [0, 1])
def test_window_gradient_against_central_differences(which):
    from a5_streamer_amd import adjoint as AD
    print("window gradient, loss term", which, flush=True)
    run, win = _window(6, 4)
    geo = AD.LeafGeometry(win.T)
    gt, gv = _grads(win, geo, which)
    assert np.isfinite(gt).all() and np.isfinite(gv)
    assert np.abs(gt).max() > 0 and gv != 0
    rng = np.random.default_rng(20261204 + which)
    t0, v0 = win.table0, win.voltage0
    h = 1e-6
    worst = 0.0
    for k in range(6):
        dt_ = t0 * rng.standard_normal(t0.shape) * (k != 1)
        dv = v0 * rng.standard_normal() * (k != 0)
        vals = []
        for s in (1.0, -1.0):
            win.forward(t0 + s * h * dt_, v0 + s * h * dv)
            vals.append(_losses(win, geo)[which])
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = float(np.sum(gt * dt_)) + gv * dv
        scale = float(np.sum(np.abs(gt * dt_))) + abs(gv * dv)
        print("  direction %d: fd %.12e ad %.12e rel %.2e" % (k, fd, ad, abs(fd - ad) / scale),
              flush=True)
        worst = max(worst, abs(fd - ad) / scale)
    assert worst <= 1e-5, worst


@pytest.mark.parametrize("which", [0, 1])
def test_window_gradient_matches_oracle_reverse_mode(which):
    """The same window's gradient computed by the oracle's FP64 reverse mode
    (the same FrozenWindow code over OStreamTree, whose reverse routines are
    pinned by the dot-product identity and by central differences on the
    CPU, tests/test_oracle_step_adjoint.py) on the same time steps: the
    forward passes agree bit for bit, and the reverse passes differ only in
    the order of the library's atomic sums (table rows, the potential's
    adjoint through ghost and restriction transposes), so the gradients of
    the table and of the start densities agree to 1e-12 of their largest
    entry (measured: 8e-15 at most). V's adjoint is a sum of boundary
    contributions of both signs that largely cancel, so a reordered sum
    moves it more relative to itself: bound 1e-10 (measured 1e-13 here,
    1.4e-12 in 3D, 2.3e-12 on the 1.5M-cell window of the bench)."""
    import torch
    from a5_streamer_amd import adjoint as AD
    print("window gradient vs oracle, loss term", which, flush=True)
    run, win = _window(6, 4)
    geo = AD.LeafGeometry(win.T)
    win.forward()
    ref = S.StreamerRun(O.OStreamTree, S.Config(**CFG))
    ref.initialise()
    rwin = AD.FrozenWindow(ref, 6, every=4, dts=win.dts)
    rwin.forward()
    c = run.c
    grads = []
    for w in (win, rwin):
        ne = torch.from_numpy(w.T.get_cc_all(S.I_ELEC)).requires_grad_(True)
        E = torch.from_numpy(w.T.get_cc_all(S.I_FLD)).requires_grad_(True)
        AD.window_loss(ne, E, geo, c.domain_len, c.applied_fld)[1 + which].backward()
        gne = ne.grad.numpy() if ne.grad is not None else np.zeros(ne.shape)
        gE = E.grad.numpy() if E.grad is not None else np.zeros(E.shape)
        grads.append(w.backward(gne, gE))
    (gt, gv, gn, gi), (rt, rv, rn, ri) = grads
    assert np.abs(rt).max() > 0 and rv != 0 and np.abs(rn).max() > 0
    for a, r, name in ((gt, rt, "table"), (gn, rn, "n_e start"), (gi, ri, "n_i start")):
        err = np.abs(a - r).max() / np.abs(r).max()
        print("  %s: max |gpu - oracle| / max |oracle| = %.2e" % (name, err), flush=True)
        assert err <= 1e-12, (name, err)
    print("  V: %.16e %.16e" % (gv, rv), flush=True)
    assert abs(gv - rv) <= 1e-10 * abs(rv)


@pytest.mark.parametrize("name", sorted(WINDOWS))
def test_window_other_drivers_match_oracle(name):
    """The 3D driver's step, the cylindrical driver's and the 2D driver's
    with a dielectric (the field solve's permittivity forms on tagged boxes)
    in a window (tests/window_cases.py; a checkpoint every 2 steps): the
    forward pass equals the oracle's window bit for bit, and the gradient of
    both loss terms with respect to the table and the start densities
    equals the oracle's FP64 reverse mode (pinned by central differences on
    the CPU, tests/test_oracle_step_adjoint.py) to 1e-12 of the largest
    entry, and V's to 1e-10 of itself (a sum that cancels; see the 2D
    test)."""
    import torch
    import a5_streamer_amd as A
    from a5_streamer_amd import adjoint as AD
    print("window vs oracle,", name, flush=True)
    cfg, n_steps = WINDOWS[name]
    run = S.StreamerRun(A.Tree, S.Config(**cfg))
    run.initialise()
    win = AD.FrozenWindow(run, n_steps, every=2)
    win.forward()
    ref = S.StreamerRun(O.OStreamTree, S.Config(**cfg))
    ref.initialise()
    rwin = AD.FrozenWindow(ref, n_steps, every=2, dts=win.dts)
    rwin.forward()
    for iv in (S.I_ELEC, S.I_PION, S.I_PHI, S.I_FLD):
        np.testing.assert_array_equal(run.T.get_cc_all(iv), ref.T.get_cc_all(iv), err_msg=str(iv))
    geo = AD.LeafGeometry(run.T)
    c = run.c
    for which in (0, 1):
        grads = []
        for w in (win, rwin):
            w.forward()
            ne = torch.from_numpy(w.T.get_cc_all(S.I_ELEC)).requires_grad_(True)
            E = torch.from_numpy(w.T.get_cc_all(S.I_FLD)).requires_grad_(True)
            AD.window_loss(ne, E, geo, c.domain_len, c.applied_fld)[1 + which].backward()
            gne = ne.grad.numpy() if ne.grad is not None else np.zeros(ne.shape)
            gE = E.grad.numpy() if E.grad is not None else np.zeros(E.shape)
            grads.append(w.backward(gne, gE))
        (gt, gv, gn, gi), (rt, rv, rn, ri) = grads
        assert np.abs(rt).max() > 0 and rv != 0
        for a, r, what in ((gt, rt, "table"), (gn, rn, "n_e start"), (gi, ri, "n_i start")):
            err = np.abs(a - r).max() / np.abs(r).max()
            print("  term %d %s: %.2e" % (which, what, err), flush=True)
            assert err <= 1e-12, (name, which, what, err)
        assert abs(gv - rv) <= 1e-10 * abs(rv), (name, which, gv, rv)


def test_window_gradient_does_not_depend_on_checkpoint_spacing():
    from a5_streamer_amd import adjoint as AD
    print("window checkpoint spacing", flush=True)
    run, win = _window(5, 5)
    geo = AD.LeafGeometry(win.T)
    ga = _grads(win, geo, 0)
    for every in (1, 2, "all"):
        # a new window on the same tree (the same store slots, used in turn)
        # starts from the state and parameters the run holds: rewind it
        win.reset()
        w2 = (AD.FrozenWindow(run, 5, keep_all=True, dts=win.dts) if every == "all"
              else AD.FrozenWindow(run, 5, every=every, dts=win.dts))
        gb = _grads(w2, geo, 0)
        np.testing.assert_allclose(gb[0], ga[0], rtol=0, atol=1e-12 * np.abs(ga[0]).max())
        assert abs(gb[1] - ga[1]) <= 1e-12 * abs(ga[1])


def test_window_through_torch_autograd():
    """StreamerWindow as a torch.autograd.Function: the mobility (a scalar
    filling the mobility column) and the alpha column as leaves of the
    graph, gradients against the window's own backward."""
    import torch
    from a5_streamer_amd import adjoint as AD
    print("window through torch", flush=True)
    run, win = _window(4, 2)
    geo = AD.LeafGeometry(win.T)
    t0 = torch.from_numpy(win.table0.copy())
    mob = torch.tensor(float(t0[S.T_MOB - 1, 0]), dtype=torch.float64, requires_grad=True)
    alpha = t0[S.T_ALPHA - 1].clone().requires_grad_(True)
    volt = torch.tensor(win.voltage0, dtype=torch.float64, requires_grad=True)
    rows = [mob * torch.ones_like(t0[0]), t0[1], alpha, t0[3]]
    table = torch.stack(rows)
    ne, E = AD.streamer_window(win, table, volt)
    c = run.c
    loss, count, head = AD.window_loss(ne, E, geo, c.domain_len, c.applied_fld)
    loss.backward()
    gt, gv = _grads(win, geo, 0)
    gt1, gv1 = _grads(win, geo, 1)
    np.testing.assert_allclose(alpha.grad.numpy(), gt[S.T_ALPHA - 1] + gt1[S.T_ALPHA - 1],
                               rtol=1e-10, atol=1e-12 * np.abs(gt[S.T_ALPHA - 1]).max())
    assert mob.grad.item() == pytest.approx(float(np.sum(gt[S.T_MOB - 1] + gt1[S.T_MOB - 1])),
                                            rel=1e-10)
    assert volt.grad.item() == pytest.approx(gv + gv1, rel=1e-10)
