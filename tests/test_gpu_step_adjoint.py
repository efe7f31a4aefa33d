"""The reverse mode of the rest of a streamer time step (a5g_tree_copy_cc_,
average_dens_, space_charge_, restrict_tree_, gc_tree_ and
fld_from_pot_adjoint) against the forward calls of the library, which are
bit for bit the oracle's and the reference's.

Every forward call F here is positively homogeneous of degree one in the
values it reads: linear for the copy, the average, the space charge and
the restriction; piecewise linear with a data-dependent case for the
limited interpolation (the case of s x is the case of x for s > 0); and
the field norm sqrt(sum F^2) / 2 of faces linear in the potential. Its
Jacobian J at x then satisfies J x = F(x), so the reverse call, which
applies J^T at x, must give <w, F(x)> = <J^T w, x> for seeded x and w over
every cell of every box slot (ghosts included; values the call leaves
alone pass through both sides unchanged). A transpose that drops, doubles,
misplaces or mis-signs any term breaks the identity far above the bound
(1e-12 of the sum of the absolute products). The field adjoint is also
checked against a central difference along a random direction, which
pins its Jacobian off the ray through x."""
import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu

# tree cases: dim, coord, base boxes
CASES = [(2, 0, (2, 1, 1)), (2, 1, (1, 2, 1)), (3, 0, (1, 1, 1))]
N_CC, N_FC = 20, 4
# primal variables (the streamer's) and adjoint variables
I_E, I_I, I_EO, I_IO, I_PHI, I_FLD, I_RHS, I_EPS = 1, 2, 3, 4, 5, 6, 7, 9
L_A, L_B, L_C, L_D = 11, 12, 13, 14
F_FLD, LF_FLD = 2, 4


def _rules(lvl, c):
    if lvl < 2:
        return 1
    return 1 if (lvl < 4 and np.linalg.norm(c - 0.4) < 0.22) else -1


def _tree(dim, coord, base):
    import a5_streamer_amd as A
    nc = 8
    T = A.Tree(dim, nc, N_CC, N_FC, 1.0 / nc, coarsen_to=2, coord=coord)
    S.build_tree(T, dim, nc, _rules, n_base=base)
    T.streamer_setup()
    return T


def _shape(T):
    return (T.info()["max_id"], T.box_len(T.n_cell))


def _dot(a, b):
    return float(np.sum(a * b)), float(np.sum(np.abs(a * b)))


def _check(pairs_y, pairs_x):
    """<w, F(x)> against <J^T w, x>: lists of (w, y) and (xbar, x)."""
    lhs = sum(_dot(w, y)[0] for w, y in pairs_y)
    rhs = sum(_dot(xb, x)[0] for xb, x in pairs_x)
    scale = sum(_dot(w, y)[1] for w, y in pairs_y) + sum(_dot(xb, x)[1] for xb, x in pairs_x)
    assert scale > 0
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


def _faces(T, ifc):
    b = T.boxes()
    out = []
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            for d in range(1, T.dim + 1):
                out.append(T.get_fc(i + 1, d, ifc))
    return np.concatenate(out)


def _set_faces(T, ifc, rng):
    b = T.boxes()
    nf = (T.n_cell + 1) * T.n_cell ** (T.dim - 1)
    vals = []
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            for d in range(1, T.dim + 1):
                v = rng.standard_normal(nf)
                T.set_fc(i + 1, d, ifc, v)
                vals.append(v)
    return np.concatenate(vals)


@pytest.mark.parametrize("dim,coord,base", CASES)
def test_copy_average_space_charge_dot_product(dim, coord, base):
    print("copy/average/space charge adjoint", dim, coord, base, flush=True)
    T = _tree(dim, coord, base)
    rng = np.random.default_rng(7)
    sh = _shape(T)
    # copy: (a, b) -> (a, a)
    x = [rng.standard_normal(sh) for _ in range(2)]
    w = [rng.standard_normal(sh) for _ in range(2)]
    T.set_cc_all(I_E, x[0]); T.set_cc_all(I_EO, x[1])
    T.tree_copy_cc(I_E, I_EO)
    y = [T.get_cc_all(I_E), T.get_cc_all(I_EO)]
    T.set_cc_all(L_A, w[0]); T.set_cc_all(L_B, w[1])
    T.tree_copy_cc_adjoint(L_A, L_B)
    _check(list(zip(w, y)), list(zip([T.get_cc_all(L_A), T.get_cc_all(L_B)], x)))
    # average: (e, eo, i, io) -> (0.5 (e + eo), eo, 0.5 (i + io), io)
    ivs, lvs = (I_E, I_EO, I_I, I_IO), (L_A, L_B, L_C, L_D)
    x = [rng.standard_normal(sh) for _ in range(4)]
    w = [rng.standard_normal(sh) for _ in range(4)]
    for iv, v in zip(ivs, x):
        T.set_cc_all(iv, v)
    T.average_dens()
    y = [T.get_cc_all(iv) for iv in ivs]
    for lv, v in zip(lvs, w):
        T.set_cc_all(lv, v)
    T.average_dens_adjoint(*lvs)
    _check(list(zip(w, y)), list(zip([T.get_cc_all(lv) for lv in lvs], x)))
    # space charge: rhs = fac (n_e - n_i) on the leaves
    fac = 1.6022e-19 / 8.8541878176e-12
    ivs, lvs = (I_RHS, I_E, I_I), (L_A, L_B, L_C)
    x = [rng.standard_normal(sh) for _ in range(3)]
    w = [rng.standard_normal(sh) for _ in range(3)]
    for iv, v in zip(ivs, x):
        T.set_cc_all(iv, v)
    T.space_charge(fac)
    y = [T.get_cc_all(iv) for iv in ivs]
    assert not np.array_equal(y[0], x[0])
    for lv, v in zip(lvs, w):
        T.set_cc_all(lv, v)
    T.space_charge_adjoint(fac, *lvs)
    _check(list(zip(w, y)), list(zip([T.get_cc_all(lv) for lv in lvs], x)))
    T.close()


@pytest.mark.parametrize("dim,coord,base", CASES)
def test_restrict_dot_product(dim, coord, base):
    print("restrict adjoint", dim, coord, base, flush=True)
    T = _tree(dim, coord, base)
    rng = np.random.default_rng(11)
    sh = _shape(T)
    x, w = rng.standard_normal(sh), rng.standard_normal(sh)
    T.set_cc_all(I_E, x)
    T.restrict_tree(I_E)
    y = T.get_cc_all(I_E)
    assert not np.array_equal(x, y)
    T.set_cc_all(L_A, w)
    T.restrict_tree_adjoint(L_A)
    _check([(w, y)], [(T.get_cc_all(L_A), x)])
    T.close()


@pytest.mark.parametrize("rb,bc", [(1, 1), (2, 1), (2, 2), (1, 2)])
@pytest.mark.parametrize("dim,coord,base", CASES)
def test_ghost_interpolation_dot_product(dim, coord, base, rb, bc):
    print("gc adjoint", dim, coord, base, rb, bc, flush=True)
    T = _tree(dim, coord, base)
    rng = np.random.default_rng(13 + rb + 2 * bc)
    sh = _shape(T)
    # positive, widely spread values: the limiter of interp_lim clamps many
    # refinement-boundary ghosts (v > 2 v1 where the fine cell is large)
    x = np.exp(2.0 * rng.standard_normal(sh))
    w = rng.standard_normal(sh)
    T.set_cc_all(I_E, x)
    T.gc_tree(I_E, rb, bc)
    y = T.get_cc_all(I_E)
    T.set_cc_all(L_A, w)
    T.gc_tree_adjoint(I_E, L_A, rb, bc)
    _check([(w, y)], [(T.get_cc_all(L_A), x)])
    T.close()


@pytest.mark.parametrize("dim,coord,base", CASES)
def test_field_from_potential_adjoint(dim, coord, base):
    print("field adjoint", dim, coord, base, flush=True)
    T = _tree(dim, coord, base)
    rng = np.random.default_rng(17)
    sh = _shape(T)
    phi = rng.standard_normal(sh)
    T.set_cc_all(I_EPS, 1.0 + rng.random(sh))
    fld0 = rng.standard_normal(sh)
    T.set_cc_all(I_PHI, phi)
    T.set_cc_all(I_FLD, fld0)
    F0 = _set_faces(T, F_FLD, rng)
    T.fld_from_pot()
    y_phi, y_fld, y_F = T.get_cc_all(I_PHI), T.get_cc_all(I_FLD), _faces(T, F_FLD)
    w_phi, w_fld = rng.standard_normal(sh), rng.standard_normal(sh)
    T.set_cc_all(L_A, w_phi)
    T.set_cc_all(L_B, w_fld)
    w_F = _set_faces(T, LF_FLD, rng)
    # the potential adjoint the device holds before the call (a box coarser
    # than n_cell keeps the start of its slot only)
    w_phi_dev = T.get_cc_all(L_A)
    T.fld_from_pot_adjoint(L_B, LF_FLD, L_A)
    xb_phi, xb_fld, xb_F = T.get_cc_all(L_A), T.get_cc_all(L_B), _faces(T, LF_FLD)
    _check([(w_phi, y_phi), (w_fld, y_fld), (w_F, y_F)],
           [(xb_phi, phi), (xb_fld, fld0), (xb_F, F0)])
    # central difference of <w, F(phi)> along a random direction
    delta = rng.standard_normal(sh)
    h = 1e-6
    vals = []
    for s in (1.0, -1.0):
        T.set_cc_all(I_PHI, phi + s * h * delta)
        T.fld_from_pot()
        vals.append(_dot(w_fld, T.get_cc_all(I_FLD))[0] + _dot(w_F, _faces(T, F_FLD))[0])
    fd = (vals[0] - vals[1]) / (2 * h)
    ad, ad_scale = _dot(xb_phi - w_phi_dev, delta)
    # the central difference's error at h = 1e-6 (curvature of the norm,
    # rounding of sums over ~1e4 terms) is far below 1e-6 of the sum of the
    # absolute terms; a wrong coefficient moves ad by a fraction of it
    assert abs(fd - ad) <= 1e-6 * ad_scale, (fd, ad, ad_scale)
    T.close()


def test_store_round_trip_and_refusals():
    print("state store", flush=True)
    T = _tree(2, 0, (2, 1, 1))
    rng = np.random.default_rng(3)
    sh = _shape(T)
    T.set_cc_all(I_E, rng.standard_normal(sh))
    T.set_cc_all(I_PHI, rng.standard_normal(sh))
    _set_faces(T, F_FLD, rng)
    # what the device holds (a box coarser than n_cell keeps the start of
    # its slot only)
    a, b, fa = T.get_cc_all(I_E), T.get_cc_all(I_PHI), _faces(T, F_FLD)
    T.state_store(5, [I_E, I_PHI], [F_FLD])
    T.set_cc_all(I_E, np.zeros(sh))
    T.set_cc_all(I_PHI, np.zeros(sh))
    _set_faces(T, F_FLD, rng)
    T.state_store(5, [I_E, I_PHI], [F_FLD], load=True)
    np.testing.assert_array_equal(T.get_cc_all(I_E), a)
    np.testing.assert_array_equal(T.get_cc_all(I_PHI), b)
    np.testing.assert_array_equal(_faces(T, F_FLD), fa)
    T.state_store(0, [I_PHI], [F_FLD], clear=True)
    assert not T.get_cc_all(I_PHI).any() and not _faces(T, F_FLD).any()
    np.testing.assert_array_equal(T.get_cc_all(I_E), a)
    with pytest.raises(Exception):
        T.state_store(5, [I_E], [F_FLD], load=True)    # other variables
    with pytest.raises(Exception):
        T.state_store(6, [I_E], [], load=True)         # nothing saved there
    with pytest.raises(Exception):
        T.gc_tree_adjoint(I_E, I_E, 2, 1)              # the adjoint must be its own
    with pytest.raises(Exception):
        T.gc_tree_adjoint(I_E, L_A, 0, 1)              # multigrid rule: not covered
    T.close()
