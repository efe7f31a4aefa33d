"""The Helmholtz operator and the Helmholtz photoionization model of the
oracle. The reference has neither (its photoionization is Monte-Carlo,
src/m_photons.f90), so they are pinned by the mathematics alone: the
smoother's closed form, a direct sparse solve of the same discrete problem,
second-order convergence to a manufactured solution on a refined mesh, and
the far-field value sum_j amp_j I / kappa_j^2 of a uniform source."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
import scenarios as S
from a5_streamer_amd import streamer as ST


def _idx(dim, nc, i, j, k=0):
    n = nc + 2
    return i + n * j + (n * n * k if dim == 3 else 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_helmholtz_gsrb_closed_form(oracle_lib, dim):
    """One half-sweep sets the cells of colour cntr&1 to
    (sum of the 2*dim neighbours - h^2 rhs) / (2*dim + kappa2 h^2)."""
    nc, kappa2 = 4, 37.0
    T = O.OTree(dim, nc, 4, 0, 1.0 / nc)
    S.build_tree(T, dim, nc, lambda l, c: 0)
    T.mg_setup(1, 2, 3)
    T.mg_set_helmholtz(kappa2)
    rng = np.random.default_rng(3)
    bl = (nc + 2) ** dim
    phi, rhs = rng.standard_normal(bl), rng.standard_normal(bl)
    T.set_cc(1, 1, phi)
    T.set_cc(1, 3, rhs)
    T.box_gsrb(1, 1)
    out = T.get_cc(1, 1)
    h2 = (1.0 / nc) ** 2
    rr = range(1, nc + 1)
    for k in (rr if dim == 3 else [0]):
        for j in rr:
            for i in rr:
                c = _idx(dim, nc, i, j, k)
                if (i + j + k) % 2 == 1:
                    nbs = [c + 1, c - 1, c + nc + 2, c - nc - 2]
                    if dim == 3:
                        nbs += [c + (nc + 2) ** 2, c - (nc + 2) ** 2]
                    want = (sum(phi[m] for m in nbs) - h2 * rhs[c]) / (2 * dim + kappa2 * h2)
                    assert abs(out[c] - want) <= 1e-14 * max(1, abs(want))
                else:
                    assert out[c] == phi[c]


def _leaf_grid(T, dim, nc, iv):
    """Values of iv on a uniform leaf level as one (N,)*dim array, x first."""
    b = T.boxes()
    lvl = T.info()["max_lvl"]
    ids = T.level(lvl)[1]
    h = b["dr"][ids[0] - 1]
    N = int(round(1.0 / h))
    out = np.zeros([N] * dim)
    for i in ids:
        o = np.rint(b["r_min"][i - 1] / h).astype(int)
        v = T.get_cc(int(i), iv).reshape([nc + 2] * dim, order="F")[(slice(1, -1),) * dim]
        out[tuple(slice(o[d], o[d] + nc) for d in range(dim))] = v
    return out, h


@pytest.mark.parametrize("dim,lvls", [(2, 3), (3, 2)])
def test_helmholtz_matches_direct_solve(oracle_lib, dim, lvls):
    """FMG and V-cycles with the Helmholtz operator converge to the solution
    of the discrete problem solved directly: (sum of neighbours - 2 dim phi)
    / h^2 - kappa2 phi = rhs on a uniform grid, a ghost being minus its
    interior neighbour on every domain face."""
    nc, kappa2 = 8, 150.0
    T = O.OTree(dim, nc, 4, 0, 1.0 / nc, coarsen_to=2)
    S.build_tree(T, dim, nc, lambda l, c: 1 if l < lvls else 0)
    rng = np.random.default_rng(11)
    b = T.boxes()
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            n = int(b["n_cell"][i]) + 2
            T.set_cc(i + 1, 3, rng.standard_normal(n ** dim))
    T.mg_setup(1, 2, 4, bc_kind=np.ones(2 * dim, np.int32))
    T.mg_set_helmholtz(kappa2)
    T.fas_fmg(False, True)
    for _ in range(40):
        T.fas_vcycle(T.info()["max_lvl"], False)
    rhs, h = _leaf_grid(T, dim, nc, 4)
    phi, _ = _leaf_grid(T, dim, nc, 1)
    N = rhs.shape[0]
    n = N ** dim
    main = np.full(n, -2 * dim / h ** 2 - kappa2)
    A = sp.diags(main).tolil()
    strides = [N ** d for d in range(dim)]
    idx = np.arange(n).reshape([N] * dim, order="F")
    for d in range(dim):
        for s in (-1, 1):
            src = idx.take(range(max(0, -s), N - max(0, s)), axis=d).ravel(order="F")
            A[src, src + s * strides[d]] = 1 / h ** 2
            edge = idx.take([0 if s < 0 else N - 1], axis=d).ravel(order="F")
            A[edge, edge] = A[edge, edge].toarray().ravel() - 1 / h ** 2
    x = spla.spsolve(A.tocsr(), rhs.ravel(order="F"))
    ref = x.reshape([N] * dim, order="F")
    assert np.abs(phi - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("dim,levels,partial", [(2, (4, 5), True), (3, (2, 3), False)])
def test_helmholtz_manufactured_second_order(oracle_lib, dim, levels, partial):
    """S = prod sin(pi x_d) on the unit domain (zero on its faces) with
    rhs = -(dim pi^2 + kappa2) S, on a mesh refined once more around the
    middle (2D) or uniformly (3D): the error falls about 4x when every leaf
    is refined once."""
    kappa2 = 60.0
    errs = []
    for L in levels:
        nc = 8
        T = O.OTree(dim, nc, 4, 0, 1.0 / nc, coarsen_to=2)
        S.build_tree(T, dim, nc, lambda l, c: 1 if (l < L - 1 or (
            l < L and (not partial or np.all(np.abs(c - 0.5) < 0.3)))) else 0)
        b = T.boxes()
        for i in range(len(b["lvl"])):
            if not b["in_use"][i]:
                continue
            n = b["n_cell"][i] + 2
            ax = [b["r_min"][i][d] + (np.arange(n) - 0.5) * b["dr"][i] for d in range(dim)]
            g = np.meshgrid(*ax, indexing="ij")
            sol = np.prod([np.sin(np.pi * a) for a in g], axis=0)
            T.set_cc(i + 1, 3, (-(dim * np.pi ** 2 + kappa2) * sol).ravel(order="F"))
            T.set_cc(i + 1, 4, sol.ravel(order="F"))
        T.mg_setup(1, 2, 3, bc_kind=np.ones(2 * dim, np.int32))
        T.mg_set_helmholtz(kappa2)
        T.fas_fmg(False, True)
        for _ in range(20):
            T.fas_vcycle(T.info()["max_lvl"], False)
        assert T.max_residual() < 1e-9
        err = 0.0
        for lvl in range(1, T.info()["max_lvl"] + 1):
            for i in T.level(lvl)[1]:
                d = T.get_cc(int(i), 1) - T.get_cc(int(i), 4)
                d = d.reshape([nc + 2] * dim, order="F")[(slice(1, -1),) * dim]
                err = max(err, np.abs(d).max())
        errs.append(err)
    assert 2.5 < errs[0] / errs[1] < 6.0, errs


def test_photoi_uniform_source_far_field(oracle_lib):
    """With a uniform production rate I over a domain many absorption lengths
    wide, each term (lpl - kappa_j^2) S_j = -amp_j I gives S_j = amp_j I /
    kappa_j^2 away from the boundary; I = |E| mu alpha n_e coeff from the
    transport table."""
    cfg = ST.Config(variant="2d", photoi=True)
    coeff, amp, kappa = cfg.photoi_model()
    nc, nb = 8, 8
    dr = 32e-3 / (nc * nb)
    T = O.OStreamTree(2, nc, 18, 2, dr, coarsen_to=2, n_boxes=200)
    T.set_base(*ST.base_grid(2, nb))
    T.table_create(0.0, cfg.table_max_fld, cfg.table_size, 4)
    x, cols = cfg.transport_columns()
    for k, y in enumerate(cols, start=1):
        T.table_set_col(k, x, y)
    T.streamer_setup(True, pho=ST.I_PHO)
    E0, ne0 = 5e6, 1e18
    bl = (nc + 2) ** 2
    for i in range(1, T.info()["max_id"] + 1):
        T.set_cc(i, ST.I_FLD, np.full(bl, E0))
        T.set_cc(i, ST.I_ELEC, np.full(bl, ne0))
    T.mg_setup(ST.I_PHI, ST.I_FLD, ST.I_RHS, bc_kind=np.zeros(4, np.int32))
    T.photoi_helmholtz(coeff, np.array(amp), np.array(kappa), ST.I_PH_TERMS[0],
                       ST.I_PH_RHS[0], ST.I_PH_TMP[0], tol=1e-10, max_cycles=60)
    tb = ST.Table(*T.table_get())
    I0 = E0 * tb.col(ST.T_MOB, E0) * tb.col(ST.T_ALPHA, E0) * ne0 * coeff
    want = sum(a * I0 / k ** 2 for a, k in zip(amp, kappa))
    # the box at the centre of the domain: 16 mm from every face, 13
    # absorption lengths of the longest term
    b = T.boxes()
    centre = [i + 1 for i in range(len(b["lvl"])) if b["in_use"][i] and b["lvl"][i] == 1
              and np.allclose(b["r_min"][i], [16e-3 - nc * dr, 16e-3 - nc * dr])]
    assert len(centre) == 1
    got = T.get_cc(centre[0], ST.I_PHO).reshape(nc + 2, nc + 2, order="F")[1:-1, 1:-1]
    np.testing.assert_allclose(got, want, rtol=1e-4)
    # the multigrid options of the field solve are restored
    T.fas_fmg(False, True)


def _leaf_interiors(T, dim, iv):
    b = T.boxes()
    out = []
    for i in range(len(b["lvl"])):
        if b["in_use"][i] and b["children"][i][0] == 0:
            nc = int(b["n_cell"][i])
            a = T.get_cc(i + 1, iv).reshape([nc + 2] * dim, order="F")
            out.append(a[(slice(1, -1),) * dim].ravel())
    return np.concatenate(out)


def _warm_cold_tight(r, dim):
    """Each term's leaf values after a warm solve and a cold one at the
    configured tolerance, then a cold one to 1e-12, all on the same state."""
    out = []
    tol = r.c.photoi_tol
    for warm, t in ((True, tol), (False, tol), (False, 1e-12)):
        r.c.photoi_warm, r.c.photoi_tol = warm, t
        cyc = r.set_photoionization()
        assert cyc > 0
        out.append([_leaf_interiors(r.T, dim, iv) for iv in ST.I_PH_TERMS])
    r.c.photoi_warm, r.c.photoi_tol = True, tol
    return out


def check_warm_as_accurate(warm, cold, tight):
    """The warm start is as accurate as the cold solve at the same residual
    tolerance: per term, its error against the tightly converged solution is
    within 4 times the cold solve's, or below 1e-6 of the solution's size.
    (On these refined meshes the discrete maximum principle does not bound
    the error by max|r| / kappa^2: the refinement-boundary ghosts are not
    monotone, and the 3D case is hundreds of times over that bound for both
    starts.)"""
    for j in range(len(tight)):
        scale = np.abs(tight[j]).max()
        ew = np.abs(warm[j] - tight[j]).max()
        ec = np.abs(cold[j] - tight[j]).max()
        assert ew <= max(4 * ec, 1e-6 * scale), (j, ew / scale, ec / scale)


def test_warm_start_of_the_terms():
    """Each term kept in its own variable: here a warm call on an unchanged
    source runs no V-cycle (restricting the stored solution to the parents
    and refilling its ghosts leaves the residual under the tolerance on this
    mesh) and gives i_pho bit for bit. After two step blocks (the mesh
    refined, the term variables prolonged to the new boxes) the warm start
    is as accurate as a cold one (FMG from zero)."""
    cfg = ST.Config(variant="2d", photoi=True, lvls_max=6)
    r = ST.StreamerRun(O.OStreamTree, cfg)
    r.initialise()
    p0 = _leaf_interiors(r.T, 2, ST.I_PHO)
    assert r.set_photoionization() == 0
    assert np.array_equal(_leaf_interiors(r.T, 2, ST.I_PHO), p0)
    for _ in range(2):
        r.step_block()
    assert r.photoi_solved
    check_warm_as_accurate(*_warm_cold_tight(r, 2))
