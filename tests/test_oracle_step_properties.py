"""Closed forms that the step operators must reproduce, checked on the oracle
(which the GPU matches bit for bit) on seeded 2D, cylindrical and 3D trees
with refinement boundaries. Each case fixes one term of a step operator:

- fld_from_pot (src/streamer_2d.f90:345-382): the potential a.x + c gives
  -a_d on every face of every box and |a| on every cell;
- fluxes_koren (src/streamer_2d.f90:384-489): a uniform density drifts at
  -mu E n, or sits at the clamp 0.8 eps0 E / (e dt) when mu n exceeds it; a
  linear density gives the central branch of the Koren limiter, the face
  value n(x_face), and the diffusion flux -D dn/dx (sign and factor);
- update_solution (src/streamer_2d.f90:499-542, cylindrical
  src/streamer_cyl.f90:531-570): with the same flux F on every face the
  divergence vanishes in Cartesian coordinates and is F_r / r in cylindrical
  ones, so n_e gains dt (|F| (alpha - eta) + S_pho - F_r / r) and n_i gains
  dt (|F| (alpha - eta) + S_pho); with zero fluxes both gain dt S_pho; cells
  inside the dielectric then hand their electrons to the ions.

The transport table has constant mobility and diffusion and alpha, eta linear
in |E|, which linear interpolation reproduces up to rounding."""
import numpy as np
import pytest

import adjoint_cases as AC
import oracle as O

I_E, I_I, I_PHI, I_FLD, I_PHO, I_EPS = 1, 2, 5, 6, 8, 9   # cell variables
F_E, F_FLD = 1, 2                                          # face variables
MOB, DIF = 0.03, 0.2


def _table(alpha=0.0, eta=0.0):
    x = np.linspace(0.0, 3e7, 301)
    tbl = O.OTable(0.0, 3e7, 1000, 4)
    cols = (np.full_like(x, MOB), np.full_like(x, DIF), alpha * x, eta * x)
    for c, y in enumerate(cols, start=1):
        tbl.set_col(c, x, y)
    return tbl


def _boxes(T):
    """(id, n_cell, dr, r_min, cell centres (dim, (nc+2)^dim) with the ghost
    layer, is_leaf) of every box at level >= 1, flat in the tree's order."""
    b = T.boxes()
    leaves = {i for i, _ in AC.leaf_cells(T)}
    out = []
    for i, nc, lvl in AC.ids_in_use(T):
        if lvl < 1:
            continue
        dr, rmin = b["dr"][i - 1], b["r_min"][i - 1]
        ax = [rmin[d] + (np.arange(nc + 2) - 0.5) * dr for d in range(T.dim)]
        cen = np.stack([g.ravel(order="F") for g in np.meshgrid(*ax, indexing="ij")])
        out.append((i, nc, dr, rmin, cen, i in leaves))
    return out


def _interior(dim, nc):
    m = np.zeros([nc + 2] * dim, bool)
    m[(slice(1, -1),) * dim] = True
    return m.ravel(order="F")


def _face_shape(dim, nc, d):
    return [nc + 1 if e == d else nc for e in range(dim)]


def _face_centres(dim, nc, dr, rmin, d):
    """Centres of the faces normal to axis d, shape (dim, *face shape)."""
    ax = [rmin[e] + ((np.arange(nc + 1) if e == d else np.arange(1, nc + 1) - 0.5)) * dr
          for e in range(dim)]
    return np.stack(np.meshgrid(*ax, indexing="ij"))


@pytest.mark.parametrize("name", ["2d", "cyl", "3d"])
def test_field_of_a_linear_potential(oracle_lib, name):
    T, _, _ = AC.build(O.OTree, name, 3)
    v = O.streamer_vars()
    a = np.array([2.0e6, -3.5e6, 1.25e6])[:T.dim]
    boxes = _boxes(T)
    for i, nc, dr, rmin, cen, _ in boxes:
        T.set_cc(i, I_PHI, a @ cen + 40.0)
        T.set_cc(i, I_EPS, np.ones(cen.shape[1]))
    O.fld_from_pot(T, v)
    for i, nc, dr, rmin, cen, _ in boxes:
        for d in range(T.dim):
            np.testing.assert_allclose(T.get_fc(i, d + 1, F_FLD), -a[d], rtol=1e-11)
        E = T.get_cc(i, I_FLD, nc)[_interior(T.dim, nc)]
        np.testing.assert_allclose(E, np.linalg.norm(a), rtol=1e-11)


@pytest.mark.parametrize("dt, n0, clamped", [(3e-12, 1e18, False), (1e-9, 1e19, True)])
@pytest.mark.parametrize("name", ["2d", "cyl", "3d"])
def test_flux_of_a_uniform_density_is_pure_drift(oracle_lib, name, dt, n0, clamped):
    T, _, _ = AC.build(O.OTree, name, 4)   # seeded |E| and face fields
    v = O.streamer_vars()
    tbl = _table()
    fac = -0.8 * AC.EPS0 / (AC.QE * dt)
    assert (MOB * n0 > -fac) == clamped
    for i, nc, _ in AC.ids_in_use(T):
        T.set_cc(i, I_E, np.full((nc + 2) ** T.dim, n0))
    O.fluxes_koren(T, v, tbl, dt)
    n_leaf = 0
    for i, nc, dr, rmin, cen, leaf in _boxes(T):
        if not leaf:
            continue
        n_leaf += 1
        for d in range(1, T.dim + 1):
            fld = T.get_fc(i, d, F_FLD)
            want = fac * fld if clamped else -MOB * fld * n0
            np.testing.assert_allclose(T.get_fc(i, d, F_E), want, rtol=1e-13, atol=0)
    assert n_leaf > 4


@pytest.mark.parametrize("drift", [True, False])
@pytest.mark.parametrize("name", ["2d", "cyl", "3d"])
def test_flux_of_a_linear_density(oracle_lib, name, drift):
    """Faces 2..nc of each leaf along each axis, whose upwind stencils stay
    in the box and its first ghost layer."""
    T, _, _ = AC.build(O.OTree, name, 6)
    v = O.streamer_vars()
    tbl = _table()
    dt, n0 = 3e-12, 1e18
    g = np.array([300.0, -200.0, 150.0])[:T.dim]
    boxes = _boxes(T)
    for i, nc, dr, rmin, cen, _ in boxes:
        T.set_cc(i, I_E, n0 * (1 + g @ cen))
        if not drift:
            for d in range(1, T.dim + 1):
                T.set_fc(i, d, F_FLD, np.zeros((nc + 1) * nc ** (T.dim - 1)))
    O.fluxes_koren(T, v, tbl, dt)
    for i, nc, dr, rmin, cen, leaf in boxes:
        if not leaf:
            continue
        for d in range(T.dim):
            shp = _face_shape(T.dim, nc, d)
            fc = _face_centres(T.dim, nc, dr, rmin, d)
            n_face = n0 * (1 + np.tensordot(g, fc, axes=1))
            fld = T.get_fc(i, d + 1, F_FLD).reshape(shp, order="F")
            got = T.get_fc(i, d + 1, F_E).reshape(shp, order="F")
            want = -MOB * fld * n_face - DIF * n0 * g[d]
            sel = tuple(slice(1, nc) if e == d else slice(None) for e in range(T.dim))
            scale = MOB * np.abs(fld[sel]) * n_face[sel] + DIF * n0 * abs(g[d])
            assert np.all(np.abs(got[sel] - want[sel]) <= 1e-11 * scale), (i, d)


@pytest.mark.parametrize("flux", ["zero", "uniform"])
@pytest.mark.parametrize("name", ["2d", "2d_photoi", "cyl", "3d"])
def test_update_with_uniform_fluxes_is_the_source_term(oracle_lib, name, flux):
    T, diel, photoi = AC.build(O.OTree, name, 5)
    v = O.streamer_vars(dielectric_rule=diel, photoi=photoi)
    a1, b1 = 1e-2, 2e-4
    tbl = _table(alpha=a1, eta=b1)
    dt = 3e-12
    F = np.array([2e23, -1.5e23, 1e23])[:T.dim] * (flux == "uniform")
    cyl = T.dim == 2 and AC.CASES[name][1] == 1
    before = {}
    for i, nc, dr, rmin, cen, leaf in _boxes(T):
        if not leaf:
            continue
        for d in range(T.dim):
            T.set_fc(i, d + 1, F_E, np.full((nc + 1) * nc ** (T.dim - 1), F[d]))
        before[i] = [T.get_cc(i, iv, nc) for iv in (I_E, I_I, I_FLD, I_PHO, I_EPS)] + [cen]
    O.update_solution(T, v, tbl, dt)
    n_diel = 0
    for i, (ne, ni, E, ph, eps, cen) in before.items():
        nc = round(len(ne) ** (1 / T.dim)) - 2
        m = _interior(T.dim, nc)
        src = np.linalg.norm(F) * (a1 - b1) * E + (ph if photoi else 0.0)
        div = -F[0] / cen[0] if cyl else np.zeros_like(ne)
        ne_w, ni_w = ne + dt * (src + div), ni + dt * src
        scale = 1e-12 * (np.abs(ne) + np.abs(ni) + dt * (np.abs(src) + np.abs(div)))
        if diel:
            d = eps > 1.0
            n_diel += int((d & m).sum())
            ni_w = np.where(d, ni_w - ne_w, ni_w)
            ne_w = np.where(d, 0.0, ne_w)
        got_e, got_i = T.get_cc(i, I_E, nc), T.get_cc(i, I_I, nc)
        assert np.all(np.abs(got_e - ne_w)[m] <= scale[m]), i
        assert np.all(np.abs(got_i - ni_w)[m] <= scale[m]), i
        assert np.array_equal(got_e[~m], ne[~m]) and np.array_equal(got_i[~m], ni[~m])
    assert (n_diel > 0) == diel
