"""The oracle's reverse-mode transport substep (fluxes_koren,
a$D_consistent_fluxes, update_solution on a frozen mesh) against central
finite differences of its forward routines, which are pinned to the
reference's drivers: on seeded 2D, cylindrical and 3D trees with refinement
boundaries, with the dielectric rule and the photoionization source, at a
short time step and at one long enough for the flux clamp to act, along 20
seeded directions that move every input by a relative 1e-6.

The substep is continuous and piecewise smooth in its inputs (limiter cases,
upwind side, clamp, table rows), so a central difference can straddle a kink
and then differs from the derivative by O(h) of the term's scale; in 3D at
h = 1e-6 several table rows are crossed along most directions. The bounds,
relative to the directional derivative's scale sum |g| |dx|: at h = 1e-6
every direction within 1e-5, and at h = 1e-8, where crossings are rare and
rounding still small, every direction within 1e-8 (measured: about 1e-9). A
dropped term or a wrong sign fails them by orders of magnitude."""
import numpy as np
import pytest

import adjoint_cases as AC
import oracle as O

STEPS = ((1e-6, 1e-5), (1e-8, 1e-8))  # (h, bound)


def _oracle_case(name, seed):
    T, diel, photoi = AC.build(O.OTree, name, seed)
    x, cols = AC.table_data()
    tbl = O.OTable(0.0, 3e7, 1000, 4)
    for c, y in enumerate(cols, start=1):
        tbl.set_col(c, x, y)
    v = O.streamer_vars(dielectric_rule=diel, photoi=photoi)
    return T, tbl, v


def _forward(T, tbl, v, dt):
    O.fluxes_koren(T, v, tbl, dt)
    T.consistent_fluxes([1])
    O.update_solution(T, v, tbl, dt)


@pytest.mark.parametrize("dt", [3e-12, 3e-9])
@pytest.mark.parametrize("name", sorted(AC.CASES))
def test_adjoint_matches_central_differences(oracle_lib, name, dt):
    T, tbl, v = _oracle_case(name, 5)
    leaves = AC.leaf_cells(T)
    w = AC.weights(leaves, 6)
    x0 = AC.get_state(T, tbl.data()[0])
    set_table = tbl.set_rows

    def L(x):
        AC.set_state(T, x, set_table)
        _forward(T, tbl, v, dt)
        return AC.loss(T, w)

    # reverse mode at x0
    AC.set_state(T, x0, set_table)
    _forward(T, tbl, v, dt)
    AC.set_state(T, x0, set_table)
    O.fluxes_koren(T, v, tbl, dt)    # the fluxes the update reads
    T.consistent_fluxes([1])
    AC.seed_adjoint(T, w)
    a = O.OAdj(AC.LE, AC.LI, AC.LE_FLD, AC.LPHO, AC.LF, AC.LFLD)
    tg = np.zeros(4 * 1000)
    O.update_solution_adj(T, v, tbl, dt, a, tg)
    O.consistent_fluxes_adj(T, AC.LF)
    O.fluxes_koren_adj(T, v, tbl, dt, a, tg)
    g = AC.gradient(T, tg.reshape(4, 1000))
    dirs = AC.directions(x0, 7, 20)
    for h, bound in STEPS:
        errs = []
        for dx in dirs:
            xp = {k: x0[k] + h * dx[k] for k in x0}
            xm = {k: x0[k] - h * dx[k] for k in x0}
            fd = (L(xp) - L(xm)) / (2 * h)
            an = AC.dot(g, dx)
            scale = sum(float(np.dot(np.abs(g[k]).ravel(), np.abs(dx[k]).ravel())) for k in g)
            errs.append(abs(fd - an) / scale)
        assert max(errs) <= bound, (h, errs)
