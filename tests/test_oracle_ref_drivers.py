"""The streamer procedures and the run loop against the reference's own
driver programs.

The reference's drivers (src/streamer_2d.f90, src/streamer_3d.f90,
src/streamer_cyl.f90) are built from their sources into oracle/_ref as
modules (oracle/drv2mod.awk, oracle/ref_drivers.f90): their internal
procedures act on the harness tree, and their main program runs from
init_tree to the end of its time loop. The C restatement (the oracle) and
the host run loop (a5_streamer_amd.streamer, which drives the GPU library
the same way) must give the same doubles:

- each box procedure on seeded random data, with and without a dielectric
  and photoionization: fld_from_pot, fluxes_koren, update_solution,
  average_dens, set_photoi_rate, set_ref_flags, get_max_dt, set_init_cond
  and compute_fld (space charge, FMG with the drivers' sides_bc_pot, field,
  field ghosts);
- whole runs: the initial refinement, time steps, mesh updates and tidy-ups
  of the drivers' main loop, every box after the last step.

Not compared, because the reference never writes them and they hold
whatever its storage held (boxes are allocated without clearing,
afivo/src/m_aX_core.f90:363-381; a run that went differently before leaves
other values there): the photoionization variable when photoionization is
off, the corner ghost cells in 2D and the edge and corner ghost cells in 3D
("corner ghost cells are not used in Afivo", afivo/src/m_aX_gc.f90:2), and
the ghost cells of the right-hand side and the electron face fluxes on boxes
that are not leaves (the multigrid writes only the interior of the first,
fluxes_koren only leaves)."""
import numpy as np
import pytest

import oracle as O
from a5_streamer_amd import streamer as S


def _driver(ref, cfg, end_time=-1.0, time=0.0, photoi=False):
    D = ref.RefDriver(cfg.variant)
    D.config(cfg.domain_len, cfg.box_size, cfg.applied_fld, dt_max=cfg.dt_max,
             end_time=end_time, epsilon_diel=cfg.epsilon_diel, gas_pressure=cfg.gas_pressure,
             fld_mod_t0=cfg.fld_mod_t0, fld_sin_amplitude=cfg.fld_sin_amplitude,
             fld_sin_freq=cfg.fld_sin_freq, fld_lin_deriv=cfg.fld_lin_deriv, time=time,
             steps_amr=cfg.steps_amr, photoi=photoi)
    D.init_cond(cfg.bg_dens, cfg.seed_r0, cfg.seed_r1, cfg.seed_dens, cfg.seed_width,
                cfg.seed_falloff)
    x, cols = cfg.transport_columns()
    D.table(cfg.table_size, cfg.table_max_fld, x, cols)
    return D


def _mask(dim, nc, interior=False):
    """Cells compared: the interior and the face ghosts (or the interior
    only), in Fortran order."""
    n = nc + 2
    ix = np.meshgrid(*([np.arange(n)] * dim), indexing="ij")
    n_ghost = sum(((x == 0) | (x == n - 1)).astype(int) for x in ix)
    return (n_ghost <= (0 if interior else 1)).ravel(order="F")


def _same_tree(ta, tb):
    ba, bb = ta.boxes(), tb.boxes()
    for k in ba:
        np.testing.assert_array_equal(ba[k], bb[k], err_msg="tree: " + k)
    assert ta.info()["max_lvl"] == tb.info()["max_lvl"]
    return ba


def _same_cc(ta, tb, dim, ivs, lvl_min=-10, what=""):
    b = ta.boxes()
    for i in range(len(b["lvl"])):
        if not b["in_use"][i] or b["lvl"][i] < lvl_min:
            continue
        nc = int(b["n_cell"][i])
        leaf = b["children"][i][0] == 0
        for iv in ivs:
            m = _mask(dim, nc, interior=(iv == S.I_RHS and not leaf))
            np.testing.assert_array_equal(
                ta.get_cc(i + 1, iv, nc)[m], tb.get_cc(i + 1, iv, nc)[m],
                err_msg="%s: box %d lvl %d var %d" % (what, i + 1, b["lvl"][i], iv))


def _same_fc(ta, tb, dim, f, what=""):
    """Face variable f on the boxes of levels >= 1 that hold it: the field
    (fld_from_pot) on all of them, the electron flux on the leaves, where
    fluxes_koren and a$D_consistent_fluxes write it (parents keep whatever
    their storage held)."""
    b = ta.boxes()
    for i in range(len(b["lvl"])):
        if f == S.F_ELEC and b["children"][i][0] != 0:
            continue
        if b["in_use"][i] and b["lvl"][i] >= 1:
            for d in range(1, dim + 1):
                np.testing.assert_array_equal(
                    ta.get_fc(i + 1, d, f), tb.get_fc(i + 1, d, f),
                    err_msg="%s: box %d dim %d face var %d" % (what, i + 1, d, f))


# Small versions of the three drivers' runs: one base box, as init_tree
# builds it, and the reference's default of 30 levels
CFGS = {
    "2d": dict(variant="2d", n_base=1, lvls_max=30),
    "2d_dielectric": dict(variant="2d", n_base=1, lvls_max=30, epsilon_diel=1.5,
                          seed_r0=[[16e-3, 12e-3]], seed_r1=[[16e-3, 12e-3]]),
    "2d_modulated": dict(variant="2d", n_base=1, lvls_max=30, fld_mod_t0=0.0,
                         fld_lin_deriv=1e15, fld_sin_amplitude=1e5),
    "cyl": dict(variant="cyl", n_base=1, lvls_max=30, domain_len=16e-3,
                seed_r0=[[0.0, 12e-3]], seed_r1=[[0.0, 12e-3]], applied_fld=2.5e6),
    "3d": dict(variant="3d", n_base=1, lvls_max=30, domain_len=8e-3, n_boxes=5000,
               seed_r0=[[4e-3, 4e-3, 5e-3]], seed_r1=[[4e-3, 4e-3, 6e-3]],
               seed_width=[1.0e-3], applied_fld=2.5e6, epsilon_diel=1.5),
}
END_TIME = {"2d": 4e-9, "2d_dielectric": 3e-9, "2d_modulated": 3e-9, "cyl": 4e-9, "3d": 2e-9}


@pytest.mark.parametrize("name", sorted(CFGS))
def test_run_matches_reference_driver(ref, name):
    """The drivers' main loop against the host run loop over the oracle."""
    cfg = S.Config(**CFGS[name])
    end = END_TIME[name]
    D = _driver(ref, cfg, end_time=end)
    D.counters(reset=True)
    RT = D.run(cfg.box_size)
    r = S.StreamerRun(O.OStreamTree, cfg)
    r.initialise()
    blocks = 0
    cell_steps = 0
    while r.time <= end:
        cell_steps += r.leaf_cells() * cfg.steps_amr
        r.step_block()
        blocks += 1
    assert blocks >= 2
    # the step and leaf-cell counters of the CPU baseline
    assert D.counters() == (blocks * cfg.steps_amr, float(cell_steps))
    t_ref, dt_ref = D.time()
    assert r.time == t_ref
    # the reference computed the next time step before leaving its loop
    assert r.max_dt() == dt_ref
    _same_tree(RT, r.T)
    # the old densities (3, 4) are scratch of a time step: boxes created by
    # the last mesh update have never had them written
    _same_cc(RT, r.T, r.dim, (1, 2, 5, 6, 7, 9), what=name)
    # the field faces; the electron fluxes are scratch of a time step too
    # (leaves created by the last mesh update have none yet), compared where
    # they are computed, in test_box_procedures_match_reference
    _same_fc(RT, r.T, r.dim, 2, what=name)


def _random_state(trees, dim, seed):
    rng = np.random.default_rng(seed)
    b = trees[0].boxes()
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        n = (b["n_cell"][i] + 2) ** dim
        vals = {1: 1e18 * (1 + rng.random(n)), 2: 1e18 * (1 + rng.random(n)),
                3: 1e18 * rng.random(n), 4: 1e18 * rng.random(n),
                5: 1e3 * rng.standard_normal(n), 6: 1e7 * (0.5 + rng.random(n)),
                7: rng.standard_normal(n), 8: 1e20 * rng.random(n),
                9: np.where(rng.random(n) < 0.1, 1.5, 1.0)}
        for iv, v in vals.items():
            for T in trees:
                T.set_cc(i + 1, iv, v)
        if b["lvl"][i] >= 1:
            nf = (b["n_cell"][i] + 1) * b["n_cell"][i] ** (dim - 1)
            for d in range(1, dim + 1):
                for f in (1, 2):
                    # field faces of both signs, so fluxes go both ways
                    v = rng.standard_normal(nf) * (1e7 if f == 2 else 1e20)
                    for T in trees:
                        T.set_fc(i + 1, d, f, v)


PROC_CASES = ["2d", "2d_dielectric", "cyl", "3d"]


def _settled(ref, name):
    """The reference's initial refinement (its main loop left before the
    first step) and the host's, which must be the same tree."""
    cfg = S.Config(**CFGS[name])
    D = _driver(ref, cfg)
    RT = D.run(cfg.box_size)
    r = S.StreamerRun(O.OStreamTree, cfg)
    r.initialise()
    _same_tree(RT, r.T)
    _same_cc(RT, r.T, r.dim, (1, 2, 5, 6, 7, 9), what="initial state")
    return cfg, D, RT, r


@pytest.mark.parametrize("name", PROC_CASES)
def test_box_procedures_match_reference(ref, name):
    cfg, D, RT, r = _settled(ref, name)
    T, dim = r.T, r.dim
    v = T.v
    _random_state((RT, T), dim, 11)
    # fld_from_pot on every box
    D.fld_from_pot()
    O.fld_from_pot(T, v)
    _same_cc(RT, T, dim, (6,), what="fld_from_pot")
    _same_fc(RT, T, dim, 2, what="fld_from_pot")
    # fluxes_koren with the ghost cells the time loop has filled
    for X in (RT, T):
        X.gc_tree(S.I_FLD, S.RB_INTERP, S.BC_NEUMANN)
        X.gc_tree(S.I_ELEC, S.RB_INTERP_LIM, S.BC_NEUMANN)
    # the last time step is long enough for the clamp at the dielectric
    # relaxation flux -0.8 eps0 E / (e dt) to act in both directions
    for dt in (2e-12, 3e-11, 1e-8):
        D.fluxes_koren(dt)
        O.fluxes_koren(T, v, T.tbl, dt)
        _same_fc(RT, T, dim, 1, what="fluxes_koren dt %g" % dt)
    # update_solution without and with the photoionization source
    dt = 2.5e-12
    D.update_solution(dt)
    O.update_solution(T, v, T.tbl, dt)
    _same_cc(RT, T, dim, (1, 2), what="update_solution")
    vp = O.OVars(*[getattr(v, f) for f, _ in O.OVars._fields_])
    vp.photoi = 1
    D.config(cfg.domain_len, cfg.box_size, cfg.applied_fld, dt_max=cfg.dt_max,
             epsilon_diel=cfg.epsilon_diel, photoi=True)
    D.update_solution(dt)
    O.update_solution(T, vp, T.tbl, dt)
    _same_cc(RT, T, dim, (1, 2), what="update_solution with photoionization")
    # average_dens on every box
    D.average_dens()
    O.average_dens(T, v)
    _same_cc(RT, T, dim, (1, 2), what="average_dens")
    # set_photoi_rate on leaves
    coeff = cfg.photoi_model()[0]
    D.photoi_rate(coeff)
    O.photoi_rate(T, v, T.tbl, coeff)
    _same_cc(RT, T, dim, (8,), what="set_photoi_rate")


@pytest.mark.parametrize("name", PROC_CASES)
def test_flags_and_time_step_match_reference(ref, name):
    cfg, D, RT, r = _settled(ref, name)
    T, dim = r.T, r.dim
    n = T.info()["max_id"]
    # on the settled state, then on random data (fields of every size)
    for seed in (None, 12, 13):
        if seed is not None:
            _random_state((RT, T), dim, seed)
            for X in (RT, T):
                for iv in (S.I_FLD, S.I_ELEC):
                    # set_ref_flags and get_max_dt read the interior only
                    X.gc_tree(iv, S.RB_INTERP, S.BC_NEUMANN)
        for time in (0.0, 1e-9, 3e-9, 1e-8):
            D.config(cfg.domain_len, cfg.box_size, cfg.applied_fld, dt_max=cfg.dt_max,
                     epsilon_diel=cfg.epsilon_diel, time=time)
            r.time = time
            np.testing.assert_array_equal(r.ref_flags(), D.ref_flags(n),
                                          err_msg="set_ref_flags at t = %g" % time)
        assert r.max_dt() == D.max_dt()


@pytest.mark.parametrize("name", PROC_CASES)
def test_init_cond_and_field_match_reference(ref, name):
    cfg, D, RT, r = _settled(ref, name)
    T, dim = r.T, r.dim
    _random_state((RT, T), dim, 14)
    D.set_init_cond()
    r.set_init_cond()
    _same_cc(RT, T, dim, (1, 2, 5, 9), what="set_init_cond")
    _random_state((RT, T), dim, 15)
    D.mg_init()
    for no_guess in (True, False):
        D.compute_fld(1, no_guess)
        r.compute_fld(no_guess)
        _same_cc(RT, T, dim, (5, 6, 7), what="compute_fld no_guess=%s" % no_guess)
        _same_fc(RT, T, dim, 2, what="compute_fld")
