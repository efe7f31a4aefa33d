"""Host side of a streamer run, following the control flow of the reference
drivers src/streamer_2d.f90:23-143, src/streamer_3d.f90:23-143 and
src/streamer_cyl.f90:26-152 over the a5g C ABI.

The per-box physics runs on the GPU; the host keeps what the drivers keep:
configuration, initial conditions (evaluated on the host and uploaded, as
set_init_cond/set_box_eps do per box), the refinement criterion (evaluated
from per-box maxima that the GPU reduces), the time-step rule and the loop.
The run is written against the Tree interface of a5_streamer_amd.lib, so any
object with the same methods can drive it; the tests use that to step the
same run through the CPU restatement and compare every state bit for bit.

The streamer diagnostics of the cylindrical driver (streamer_properties,
cc_axis) use the GPU's located reductions. Silo output and Monte-Carlo
photoionization are not part of this path. Photoionization is provided by the three-term
Helmholtz model instead (Config.photoi), at the place where the drivers call
set_photoionization.
"""
import math

import numpy as np

EPS0 = 8.8541878176e-12  # fosito/m_units_constants.f90:34
ELEC = 1.6022e-19        # fosito/m_units_constants.f90:35
EPS_F64 = 2.220446049250313e-16  # epsilon(1.0_dp)

# variable indices of src/m_streamer.f90:16-43
I_ELEC, I_PION, I_ELEC_OLD, I_PION_OLD, I_PHI, I_FLD, I_RHS, I_PHO, I_EPS = range(1, 10)
N_VAR_CELL, N_VAR_FACE = 9, 2
# variables of the three photoionization terms, after the reference's nine:
# each term's solution (kept from one step to the next), right-hand side and
# multigrid scratch
I_PH_TERMS = (10, 11, 12)
I_PH_RHS = (13, 14, 15)
I_PH_TMP = (16, 17, 18)
UC_TORR_TO_BAR = 133.322368 * 1.0e-5  # fosito/m_units_constants.f90:44
UC_PI = math.pi                        # acos(-1), fosito/m_units_constants.f90:24
# Three-term fit of the photoionization absorption function of air (Bourdon
# et al. 2007, Plasma Sources Sci. Technol. 16, 656, table 1):
# A_j in cm^-2 Torr^-2 and lambda_j in cm^-1 Torr^-1
HELMHOLTZ_A = (1.986e-4, 0.0051, 0.4886)
HELMHOLTZ_LAMBDA = (0.0553, 0.1460, 0.89)
F_ELEC, F_FLD = 1, 2
T_MOB, T_DIF, T_ALPHA, T_ETA = 1, 2, 3, 4
RB_MG, RB_INTERP, RB_INTERP_LIM = 0, 1, 2   # mg_sides_rb, gc_interp, gc_interp_lim
BC_MG, BC_NEUMANN = 0, 1                    # the solver's sides_bc, gc_neumann
# sides_bc_pot rules per face: copy, -c, 2V - c on the face, and 2V - c over
# the whole ghost row with its corner ghosts (the cylindrical driver's high
# z face, src/streamer_cyl.f90:408-410)
BC_COPY, BC_NEG, BC_2V, BC_ROW_2V = 0, 1, 2, 3
N_FMG_CYCLES = 1                            # src/m_streamer.f90:46
# variables of the reverse mode (Config.adjoint), after the run's own
N_ADJ_CELL, N_ADJ_FACE = 8, 2


def norm2(v):
    """NORM2 as the flang runtime evaluates it in double precision: running
    maximum m and s = sum((x/m)^2) over the other entries, result
    m sqrt(1 + s)."""
    mx, s = 0.0, 0.0
    for x in v:
        ax = abs(x)
        if mx == 0:
            mx = ax
        elif ax > mx:
            q = mx / ax
            qq = q * q
            s = s * qq
            s = s + qq
            mx = ax
        else:
            q = ax / mx
            s = s + q * q
    return mx * math.sqrt(1 + s)


def dist_line(r, r0, r1):
    """GM_dist_line, src/m_geom.f90:29-46."""
    d = len(r)
    line_len2 = 0.0
    for k in range(d):
        t = r1[k] - r0[k]
        line_len2 = line_len2 + t * t
    temp = 0.0
    for k in range(d):
        temp = temp + (r[k] - r0[k]) * (r1[k] - r0[k])
    if temp <= 0.0:
        return norm2([r[k] - r0[k] for k in range(d)])
    if temp >= line_len2:
        return norm2([r[k] - r1[k] for k in range(d)])
    proj = [r0[k] + temp / line_len2 * (r1[k] - r0[k]) for k in range(d)]
    return norm2([r[k] - proj[k] for k in range(d)])


def norm2_rows(v):
    """norm2 of each row of v (n, d), entry by entry in the same order and
    with the same operations as norm2, so each result is the same double."""
    v = np.asarray(v, float)
    mx = np.zeros(v.shape[0])
    s = np.zeros(v.shape[0])
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(v.shape[1]):
            ax = np.abs(v[:, k])
            first = mx == 0
            up = ~first & (ax > mx)
            rest = ~first & ~up
            q = mx / ax
            qq = q * q
            p = ax / mx
            s = np.where(up, s * qq + qq, np.where(rest, s + p * p, s))
            mx = np.where(first | up, ax, mx)
    return mx * np.sqrt(1 + s)


def dist_line_rows(r, r0, r1):
    """dist_line of each row of r (n, d) to the segment r0-r1, with the
    operations of dist_line."""
    r = np.asarray(r, float)
    d = r.shape[1]
    line_len2 = 0.0
    for k in range(d):
        t = r1[k] - r0[k]
        line_len2 = line_len2 + t * t
    temp = np.zeros(r.shape[0])
    for k in range(d):
        temp = temp + (r[:, k] - r0[k]) * (r1[k] - r0[k])
    to0 = norm2_rows(np.stack([r[:, k] - r0[k] for k in range(d)], axis=1))
    to1 = norm2_rows(np.stack([r[:, k] - r1[k] for k in range(d)], axis=1))
    with np.errstate(divide="ignore", invalid="ignore"):
        proj = [r0[k] + temp / line_len2 * (r1[k] - r0[k]) for k in range(d)]
    top = norm2_rows(np.stack([r[:, k] - proj[k] for k in range(d)], axis=1))
    return np.where(temp <= 0.0, to0, np.where(temp >= line_len2, to1, top))


def dist_line_rows_per(r, r0, r1):
    """dist_line of each row of r (n, d) to its own segment r0[row]-r1[row],
    with the operations of dist_line."""
    r, r0, r1 = (np.asarray(x, float) for x in (r, r0, r1))
    d = r.shape[1]
    line_len2 = np.zeros(r.shape[0])
    for k in range(d):
        t = r1[:, k] - r0[:, k]
        line_len2 = line_len2 + t * t
    temp = np.zeros(r.shape[0])
    for k in range(d):
        temp = temp + (r[:, k] - r0[:, k]) * (r1[:, k] - r0[:, k])
    to0 = norm2_rows(np.stack([r[:, k] - r0[:, k] for k in range(d)], axis=1))
    to1 = norm2_rows(np.stack([r[:, k] - r1[:, k] for k in range(d)], axis=1))
    with np.errstate(divide="ignore", invalid="ignore"):
        proj = [r0[:, k] + temp / line_len2 * (r1[:, k] - r0[:, k]) for k in range(d)]
    top = norm2_rows(np.stack([r[:, k] - proj[k] for k in range(d)], axis=1))
    return np.where(temp <= 0.0, to0, np.where(temp >= line_len2, to1, top))


def seed_hits(centre, boxlen, dr, r0, r1, w):
    """The seed rule of set_ref_flags for boxes (n rows) and one seed each
    (r0, r1 (n, d), w (n,)): dist_line(centre, r0, r1) - w < boxlen and
    dr > 0.2 w (src/streamer_2d.f90:200-211). The distance is computed only
    for boxes that the cheap tests leave open: the second condition, and the
    box centre's gap to the segment's bounding box (a lower bound of the
    distance) below (w + boxlen)(1 + 1e-9); every other box is provably no
    hit, so each flag equals the one computed for every box."""
    hit = np.zeros(len(boxlen), bool)
    idx = np.nonzero(dr > 0.2 * w)[0]
    if len(idx) == 0:
        return hit
    c, a, e = centre[idx], r0[idx], r1[idx]
    gap = np.max(np.maximum(0.0, np.maximum(np.minimum(a, e) - c, c - np.maximum(a, e))),
                 axis=1)
    idx = idx[gap < (w[idx] + boxlen[idx]) * (1 + 1e-9)]
    if len(idx):
        dist = dist_line_rows_per(centre[idx], r0[idx], r1[idx])
        hit[idx] = dist - w[idx] < boxlen[idx]
    return hit


_POOL = None


def _chunked(n, fn, chunk=32768):
    """fn(lo, hi) over [0, n) in chunks, on a small thread pool when there
    is more than one chunk."""
    global _POOL
    bounds = [(lo, min(n, lo + chunk)) for lo in range(0, n, chunk)]
    if len(bounds) <= 1:
        for lo, hi in bounds:
            fn(lo, hi)
        return
    if _POOL is None:
        import concurrent.futures
        import os
        _POOL = concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1))
    for fut in [_POOL.submit(fn, lo, hi) for lo, hi in bounds]:
        fut.result()


def group_max(g, v, n, empty):
    """max of v over the rows of each group 0 <= g < n (empty for a group
    with no rows). np.maximum.at with the values in the result's dtype (a
    mixed-dtype call takes numpy's slow generic loop)."""
    v = np.ascontiguousarray(v)
    res = np.full(n, empty, dtype=v.dtype)
    np.maximum.at(res, np.asarray(g, np.int64), v)
    return res


_LOG_HALF_HUGE = math.log(0.5 * 1.7976931348623157e308)


def dens_line(r, r0, r1, width, kind):
    """GM_dens_line with GM_sigmoid/_gaussian/_smoothstep/_step
    (src/m_geom.f90:48-110); kind 1..4 as GM_*_t (src/m_geom.f90:8-11)."""
    dist = dist_line(r, r0, r1)
    if kind == 1:
        tmp = dist / width
        if tmp > _LOG_HALF_HUGE:
            return 0.0
        return 2 / (1 + math.exp(tmp))
    if kind == 2:
        q = dist / width
        return math.exp(-(q * q))
    if kind == 3:
        if dist < width:
            return 1.0
        if dist < 2 * width:
            t = dist / width - 1
            return 1 - (3 * (t * t) - 2 * (t * t * t))
        return 0.0
    if kind == 4:
        return 1.0 if dist < width else 0.0
    return 0.0


class Table:
    """Host copy of the uniform transport table for the scalar lookups the
    drivers make (LT_get_col, fosito/m_lookup_table.f90:100-144)."""

    def __init__(self, rows_cols, x_min, inv_dx):
        self.rc = np.asarray(rows_cols)
        self.n_rows = self.rc.shape[1]
        self.x_min, self.inv_dx = x_min, inv_dx

    def col(self, c, x):
        frac = (x - self.x_min) * self.inv_dx
        li = math.ceil(frac)
        lf = li - frac
        if li < 1:
            li, lf = 1, 1.0
        elif li >= self.n_rows:
            li, lf = self.n_rows - 1, 0.0
        t = self.rc[c - 1]
        return lf * t[li - 1] + (1 - lf) * t[li]

    def col_rows(self, c, x):
        """col for an array of x, with the same operations."""
        frac = (np.asarray(x, float) - self.x_min) * self.inv_dx
        li = np.ceil(frac)
        lf = li - frac
        low, high = li < 1, li >= self.n_rows
        li = np.where(low, 1, np.where(high, self.n_rows - 1, li)).astype(np.int64)
        lf = np.where(low, 1.0, np.where(high, 0.0, lf))
        t = self.rc[c - 1]
        return lf * t[li - 1] + (1 - lf) * t[li]


# per-driver constants (the lines in each src/streamer_*.f90 that differ)
VARIANTS = {
    "2d": dict(dim=2, coord=0, n_cycle_base=8, init_window=2.0e-9, crv_refine=1.0,
               exit_on_rm=False, time_in_substep=False, dielectric_rule=True, bc_top=BC_2V),
    "3d": dict(dim=3, coord=0, n_cycle_base=8, init_window=2.5e-9, crv_refine=0.1,
               exit_on_rm=True, time_in_substep=False, dielectric_rule=True, bc_top=BC_2V),
    "cyl": dict(dim=2, coord=1, n_cycle_base=4, init_window=5.0e-9, crv_refine=0.0,
                exit_on_rm=False, time_in_substep=True, dielectric_rule=False,
                bc_top=BC_ROW_2V),
}


class Config:
    """Run parameters the drivers read from their config files
    (ST_create_cfg, src/m_streamer.f90:145-228). The defaults here are the 2D
    air streamer: 32 mm square domain of 8x8-cell boxes on a 4x4-box base
    (1 mm cells), 2.5 MV/m, a Gaussian seed of peak 1e20 m^-3 and width 0.4 mm
    at (16, 28) mm over 1e15 m^-3, no dielectric, 9 levels (3.9 um)."""

    def __init__(self, variant="2d", **kw):
        if variant not in VARIANTS:
            raise KeyError("variant must be one of %s" % sorted(VARIANTS))
        self.variant = variant
        self.domain_len = 32e-3
        self.box_size = 8
        self.n_base = 4                # base boxes per direction, int or tuple;
        #                                domain_len spans the first direction
        self.lvls_max = 9
        self.applied_fld = 2.5e6
        # time dependence of the applied field (ST_get_fld,
        # src/m_streamer.f90:419-432; defaults :167-174)
        self.fld_mod_t0 = 1.0e99
        self.fld_sin_amplitude = 0.0
        self.fld_sin_freq = 0.2e9
        self.fld_lin_deriv = 0.0
        self.bg_dens = 1e15
        self.seed_dens = [1e20]
        self.seed_r0 = [[16e-3, 28e-3]]
        self.seed_r1 = [[16e-3, 28e-3]]
        self.seed_width = [0.4e-3]
        self.seed_falloff = [2]
        self.dt_max = 1e-8
        self.steps_amr = 2
        self.epsilon_diel = 1.0
        self.photoi = False            # Helmholtz photoionization
        self.photoi_eta = 0.05         # src/m_streamer.f90:196-203
        self.photoi_frac_o2 = 0.2
        self.gas_pressure = 1.0        # bar, src/m_streamer.f90:160-161
        self.photoi_tol = 1e-6         # relative residual of each term's solve
        self.photoi_warm = True        # V-cycles from the last step's solutions
        self.table_size = 1000
        self.table_max_fld = 3e7
        self.n_boxes = 1000
        self.transport = None          # (x, [mob, dif, alpha, eta]) or None
        # Driver variant for large runs (not in the reference's drivers):
        # boxes whose centre lies in [lo, hi] are refined while their level
        # is below lvl, for each (lvl, lo, hi) of refine_regions, on top of
        # set_ref_flags' criteria
        self.refine_regions = []
        # Initial densities evaluated on the device (a5g_set_gaussians: the
        # Gaussian seeds as exp(-|r - r0|^2 / w^2) with a plain sum of
        # squares, equal to GM_gaussian of set_init_cond within rounding)
        # instead of on the host in the reference's arithmetic; point seeds
        # with falloff 2 and no dielectric only
        self.init_on_device = False
        # The variables of the reverse mode (adjoint.FrozenWindow): eight
        # cell and two face variables after the run's own
        self.adjoint = False
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError("unknown config key %s" % k)
            setattr(self, k, v)

    def get_fld(self, time):
        """ST_get_fld (src/m_streamer.f90:419-432): the applied field at a
        time, with its linear and sinusoidal modulation after fld_mod_t0."""
        if time > self.fld_mod_t0:
            return (self.applied_fld + (time - self.fld_mod_t0) * self.fld_lin_deriv
                    + self.fld_sin_amplitude * math.sin(
                        (time - self.fld_mod_t0) * 2 * UC_PI * self.fld_sin_freq))
        return self.applied_fld

    def photoi_model(self):
        """(coeff, amp, kappa) of the Helmholtz photoionization: coeff = eta
        times the quenching factor of set_photoionization
        (src/streamer_2d.f90:551-560), amp_j = A_j p_O2^2 (1/m^2) and kappa_j =
        lambda_j p_O2 (1/m) with p_O2 in Torr."""
        p_quench = 30.0 * UC_TORR_TO_BAR
        coeff = self.photoi_eta * (p_quench / (self.gas_pressure + p_quench))
        p_o2 = self.photoi_frac_o2 * self.gas_pressure / UC_TORR_TO_BAR
        amp = [a * p_o2 ** 2 * 1e4 for a in HELMHOLTZ_A]
        kappa = [lam * p_o2 * 1e2 for lam in HELMHOLTZ_LAMBDA]
        return coeff, amp, kappa

    def transport_columns(self):
        """Input rows of the transport table. The default is the air model of
        the 2D benchmark: mobility 0.03 m^2/Vs, diffusion 0.2 m^2/s,
        alpha(E) = 4.3e5 exp(-2e7/E) 1/m, eta = 0, given as 500 rows on
        0..3e7 V/m (the reader's row limit, src/m_transport_data.f90:9)."""
        if self.transport is not None:
            return self.transport
        x = np.array([self.table_max_fld * k / 499 for k in range(500)])
        alpha = np.array([4.3e5 * math.exp(-2e7 / v) if v > 0 else 0.0 for v in x])
        return x, [np.full(500, 0.03), np.full(500, 0.2), alpha, np.zeros(500)]


def base_grid(dim, n):
    """ix and neighbour lists of a block of base boxes, n per direction (an
    int, or one count per direction), physical boundaries (-1) outside; ids
    in lexicographic order, x fastest."""
    shape = [n] * dim if np.isscalar(n) else list(n)
    ix, idx = [], {}
    for p in np.ndindex(*reversed(shape)):
        q = tuple(reversed(p))
        idx[q] = len(ix) + 1
        ix.append([c + 1 for c in q])
    nbl = []
    for q in (tuple(c - 1 for c in r) for r in ix):
        row = []
        for d in range(dim):
            for s in (-1, 1):
                o = list(q)
                o[d] += s
                row.append(idx.get(tuple(o), -1))
        nbl.append(row)
    return np.array(ix, np.int32), np.array(nbl, np.int32)


class StreamerRun:
    """One streamer simulation on a tree made by ``tree_cls`` (the HIP
    library's Tree, or any class with its constructor and methods)."""

    def __init__(self, tree_cls, cfg):
        c = cfg
        v = VARIANTS[c.variant]
        self.c, self.v = c, v
        self.dim = v["dim"]
        self.time = 0.0
        self.dt = 0.0
        self.step_no = 0
        # init_tree: box_size cells per base box; one base box of the whole
        # domain uses coarsen_to=2 as the drivers do
        n0 = c.n_base if np.isscalar(c.n_base) else c.n_base[0]
        dr = c.domain_len / (c.box_size * n0)
        n_cc = N_VAR_CELL + (3 * len(I_PH_TERMS) if c.photoi else 0)
        self.n_cc_run = n_cc
        n_fc = N_VAR_FACE
        if c.adjoint:
            n_cc += N_ADJ_CELL
            n_fc += N_ADJ_FACE
        self.photoi_solved = False     # the term variables hold solutions
        self.photoi_cycles = 0         # V-cycles of the photoionization solves
        self.T = T = tree_cls(self.dim, c.box_size, n_cc, n_fc, dr,
                              coarsen_to=2, lvls_max=c.lvls_max, coord=v["coord"],
                              n_boxes=c.n_boxes)
        T.set_base(*base_grid(self.dim, c.n_base))
        # ST_load_transport_data (src/m_streamer.f90:273-333)
        T.table_create(0.0, c.table_max_fld, c.table_size, 4)
        x, cols = c.transport_columns()
        for k, y in enumerate(cols, start=1):
            T.table_set_col(k, x, y)
        self.tbl = Table(*T.table_get())
        T.streamer_setup(dielectric_rule=v["dielectric_rule"],
                         pho=I_PHO if c.photoi else -1)
        # sides_bc_pot: Neumann sides, grounded at the low end of the last
        # axis, 2V - phi at its high end
        self.voltage = -c.domain_len * c.applied_fld
        kinds = [BC_COPY] * (2 * self.dim)
        kinds[-2], kinds[-1] = BC_NEG, v["bc_top"]
        vals = [0.0] * (2 * self.dim)
        vals[-1] = self.voltage
        T.mg_setup(I_PHI, I_FLD, I_RHS, i_eps=I_EPS, n_base=v["n_cycle_base"],
                   bc_kind=np.array(kinds, np.int32), bc_val=np.array(vals),
                   auto_ops=True)
        # the refinement flags on the device when the tree has them
        # (a5g_ref_flags: the same rule, each flag the same)
        self.device_flags = hasattr(T, "set_ref_rules")
        if self.device_flags:
            D = self.dim
            seeds = np.zeros((1, len(c.seed_dens), 2 * D + 1))
            for k in range(len(c.seed_dens)):
                seeds[0, k, :D] = c.seed_r0[k]
                seeds[0, k, D:2 * D] = c.seed_r1[k]
                seeds[0, k, 2 * D] = c.seed_width[k]
            regions = [[lv] + list(lo) + list(hi) for lv, lo, hi in c.refine_regions]
            T.set_ref_rules(seeds, np.array(regions, float).reshape(-1, 1 + 2 * D),
                            v["init_window"], v["crv_refine"], 1)

    # --- set_init_cond / set_box_eps (src/streamer_2d.f90:218-268) -----------
    def _eps_at(self, r):
        c = self.c
        if self.c.variant == "3d":
            x = r[0] / c.domain_len
            return c.epsilon_diel if 0.375 < x < 0.5 else 1.0
        k = 0 if c.variant == "cyl" else 1
        return c.epsilon_diel if r[k] / c.domain_len < 0.25 else 1.0

    def _cell_coords(self, nc, dr, rmin):
        n = nc + 2
        axes = [rmin[d] + (np.arange(n) - 0.5) * dr for d in range(self.dim)]
        # box layout: x fastest
        return [tuple(axes[d][p[self.dim - 1 - d]] for d in range(self.dim))
                for p in np.ndindex(*([n] * self.dim))]

    def _far_from_seeds(self, nc, dr, rmin):
        """True when every cell centre of the box (ghosts included) lies more
        than 30 widths from every Gaussian seed segment: exp(-(d/w)^2) is then
        exactly 0 in double precision (it underflows below d/w = 27.3), so each
        density is exactly bg_dens."""
        c = self.c
        lo = np.asarray(rmin) - 0.5 * dr
        hi = np.asarray(rmin) + (nc + 0.5) * dr
        for k in range(len(c.seed_dens)):
            if c.seed_falloff[k] != 2:
                return False
            s_lo = np.minimum(c.seed_r0[k], c.seed_r1[k])
            s_hi = np.maximum(c.seed_r0[k], c.seed_r1[k])
            gap = np.maximum(0.0, np.maximum(s_lo - hi, lo - s_hi))
            if math.sqrt(float(np.sum(gap * gap))) <= 30 * c.seed_width[k]:
                return False
        return True

    def _box_fields(self, b, i, want_dens):
        c = self.c
        nc, dr, rmin = int(b["n_cell"][i]), b["dr"][i], b["r_min"][i]
        n_all = (nc + 2) ** self.dim
        if c.epsilon_diel == 1.0:
            eps = np.ones(n_all)
        else:
            eps = np.array([self._eps_at(r) for r in self._cell_coords(nc, dr, rmin)])
        if not want_dens:
            return None, eps
        if self._far_from_seeds(nc, dr, rmin):
            return np.full(n_all, float(c.bg_dens)), eps
        pts = self._cell_coords(nc, dr, rmin)
        elec = np.empty(len(pts))
        for q, r in enumerate(pts):
            val = c.bg_dens
            for k in range(len(c.seed_dens)):
                val = val + c.seed_dens[k] * dens_line(
                    r, c.seed_r0[k], c.seed_r1[k], c.seed_width[k], c.seed_falloff[k])
            elec[q] = val
        return elec, eps

    def _init_on_device(self):
        c, T = self.c, self.T
        if c.epsilon_diel != 1.0 or any(f != 2 for f in c.seed_falloff) or \
                any(list(a) != list(b) for a, b in zip(c.seed_r0, c.seed_r1)):
            raise ValueError("init_on_device: point seeds with falloff 2 and no dielectric only")
        blobs = [list(c.seed_r0[k]) + [c.seed_width[k], c.seed_dens[k]]
                 for k in range(len(c.seed_dens))]
        T.set_gaussians(I_ELEC, np.array(blobs, float), c.bg_dens)
        T.tree_copy_cc(I_ELEC, I_PION)
        b = T.boxes()
        ids = np.nonzero(b["in_use"])[0] + 1
        T.fill_cc(ids, I_PHI, 0.0)
        T.fill_cc(ids, I_EPS, 1.0)

    def set_init_cond(self):
        """set_init_cond (src/streamer_2d.f90:218-248) on every box in use.
        A tree with whole-variable uploads (set_cc_all) gets one upload per
        variable; box slots are box_len(n_cell) long, a coarser box filling
        the start of its slot."""
        T = self.T
        if self.c.init_on_device:
            self._init_on_device()
            return
        b = T.boxes()
        n = len(b["lvl"])
        whole = hasattr(T, "set_cc_all")
        if whole:
            bl = T.box_len(T.n_cell)
            elec_all, eps_all = np.zeros((n, bl)), np.zeros((n, bl))
        for i in range(n):
            if not b["in_use"][i]:
                continue
            elec, eps = self._box_fields(b, i, True)
            if whole:
                elec_all[i, :len(elec)] = elec
                eps_all[i, :len(eps)] = eps
                continue
            T.set_cc(i + 1, I_ELEC, elec)
            T.set_cc(i + 1, I_PION, elec)
            T.set_cc(i + 1, I_PHI, np.zeros_like(elec))
            T.set_cc(i + 1, I_EPS, eps)
        if whole:
            T.set_cc_all(I_ELEC, elec_all)
            T.set_cc_all(I_PION, elec_all)
            T.set_cc_all(I_PHI, np.zeros((n, bl)))
            T.set_cc_all(I_EPS, eps_all)

    def _set_eps(self, ids):
        if self.c.epsilon_diel == 1.0:
            self.T.fill_cc(ids, I_EPS, 1.0)
            return
        b = self.T.boxes()
        for i in ids:
            self.T.set_cc(int(i), I_EPS, self._box_fields(b, int(i) - 1, False)[1])

    # --- compute_fld (src/streamer_2d.f90:307-343) ---------------------------
    def compute_fld(self, no_guess):
        T = self.T
        T.space_charge(ELEC / EPS0)
        # ST_applied_voltage = -ST_domain_len * ST_get_fld(ST_time)
        voltage = -self.c.domain_len * self.c.get_fld(self.time)
        if voltage != self.voltage:
            self.voltage = voltage
            T.mg_set_bc_value(2 * self.dim, voltage)
        for i in range(N_FMG_CYCLES):
            T.fas_fmg(False, no_guess and i == 0)
        T.fld_from_pot()
        T.gc_tree(I_FLD, RB_INTERP, BC_NEUMANN)

    # --- set_ref_flags (src/streamer_2d.f90:177-216) ------------------------
    def box_flag(self, time, summary, dr, nc, rmin):
        """The refinement flag of one box from its maxima (dr^2 max|rhs|,
        max |E|), its size and position."""
        c, v = self.c, self.v
        crv_phi, max_fld = summary[0], summary[1]
        adx = dr * self.tbl.col(T_ALPHA, max_fld)
        f = 0
        if adx < 0.1 and dr < 2.0e-5:
            f = -1
        if adx < 0.1 and crv_phi < 4.0 and dr < 5.0e-5:
            f = -1
        if time < v["init_window"]:
            boxlen = nc * dr
            centre = [rmin[k] + 0.5 * nc * dr for k in range(self.dim)]
            for k in range(len(c.seed_dens)):
                dist = dist_line(centre, c.seed_r0[k], c.seed_r1[k])
                if dist - c.seed_width[k] < boxlen and dr > 0.2 * c.seed_width[k]:
                    f = 1
        if adx > 1.0 and crv_phi > v["crv_refine"]:
            f = 1
        return f

    def flags_rows(self, time, s, dr, nc, rmin, ok, lvl=None):
        """box_flag of many boxes of this run at once, with array forms of
        the same operations (each flag the same): s the per-box maxima, dr,
        nc, rmin (n, dim) the boxes, ok the boxes to flag (others keep)."""
        c, v = self.c, self.v
        crv_phi, max_fld = s[:, 0], s[:, 1]
        adx = dr * self.tbl.col_rows(T_ALPHA, max_fld)
        f = np.zeros(len(dr), np.int32)
        f[(adx < 0.1) & (dr < 2.0e-5)] = -1
        f[(adx < 0.1) & (crv_phi < 4.0) & (dr < 5.0e-5)] = -1
        if time < v["init_window"] and ok.any():
            idx = np.nonzero(ok)[0]
            boxlen = nc[idx] * dr[idx]
            centre = np.stack([rmin[idx, k] + 0.5 * nc[idx] * dr[idx] for k in range(self.dim)],
                              axis=1)
            m = len(idx)
            for k in range(len(c.seed_dens)):
                hit = seed_hits(centre, boxlen, dr[idx],
                                np.broadcast_to(np.asarray(c.seed_r0[k], float), (m, self.dim)),
                                np.broadcast_to(np.asarray(c.seed_r1[k], float), (m, self.dim)),
                                np.full(m, float(c.seed_width[k])))
                f[idx[hit]] = 1
        f[(adx > 1.0) & (crv_phi > v["crv_refine"])] = 1
        for lvl_r, lo, hi in c.refine_regions:
            centre = np.stack([rmin[:, k] + 0.5 * nc * dr for k in range(self.dim)], axis=1)
            inside = np.all((centre >= np.asarray(lo)) & (centre <= np.asarray(hi)), axis=1)
            f[inside & (np.asarray(lvl) < lvl_r)] = 1
        f[~ok] = 0
        return f

    def ref_flags(self):
        T = self.T
        if self.device_flags:
            return T.ref_flags([self.time])
        b = T.boxes()
        s = T.box_summaries()
        ok = (b["in_use"] != 0) & (b["lvl"] >= 1)
        return self.flags_rows(self.time, s, b["dr"], b["n_cell"], b["r_min"], ok, b["lvl"])

    def adjust(self):
        return self.T.adjust(self.ref_flags())

    # --- prolong_to_new_boxes (src/streamer_2d.f90:592-617) ------------------
    def prolong_to_new_boxes(self):
        T = self.T
        for lvl in range(1, T.info()["max_lvl"] + 1):
            add, _ = T.ref_info_lvl(lvl)
            if len(add) == 0:
                continue
            for iv in (I_ELEC, I_PION, I_PHI):
                T.prolong1_to(add, iv)
            if self.c.photoi:
                # the next photoionization solves start from these
                for iv in I_PH_TERMS:
                    T.prolong1_to(add, iv)
            self._set_eps(add)
            T.gc_boxes(add, I_ELEC, RB_INTERP_LIM, BC_NEUMANN)
            T.gc_boxes(add, I_PION, RB_INTERP_LIM, BC_NEUMANN)
            T.gc_boxes(add, I_PHI, RB_MG, BC_MG)

    # --- get_max_dt (src/streamer_2d.f90:271-303) ----------------------------
    def max_dt(self):
        T, t = self.T, self.tbl
        min_fld, max_fld = T.tree_minmax(I_FLD)
        dr_min = T.min_dr()
        mobility = t.col(T_MOB, max_fld)
        diff_coeff = t.col(T_DIF, max_fld)
        alpha = t.col(T_ALPHA, max_fld)
        dt_cfl = dr_min / (mobility * max_fld)
        dt_dif = 0.25 * (dr_min * dr_min) / diff_coeff
        # dt_drt (dielectric relaxation) is computed by the drivers but not
        # part of the minimum
        dt_alpha = 1 / max(mobility * max_fld * alpha, EPS_F64)
        return 0.5 * min(1 / (1 / dt_cfl + 1 / dt_dif), dt_alpha, self.c.dt_max)

    # --- the run ------------------------------------------------------------
    def set_photoionization(self):
        """set_photoionization (src/streamer_2d.f90:544-565) with the
        Helmholtz model in place of Monte-Carlo photons: the production rate
        of set_photoi_rate, then one multigrid solve per term in variables
        of its own (the library runs the terms side by side). After the
        first call each solve starts from the term's solution of the call
        before (Config.photoi_warm): V-cycles only, no FMG. Returns the
        V-cycles run."""
        coeff, amp, kappa = self.c.photoi_model()
        warm = self.c.photoi_warm and self.photoi_solved
        cyc = self.T.photoi_helmholtz_terms(coeff, np.array(amp), np.array(kappa), I_PH_TERMS,
                                            I_PH_RHS, I_PH_TMP, warm, self.c.photoi_tol, 40)
        self.photoi_solved = True
        self.photoi_cycles += cyc
        return cyc

    def initialise(self, max_iter=50):
        """Initial conditions, field and refinement until the mesh settles
        (src/streamer_2d.f90:57-66)."""
        for _ in range(max_iter):
            self.set_init_cond()
            self.compute_fld(True)
            n_add, n_rm = self.adjust()
            if n_add == 0 and (n_rm == 0 or not self.v["exit_on_rm"]):
                if self.c.photoi:
                    self.set_photoionization()
                return
        raise RuntimeError("initial refinement did not settle")

    def _euler_pair(self, dt, cyl_time):
        T = self.T
        for i in (1, 2):
            if cyl_time:
                self.time = self.time + dt
            T.fluxes_koren(dt)
            T.consistent_fluxes([F_ELEC])
            T.update_solution(dt)
            T.restrict_tree(I_ELEC)
            T.restrict_tree(I_PION)
            T.gc_tree(I_ELEC, RB_INTERP_LIM, BC_NEUMANN)
            T.gc_tree(I_PION, RB_INTERP_LIM, BC_NEUMANN)
            if i == 1:
                self.compute_fld(False)
        if cyl_time:
            self.time = self.time - dt

    def step_block(self):
        """One pass of the outer loop: dt, steps_amr time steps, then the mesh
        update (src/streamer_2d.f90:68-142). Returns (n_add, n_rm)."""
        T = self.T
        self.dt = self.max_dt()
        if self.dt < 1e-14:
            raise RuntimeError("dt getting too small, instability?")
        dt = self.dt
        cyl_time = self.v["time_in_substep"]
        for _ in range(self.c.steps_amr):
            if not cyl_time:
                self.time = self.time + dt
            if self.c.photoi:
                self.set_photoionization()
            T.tree_copy_cc(I_ELEC, I_ELEC_OLD)
            T.tree_copy_cc(I_PION, I_PION_OLD)
            self._euler_pair(dt, cyl_time)
            T.average_dens()
            self.compute_fld(False)
            self.step_no += 1
        n_add, n_rm = self.adjust()
        if n_add > 0 or n_rm > 0:
            self.prolong_to_new_boxes()
            self.compute_fld(False)
            T.tidy_up(0.9, 0.25, 5000, False)
        return n_add, n_rm

    # --- output diagnostics (src/streamer_cyl.f90:647-956) --------------------
    def _cell(self, loc, iv):
        """tree%boxes(loc%id)%cc(loc%ix(1), loc%ix(2), iv): index 0 is the
        ghost layer, which a$D_get_loc returns on the low face of a box."""
        i, ix = loc
        nc = self.c.box_size
        return self.T.get_cc(i, iv, nc).reshape([nc + 2] * self.dim, order="F")[tuple(ix)]

    def _r_loc(self, loc, b):
        """a$D_r_loc / a$D_r_cc (afivo/src/m_aX_utils.f90:309-323)"""
        i, ix = loc
        return b["r_min"][i - 1] + (np.asarray(ix) - 0.5) * b["dr"][i - 1]

    def streamer_properties(self):
        """get_streamer_properties (src/streamer_cyl.f90:682-828) of the
        cylindrical run: 39 named values from located maxima of the field
        (on the GPU, a5g_reduction_loc) and point lookups. Two behaviours of
        the reference are kept: the mobility of the S values is read from
        the alpha column (:727-728, :736-737), and a location that no box
        improves on keeps its previous value (an intent(out) a2_loc_t
        without default initialisation, afivo/src/m_aX_utils.f90:502-553).
        Raises where the reference stops: the head (the location of max E_r)
        below 0.1 of the domain, or above 0.9 of it after 1 ns."""
        if self.c.variant != "cyl":
            raise ValueError("streamer properties are defined for the cylindrical run")
        T, c, tb = self.T, self.c, self.tbl
        b = T.boxes()
        L = c.domain_len
        names, props = [], []

        def add(n, v):
            names.append(n)
            props.append(float(v))

        add("E_bg", c.get_fld(self.time))
        ez_max, *loc_ez = T.reduction_loc(1, F_FLD, 2, 0.0, -1.0e99)
        add("Ez_max", ez_max)
        er_max, *loc_er = T.reduction_loc(1, F_FLD, 1, 0.0, -1.0e99)
        add("Er_max", er_max)
        rz = self._r_loc(loc_er, b)
        add("Er_max(r)", rz[0])
        add("Er_max(z)", rz[1])
        z_head = rz[1]
        er_max_norm = self._cell(loc_er, I_FLD)
        add("Er_max(E)", er_max_norm)
        for tag, e in (("Ez_max", ez_max), ("Er_max", er_max_norm)):
            alpha = tb.col(T_ALPHA, e)
            mu = tb.col(T_ALPHA, e)
            add(tag + "(alpha)", alpha)
            add(tag + "(S)", alpha * mu * e)
        if (self.time > 1.0e-9 and rz[1] > 0.9 * L) or rz[1] < 0.1 * L:
            raise RuntimeError("Simulation has reached boundary")
        loc_dens = T.get_loc([0.0, rz[1]])
        add("n_e(head)", self._cell(loc_dens, I_ELEC))
        phi_head = self._cell(loc_dens, I_PHI)
        add("dphi", phi_head - (rz[1] / L) * self.voltage)
        rz = self._r_loc(loc_ez, b)
        add("Ez_max(z)", rz[1])
        loc = None
        for dz, n in ((2.0e-3, "dphi(2mm)"), (4.0e-3, "dphi(4mm)")):
            loc = T.get_loc([0.0, rz[1] + dz])
            add(n, phi_head - self._cell(loc, I_PHI) if loc[0] > 0 else 0.0)
        for i in range(4, 15, 2):
            thr = i * 1e6
            v, *new = T.reduction_loc(2, I_FLD, 1, thr, 0.0)
            loc = tuple(new) if new[0] != -1 else loc
            add("r_max(%de6)" % i, v)
            add("dphi_r(%de6)" % i, phi_head - self._cell(loc, I_PHI) if loc[0] > 0 else 0.0)
            v, *new = T.reduction_loc(2, I_FLD, 2, thr, 0.0)
            loc = tuple(new) if new[0] != -1 else loc
            add("z_max(%de6)" % i, v - z_head)
            add("dphi_z(%de6)" % i, phi_head - self._cell(loc, I_PHI) if loc[0] > 0 else 0.0)
        return names, np.array(props)

    def cc_axis(self):
        """get_cc_axis(tree, [i_elec, i_pion], [f_fld, f_elec])
        (src/streamer_cyl.f90:659,905-956): rows z, n_e, n_i, E_z, flux_z."""
        return self.T.cc_axis([I_ELEC, I_PION], [F_FLD, F_ELEC])

    def write_streamer_properties(self, fname_stats, fname_axis, out_cnt):
        """write_streamer_properties (src/streamer_cyl.f90:647-680): the
        first call writes the header line of the statistics file, later
        calls append (out_cnt, time, properties); the axis file is rewritten
        with one point per line. Numbers are written with 17 significant
        digits (the reference uses list-directed output)."""
        names, props = self.streamer_properties()
        axis = self.cc_axis()
        if out_cnt == 1:
            with open(fname_stats, "w") as f:
                f.write(" ix_out      time     " + "".join(n.ljust(15) for n in names) + "\n")
        else:
            with open(fname_stats, "a") as f:
                f.write(" %d %.17g " % (out_cnt, self.time)
                        + " ".join("%.17g" % v for v in props) + "\n")
        with open(fname_axis, "w") as f:
            for p in range(axis.shape[1]):
                f.write(" ".join("%.17g" % v for v in axis[:, p]) + "\n")
        return names, props

    def leaf_cells(self):
        info = self.T.info()
        n = sum(len(self.T.level(l)[1]) for l in range(1, info["max_lvl"] + 1))
        return n * self.c.box_size ** self.dim


class BatchRun:
    """Many independent streamer runs advanced together in one tree
    (a5g_set_instances): run g owns the base boxes stacked along the last
    axis in [g H, (g + 1) H), H the height of one run's domain, with no
    neighbour links between runs. Every library call then treats all runs
    in one launch, and the time step and applied voltage are taken per run.
    The state of each run is the state StreamerRun reaches alone:

    - the host evaluates initial conditions and refinement flags in each
      run's own coordinates, recomputed the way a tree holding only that run
      computes them (base position from its own ix, then the children's
      offsets), so every double is the same;
    - a run whose initial refinement has settled is left out of further
      initial iterations (a5g_set_active_instances) and keeps its mesh;
    - after a mesh update the field is solved again only for the runs whose
      mesh changed.

    The runs share the variant, domain, box size, base, level limit,
    transport data and steps per mesh update; field, seeds, background and
    dt_max may differ. Photoionization, a dielectric and the 3D driver's
    stop on removal are not supported here."""

    def __init__(self, tree_cls, cfgs):
        c0 = cfgs[0]
        for c in cfgs:
            for k in ("variant", "domain_len", "box_size", "n_base", "lvls_max",
                      "steps_amr", "table_size", "table_max_fld"):
                if getattr(c, k) != getattr(c0, k):
                    raise ValueError("runs of a batch must share %s" % k)
            if c.photoi or c.epsilon_diel != 1.0:
                raise ValueError("a batch has no photoionization and no dielectric")
        self.v = v = VARIANTS[c0.variant]
        if v["exit_on_rm"]:
            raise ValueError("the %s driver is not batched" % c0.variant)
        self.cfgs = cfgs
        self.K = K = len(cfgs)
        # field solves of runs with different finest levels side by side
        # (fas_fmg_groups) rather than one group after the other
        self.groups_side_by_side = True
        self.dim = D = v["dim"]
        shape = [c0.n_base] * D if np.isscalar(c0.n_base) else list(c0.n_base)
        self.n_last = shape[-1]
        self.dr_base = c0.domain_len / (c0.box_size * shape[0])
        self.H = self.n_last * c0.box_size * self.dr_base
        self.T = T = tree_cls(D, c0.box_size, N_VAR_CELL, N_VAR_FACE, self.dr_base, coarsen_to=2,
                              lvls_max=c0.lvls_max, coord=v["coord"], n_boxes=c0.n_boxes * K)
        ix1, nb1 = base_grid(D, c0.n_base)
        n1 = len(ix1)
        ix = np.concatenate([ix1 + np.array([0] * (D - 1) + [g * self.n_last]) for g in range(K)])
        nb = np.concatenate([np.where(nb1 > 0, nb1 + g * n1, nb1) for g in range(K)])
        T.set_base(ix.astype(np.int32), nb.astype(np.int32))
        # one StreamerRun per run holds its configuration and scalar state;
        # its tree is this shared one
        self.runs = []
        for c in cfgs:
            r = StreamerRun.__new__(StreamerRun)
            r.c, r.v, r.dim, r.T = c, v, D, T
            r.time, r.dt, r.step_no = 0.0, 0.0, 0
            r.voltage = -c.domain_len * c.applied_fld
            self.runs.append(r)
        T.table_create(0.0, c0.table_max_fld, c0.table_size, 4)
        x, cols = c0.transport_columns()
        for k, y in enumerate(cols, start=1):
            T.table_set_col(k, x, y)
        tbl = Table(*T.table_get())
        for r in self.runs:
            r.tbl = tbl
        T.streamer_setup(dielectric_rule=v["dielectric_rule"], pho=-1)
        kinds = [BC_COPY] * (2 * D)
        kinds[-2], kinds[-1] = BC_NEG, VARIANTS["cyl"]["bc_top"]
        T.mg_setup(I_PHI, I_FLD, I_RHS, i_eps=I_EPS, n_base=v["n_cycle_base"],
                   bc_kind=np.array(kinds, np.int32), bc_val=np.zeros(2 * D), auto_ops=True)
        T.set_instances(K, self.H)
        self.voltages = np.array([r.voltage for r in self.runs])
        T.set_instance_values(2 * D, self.voltages)
        self.active = np.ones(K, bool)
        # refinement flags and time-step maxima on the device when the tree
        # has them (a5g_ref_flags, a5g_instance_maxima) and every run shares
        # its regions; the array forms on the host otherwise
        self.device_flags = False
        regs = {repr(c.refine_regions) for c in cfgs}
        if hasattr(T, "set_ref_rules") and len(regs) == 1:
            n_seed = {len(c.seed_dens) for c in cfgs}
            if len(n_seed) == 1:
                ns = n_seed.pop()
                seeds = np.zeros((K, ns, 2 * D + 1))
                for g, c in enumerate(cfgs):
                    for k in range(ns):
                        seeds[g, k, :D] = c.seed_r0[k]
                        seeds[g, k, D:2 * D] = c.seed_r1[k]
                        seeds[g, k, 2 * D] = c.seed_width[k]
                regions = [[lv] + list(lo) + list(hi) for lv, lo, hi in c0.refine_regions]
                T.set_ref_rules(seeds, np.array(regions, float).reshape(-1, 1 + 2 * D),
                                v["init_window"], v["crv_refine"], self.n_last)
                self.device_flags = True

    # --- geometry -------------------------------------------------------------
    def _boxes(self):
        """boxes() plus each box's run and its r_min in the run's own
        coordinates (last axis recomputed as a one-run tree computes it:
        from the base box's own ix, then the children's offsets)."""
        b = self.T.boxes()
        D = self.dim
        use = b["in_use"] != 0
        g = np.floor((b["r_min"][:, D - 1] + 0.25 * b["dr"]) / self.H).astype(np.int64)
        g = np.clip(g, 0, self.K - 1)
        loc = b["r_min"].copy()
        lv = b["lvl"]
        for lvl in range(int(lv[use].min()) if use.any() else 1, int(lv[use].max()) + 1 if use.any() else 1):
            sel = np.nonzero(use & (lv == lvl))[0]
            if lvl <= 1:
                ixl = b["ix"][sel, D - 1].astype(np.int64) - g[sel] * self.n_last
                loc[sel, D - 1] = 0.0 + (ixl - 1) * self.dr_base * self.cfgs[0].box_size
            else:
                p = b["parent"][sel] - 1
                dix = (b["ix"][sel, D - 1] - 1) & 1
                loc[sel, D - 1] = loc[p, D - 1] + 0.5 * b["dr"][p] * dix * b["n_cell"][p]
        b["run"], b["r_loc"] = g, loc
        self.inst_lvl = group_max(g[use], lv[use].astype(np.int64), self.K, 0)
        return b

    def _set_active(self, mask):
        self.active = np.asarray(mask, bool)
        self.T.set_active_instances(self.active.astype(np.int32))

    # --- the driver steps, per run where they differ ------------------------
    def set_init_cond(self, b):
        """set_init_cond on the boxes of active runs: densities and eps are
        uploaded for every box (a settled run gets the values it holds), phi
        is cleared on the active runs' boxes only."""
        T = self.T
        if all(c.init_on_device for c in self.cfgs):
            self._init_on_device(b)
            return
        n = len(b["lvl"])
        bl = T.box_len(T.n_cell)
        elec_all, eps_all = np.zeros((n, bl)), np.zeros((n, bl))
        clear = []
        for i in range(n):
            if not b["in_use"][i]:
                continue
            r = self.runs[b["run"][i]]
            bb = {"n_cell": b["n_cell"], "dr": b["dr"], "r_min": b["r_loc"]}
            elec, eps = r._box_fields(bb, i, True)
            elec_all[i, :len(elec)] = elec
            eps_all[i, :len(eps)] = eps
            if self.active[b["run"][i]]:
                clear.append(i + 1)
        T.set_cc_all(I_ELEC, elec_all)
        T.set_cc_all(I_PION, elec_all)
        T.set_cc_all(I_EPS, eps_all)
        if clear:
            T.fill_cc(np.array(clear, np.int32), I_PHI, 0.0)

    def _init_on_device(self, b):
        """set_init_cond of the active runs on the device
        (a5g_set_gaussians_instances, Config.init_on_device): every run's
        seeds moved to its place in the tree, its own background."""
        T, D = self.T, self.dim
        n_seed = {len(c.seed_dens) for c in self.cfgs}
        if len(n_seed) != 1:
            raise ValueError("runs of a batch must have the same number of seeds")
        blobs = np.zeros((self.K, n_seed.pop(), D + 2))
        for g, c in enumerate(self.cfgs):
            if any(f != 2 for f in c.seed_falloff) or \
                    any(list(a) != list(e) for a, e in zip(c.seed_r0, c.seed_r1)):
                raise ValueError("init_on_device: point seeds with falloff 2 only")
            for k in range(len(c.seed_dens)):
                blobs[g, k, :D] = c.seed_r0[k]
                blobs[g, k, D - 1] += g * self.H
                blobs[g, k, D], blobs[g, k, D + 1] = c.seed_width[k], c.seed_dens[k]
        T.set_gaussians_instances(I_ELEC, blobs, [c.bg_dens for c in self.cfgs])
        T.tree_copy_cc(I_ELEC, I_PION)
        use = b["in_use"] != 0
        T.fill_cc((np.nonzero(use)[0] + 1).astype(np.int32), I_EPS, 1.0)
        clear = np.nonzero(use & self.active[b["run"]])[0] + 1
        if len(clear):
            T.fill_cc(clear.astype(np.int32), I_PHI, 0.0)

    def compute_fld(self, no_guess):
        """compute_fld of the active runs. The FMG of a run goes up to its own
        finest level, so it is called once per group of runs with the same
        finest level, that group active."""
        T, D = self.T, self.dim
        T.space_charge(ELEC / EPS0)
        volt = np.array([-r.c.domain_len * r.c.get_fld(r.time) for r in self.runs])
        if np.any(volt != self.voltages):
            self.voltages = volt
            T.set_instance_values(2 * D, volt)
        for r, vg in zip(self.runs, volt):
            r.voltage = vg
        active = self.active.copy()
        levels = sorted(set(self.inst_lvl[active].tolist()))
        if len(levels) > 1 and self.groups_side_by_side and hasattr(T, "fas_fmg_groups"):
            # one group per finest level, the groups' FMGs side by side
            group = np.full(self.K, -1, np.int32)
            for gi, m in enumerate(levels):
                group[active & (self.inst_lvl == m)] = gi
            T.set_instance_groups(group)
            for i in range(N_FMG_CYCLES):
                T.fas_fmg_groups(False, no_guess and i == 0)
        else:
            for m in levels:
                if len(levels) > 1:
                    self._set_active(active & (self.inst_lvl == m))
                for i in range(N_FMG_CYCLES):
                    T.fas_fmg(False, no_guess and i == 0)
            if len(levels) > 1:
                self._set_active(active)
        T.fld_from_pot()
        T.gc_tree(I_FLD, RB_INTERP, BC_NEUMANN)

    def adjust(self, b):
        """Refinement flags of the active runs (the others keep their mesh);
        returns the boxes added and removed per run."""
        T = self.T
        if self.device_flags:
            T.adjust(T.ref_flags([r.time for r in self.runs]))
        else:
            T.adjust(self.flags(b, T.box_summaries()))
        add_k, rm_k = np.zeros(self.K, int), np.zeros(self.K, int)
        nb = T.boxes()
        added = []
        for lvl in range(1, T.info()["max_lvl"] + 1):
            add, rm = T.ref_info_lvl(lvl)
            for i in add:
                gi = int(np.floor((nb["r_min"][i - 1][self.dim - 1] + 0.25 * nb["dr"][i - 1])
                                  / self.H))
                add_k[min(max(gi, 0), self.K - 1)] += 1
            for i in rm:
                rm_k[b["run"][i - 1]] += 1
            added.append((lvl, add))
        return add_k, rm_k, added

    def flags(self, b, s):
        """set_ref_flags of every box of the active runs (StreamerRun.box_flag
        with array operations over all runs at once, each flag the same), keep
        for the others. The boxes are split into chunks flagged on a thread
        pool (numpy releases the interpreter lock inside its array
        operations); every operation is per box, so chunking changes no flag."""
        v = self.v
        n_seed = {len(c.seed_dens) for c in self.cfgs}
        if len(n_seed) != 1:
            raise ValueError("runs of a batch must have the same number of seeds")
        n_seed = n_seed.pop()
        # per-run tables, gathered per box in the chunks
        times = np.array([r.time for r in self.runs])
        W = np.array([[c.seed_width[k] for k in range(n_seed)] for c in self.cfgs], float)
        R0 = np.array([[c.seed_r0[k] for k in range(n_seed)] for c in self.cfgs], float)
        R1 = np.array([[c.seed_r1[k] for k in range(n_seed)] for c in self.cfgs], float)
        # refine_regions: runs with the same regions share a region group
        reg_of, regions = np.full(self.K, -1, np.int64), []
        seen = {}
        for gi, c in enumerate(self.cfgs):
            if c.refine_regions:
                key = repr(c.refine_regions)
                if key not in seen:
                    seen[key] = len(regions)
                    regions.append(c.refine_regions)
                reg_of[gi] = seen[key]
        tbl = self.runs[0].tbl
        D = self.dim
        f = np.zeros(len(b["dr"]), np.int32)

        def rows(lo, hi):
            g = b["run"][lo:hi]
            dr, nc, lvl = b["dr"][lo:hi], b["n_cell"][lo:hi], b["lvl"][lo:hi]
            rl = b["r_loc"][lo:hi]
            crv_phi, max_fld = s[lo:hi, 0], s[lo:hi, 1]
            ok = (b["in_use"][lo:hi] != 0) & (lvl >= 1) & self.active[g]
            adx = dr * tbl.col_rows(T_ALPHA, max_fld)
            fc = np.zeros(hi - lo, np.int32)
            fc[(adx < 0.1) & (dr < 2.0e-5)] = -1
            fc[(adx < 0.1) & (crv_phi < 4.0) & (dr < 5.0e-5)] = -1
            idx = np.nonzero(ok & (times[g] < v["init_window"]))[0]
            for k in range(n_seed if len(idx) else 0):
                # seed k of each box's run, for the boxes coarse enough for
                # the rule (dr > 0.2 w)
                w = W[g[idx], k]
                pre = dr[idx] > 0.2 * w
                j = idx[pre]
                if len(j) == 0:
                    continue
                gj = g[j]
                centre = np.stack([rl[j, d] + 0.5 * nc[j] * dr[j] for d in range(D)], axis=1)
                fc[j[seed_hits(centre, nc[j] * dr[j], dr[j], R0[gj, k], R1[gj, k], w[pre])]] = 1
            fc[(adx > 1.0) & (crv_phi > v["crv_refine"])] = 1
            if regions:
                rg = reg_of[g]
                for ri, regs in enumerate(regions):
                    for lvl_r, rlo, rhi in regs:
                        # boxes of this region group below the region's level,
                        # then their centres against the region, one axis at
                        # a time
                        j = np.nonzero((rg == ri) & (lvl < lvl_r))[0]
                        for d in range(D):
                            if len(j) == 0:
                                break
                            cd = rl[j, d] + 0.5 * nc[j] * dr[j]
                            j = j[(cd >= rlo[d]) & (cd <= rhi[d])]
                        fc[j] = 1
            fc[~ok] = 0
            f[lo:hi] = fc

        _chunked(len(f), rows)
        return f

    def max_dt(self, b):
        """get_max_dt per run from the per-box maxima of its leaves."""
        dts = np.zeros(self.K)
        if self.device_flags and self.active.all():
            # the same maxima reduced per run on the device
            fld_k, lvl_k = self.T.instance_maxima(I_FLD, self.K)
        else:
            s = self.T.box_summaries()
            sel = (b["in_use"] != 0) & (b["lvl"] >= 1)
            leaf = sel & (b["children"][:, 0] == 0)
            fld_k = group_max(b["run"][leaf], s[leaf, 1], self.K, -np.inf)
            lvl_k = group_max(b["run"][sel], b["lvl"][sel].astype(np.int64), self.K, 0)
        for g, r in enumerate(self.runs):
            max_fld = float(fld_k[g])
            dr_min = self.dr_base * 0.5 ** (int(lvl_k[g]) - 1)
            t = r.tbl
            mobility = t.col(T_MOB, max_fld)
            diff_coeff = t.col(T_DIF, max_fld)
            alpha = t.col(T_ALPHA, max_fld)
            dt_cfl = dr_min / (mobility * max_fld)
            dt_dif = 0.25 * (dr_min * dr_min) / diff_coeff
            dt_alpha = 1 / max(mobility * max_fld * alpha, EPS_F64)
            dts[g] = 0.5 * min(1 / (1 / dt_cfl + 1 / dt_dif), dt_alpha, r.c.dt_max)
        return dts

    def initialise(self, max_iter=50):
        """The initial loop of every run (src/streamer_2d.f90:57-66); a run
        leaves it when its refinement adds no box."""
        settled = np.zeros(self.K, bool)
        for _ in range(max_iter):
            self._set_active(~settled)
            b = self._boxes()
            self.set_init_cond(b)
            self.compute_fld(True)
            add_k, _, _ = self.adjust(b)
            settled |= self.active & (add_k == 0)
            if settled.all():
                self._set_active(np.ones(self.K, bool))
                return
        raise RuntimeError("initial refinement did not settle")

    def prolong_to_new_boxes(self, added):
        T = self.T
        for lvl, add in added:
            if len(add) == 0:
                continue
            for iv in (I_ELEC, I_PION, I_PHI):
                T.prolong1_to(add, iv)
            T.fill_cc(add, I_EPS, 1.0)
            T.gc_boxes(add, I_ELEC, RB_INTERP_LIM, BC_NEUMANN)
            T.gc_boxes(add, I_PION, RB_INTERP_LIM, BC_NEUMANN)
            T.gc_boxes(add, I_PHI, RB_MG, BC_MG)

    def step_block(self):
        """One block for every run: its own dt, steps_amr steps, the mesh
        update. Returns (added, removed) boxes per run."""
        T = self.T
        b = self._boxes()
        dts = self.max_dt(b)
        if np.any(dts < 1e-14):
            raise RuntimeError("dt getting too small, instability?")
        for r, dt in zip(self.runs, dts):
            r.dt = dt
        T.set_instance_values(0, dts)
        cyl_time = self.v["time_in_substep"]
        for _ in range(self.cfgs[0].steps_amr):
            if not cyl_time:
                for r in self.runs:
                    r.time = r.time + r.dt
            T.tree_copy_cc(I_ELEC, I_ELEC_OLD)
            T.tree_copy_cc(I_PION, I_PION_OLD)
            for i in (1, 2):
                if cyl_time:
                    for r in self.runs:
                        r.time = r.time + r.dt
                T.fluxes_koren(1.0)
                T.consistent_fluxes([F_ELEC])
                T.update_solution(1.0)
                T.restrict_tree(I_ELEC)
                T.restrict_tree(I_PION)
                T.gc_tree(I_ELEC, RB_INTERP_LIM, BC_NEUMANN)
                T.gc_tree(I_PION, RB_INTERP_LIM, BC_NEUMANN)
                if i == 1:
                    self.compute_fld(False)
            if cyl_time:
                for r in self.runs:
                    r.time = r.time - r.dt
            T.average_dens()
            self.compute_fld(False)
            for r in self.runs:
                r.step_no += 1
        # the steps change no topology: b still describes the tree
        add_k, rm_k, added = self.adjust(b)
        changed = (add_k > 0) | (rm_k > 0)
        if changed.any():
            self.prolong_to_new_boxes(added)
            self._boxes()
            self._set_active(changed)
            self.compute_fld(False)
            self._set_active(np.ones(self.K, bool))
            T.tidy_up(0.9, 0.25, 5000, False)
        return add_k, rm_k

    def leaf_cells(self):
        """Leaf cells per run."""
        b = self._boxes()
        leaf = (b["in_use"] != 0) & (b["lvl"] >= 1) & (b["children"][:, 0] == 0)
        nc = self.cfgs[0].box_size ** self.dim
        return np.bincount(b["run"][leaf], minlength=self.K).astype(np.int64) * nc

    def total_leaf_cells(self):
        """Leaf cells of all runs together (no run attribution needed)."""
        b = self.T.boxes()
        leaf = (b["in_use"] != 0) & (b["lvl"] >= 1) & (b["children"][:, 0] == 0)
        return int(leaf.sum()) * self.cfgs[0].box_size ** self.dim
