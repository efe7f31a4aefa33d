"""Differentiation of the streamer's time steps on a frozen mesh, as
torch.autograd.Functions over the a5g C ABI: one transport substep
(TransportSubstep), and a window of whole time steps with checkpoints
(FrozenWindow, StreamerWindow).

One substep of the drivers' time loop is fluxes_koren, then
a$D_consistent_fluxes, then update_solution (src/streamer_2d.f90:105-110,
:384-542). With the mesh and the field held fixed it maps the densities n_e,
n_i and the transport table to the new densities. TransportSubstep runs it
on the library's tree in the forward pass and the hand-written reverse
kernels (a5g_update_solution_adjoint, a5g_consistent_fluxes_adjoint,
a5g_fluxes_koren_adjoint) in the backward pass; the fluxes the reverse
kernels need are recomputed from the saved inputs rather than stored
(recompute instead of storing the intermediate state).

Tensors are float64 on the host, in the layout of Tree.get_cc_all: one row
per box id (1..max_id), the box's cells ghosts included in Afivo's order; the
table is (n_cols, n_rows) as Tree.table_get gives it. The gradient of a
density includes its ghost cells and the cells the second ghost layer is
taken from; rows of boxes the substep does not change pass their gradient
through.

A window of whole steps: FrozenWindow steps a run of the 2D, 3D or
cylindrical driver (src/streamer_2d.f90:91-129; src/streamer_3d.f90 alike;
src/streamer_cyl.f90 advances the time inside the substeps, which a
constant field does not see) with its mesh and its time steps frozen, and runs the reverse of every operation of a step
in the opposite order with the hand-written reverse kernels
(a5g_*_adjoint, a5g_fas_fmg_adjoint). States are checkpointed in device
memory (a5g_state_store) every `every` steps; the reverse pass recomputes
each segment from its checkpoint once, keeping the three states of every
step that its reverse reads (the start, the state after the first
substep's field solve and the state after the second substep).
StreamerWindow maps the transport table and the applied voltage to the
final densities and field; window_loss is the loss of the adjoint
workload (integrated electron count plus the softmax-smoothed head
position).

No counterpart in the reference, which computes no derivatives.
"""
import math

import numpy as np

from . import streamer as S


class AdjointLayout:
    """Where a tree keeps the adjoint variables: cell variables le, li, lE,
    lpho and face variables lF, lfld (1-based, beyond the streamer's own)."""

    def __init__(self, le, li, lE, lpho, lF, lfld):
        self.le, self.li, self.lE, self.lpho, self.lF, self.lfld = le, li, lE, lpho, lF, lfld

    def array(self):
        return np.array([self.le, self.li, self.lE, self.lpho, self.lF, self.lfld], np.int32)


def _zero_adjoints(T, lay):
    """Every adjoint variable zero (cells of all boxes, faces of levels >= 1)."""
    info = T.info()
    n = info["max_id"]
    bl = T.box_len(T.n_cell)
    for iv in (lay.le, lay.li, lay.lE, lay.lpho):
        if iv > 0:
            T.set_cc_all(iv, np.zeros((n, bl)))
    b = T.boxes()
    for i in range(n):
        if b["in_use"][i] and b["lvl"][i] >= 1:
            nc = int(b["n_cell"][i])
            nf = (nc + 1) * nc ** (T.dim - 1)
            for d in range(1, T.dim + 1):
                for f in (lay.lF, lay.lfld):
                    T.set_fc(i + 1, d, f, np.zeros(nf))


def _forward_substep(T, dt):
    T.fluxes_koren(dt)
    T.consistent_fluxes([S.F_ELEC])
    T.update_solution(dt)


def _torch():
    import torch
    return torch


def transport_substep_function():
    """The torch.autograd.Function class (built on first use, so that
    importing this module does not import torch)."""
    torch = _torch()

    class TransportSubstep(torch.autograd.Function):
        @staticmethod
        def forward(ctx, ne, ni, table, tree, dt, layout):
            T = tree
            rc, x_min, inv_dx = T.table_get()
            T.set_table(table.detach().cpu().numpy(), x_min, inv_dx)
            T.set_cc_all(S.I_ELEC, ne.detach().cpu().numpy())
            T.set_cc_all(S.I_PION, ni.detach().cpu().numpy())
            _forward_substep(T, dt)
            out_e = torch.from_numpy(T.get_cc_all(S.I_ELEC))
            out_i = torch.from_numpy(T.get_cc_all(S.I_PION))
            ctx.save_for_backward(ne.detach(), ni.detach(), table.detach())
            ctx.T, ctx.dt, ctx.lay, ctx.tbl = T, dt, layout, (x_min, inv_dx)
            return out_e, out_i

        @staticmethod
        def backward(ctx, g_e, g_i):
            ne, ni, table = ctx.saved_tensors
            T, dt, lay = ctx.T, ctx.dt, ctx.lay
            # recompute the fluxes the update read, from the saved inputs
            T.set_table(table.cpu().numpy(), *ctx.tbl)
            T.set_cc_all(S.I_ELEC, ne.cpu().numpy())
            T.set_cc_all(S.I_PION, ni.cpu().numpy())
            T.fluxes_koren(dt)
            T.consistent_fluxes([S.F_ELEC])
            _zero_adjoints(T, lay)
            T.table_adjoint(clear=True)
            T.set_cc_all(lay.le, g_e.detach().cpu().numpy())
            T.set_cc_all(lay.li, g_i.detach().cpu().numpy())
            adj = lay.array()
            T.update_solution_adjoint(dt, adj)
            T.consistent_fluxes_adjoint(lay.lF)
            T.fluxes_koren_adjoint(dt, adj)
            gne = torch.from_numpy(T.get_cc_all(lay.le))
            gni = torch.from_numpy(T.get_cc_all(lay.li))
            gtab = torch.from_numpy(T.table_adjoint(clear=True))
            return gne, gni, gtab, None, None, None

    return TransportSubstep


def transport_substep(tree, dt, layout, ne, ni, table):
    """n_e', n_i' of one transport substep, differentiable in ne, ni and the
    table (torch float64 tensors)."""
    return transport_substep_function().apply(ne, ni, table, tree, dt, layout)


# ---------------------------------------------------------------------------
# a window of whole time steps
# ---------------------------------------------------------------------------

# the state a step reads and writes: cell and face variables
PRIMAL_CC = (S.I_ELEC, S.I_PION, S.I_ELEC_OLD, S.I_PION_OLD, S.I_PHI, S.I_FLD, S.I_RHS)
PRIMAL_FC = (S.F_ELEC, S.F_FLD)


class StepAdjointVars:
    """The reverse mode's variables of a run made with Config(adjoint=True):
    cell variables le, li (densities), leo, lio (the old densities), lphi,
    lE (|E|, which is also the multigrid's tmp), lrhs, lw (multigrid work)
    after the run's own, and face variables lF (flux), lfld (field)."""

    def __init__(self, run):
        b = run.n_cc_run
        (self.le, self.li, self.leo, self.lio, self.lphi, self.lE, self.lrhs,
         self.lw) = range(b + 1, b + 1 + S.N_ADJ_CELL)
        self.lF, self.lfld = S.N_VAR_FACE + 1, S.N_VAR_FACE + 2
        self.cells = tuple(range(b + 1, b + 1 + S.N_ADJ_CELL))
        self.faces = (self.lF, self.lfld)

    def transport(self):
        """adj[6] of the transport substep's reverse kernels (no
        photoionization source)."""
        return np.array([self.le, self.li, self.lE, -1, self.lF, self.lfld], np.int32)


class FrozenWindow:
    """n_steps time steps of a run on its current mesh, with the time steps
    dts (recorded by the first forward pass when None: max_dt every
    steps_amr steps, as the drivers take it), differentiable in the
    transport table, the applied voltage and the densities it starts from.

    The run (2D, 3D or cylindrical, with or without a dielectric) must have
    been made with Config(adjoint=True), without photoionization or a
    time-dependent field. The window starts from the run's state at
    construction every time it runs.
    Windows on one tree use the same store slots: a window made later
    replaces an earlier one's start state and checkpoints, so only the
    latest is live (reset() first to start the new one where the old one
    started)."""

    SLOT_START, SLOT_FINAL, SLOT_CKPT = 0, 1, 2

    def __init__(self, run, n_steps, every=10, dts=None, keep_all=False):
        c = run.c
        if not c.adjoint:
            raise ValueError("FrozenWindow: the run needs Config(adjoint=True)")
        if c.photoi or c.fld_mod_t0 < 1e98:
            raise ValueError("FrozenWindow: no photoionization or field modulation")
        if n_steps < 1 or every < 1:
            raise ValueError("FrozenWindow: n_steps and every must be positive")
        self.run, self.T = run, run.T
        # keep_all: the forward pass keeps the three states of every step in
        # device memory (3 n_steps whole states: 72 GB for 100 steps of 1.5M
        # cells, within one MI355X's 288 GB), and the reverse pass recomputes
        # nothing; one segment spans the window
        self.keep_all = bool(keep_all)
        self.n = int(n_steps)
        self.k = self.n if self.keep_all else int(every)
        self.n_seg = (self.n + self.k - 1) // self.k
        self.v = StepAdjointVars(run)
        self.dts = None if dts is None else [float(x) for x in dts]
        rc, self.x_min, self.inv_dx = self.T.table_get()
        self.table0 = rc.copy()
        self.voltage0 = run.voltage
        self.top = 2 * run.dim                 # the face of the rule 2V - c
        self._ckpt_key = None
        self.T.state_store(self.SLOT_START, PRIMAL_CC, PRIMAL_FC)

    def _slot_step(self, j, which):
        """Store slot of state S0, S1 or S2 (which 0, 1, 2) of step j of a
        segment."""
        return self.SLOT_CKPT + self.n_seg + 3 * j + which

    def store_slots(self):
        """Store slots the window uses (each holds the whole primal state)."""
        return self.SLOT_CKPT + self.n_seg + 3 * min(self.k, self.n)

    def _save(self, slot):
        self.T.state_store(slot, PRIMAL_CC, PRIMAL_FC)

    def _load(self, slot):
        self.T.state_store(slot, PRIMAL_CC, PRIMAL_FC, load=True)

    # --- forward ----------------------------------------------------------
    def _compute_fld(self):
        """compute_fld (src/streamer_2d.f90:307-343) at the window's fixed
        boundary value."""
        T = self.T
        T.space_charge(S.ELEC / S.EPS0)
        for _ in range(S.N_FMG_CYCLES):
            T.fas_fmg(False, False)
        T.fld_from_pot()
        T.gc_tree(S.I_FLD, S.RB_INTERP, S.BC_NEUMANN)

    def _substep(self, dt):
        T = self.T
        T.fluxes_koren(dt)
        T.consistent_fluxes([S.F_ELEC])
        T.update_solution(dt)
        T.restrict_tree(S.I_ELEC)
        T.restrict_tree(S.I_PION)
        T.gc_tree(S.I_ELEC, S.RB_INTERP_LIM, S.BC_NEUMANN)
        T.gc_tree(S.I_PION, S.RB_INTERP_LIM, S.BC_NEUMANN)

    def _step(self, dt, keep=None):
        """One time step of the drivers' loop (src/streamer_2d.f90:91-129);
        keep = (slot1, slot2) saves the states after the first substep's
        field solve (S1) and after the second substep (S2)."""
        T = self.T
        T.tree_copy_cc(S.I_ELEC, S.I_ELEC_OLD)
        T.tree_copy_cc(S.I_PION, S.I_PION_OLD)
        self._substep(dt)
        self._compute_fld()
        if keep:
            self._save(keep[0])
        self._substep(dt)
        if keep:
            self._save(keep[1])
        T.average_dens()
        self._compute_fld()

    def _set_params(self, table, voltage):
        T = self.T
        T.set_table(np.asarray(table, float), self.x_min, self.inv_dx)
        T.mg_set_bc_value(self.top, float(voltage))

    def reset(self):
        """Puts the run's tree back where the window started: its start
        state, and the table and voltage the window was made with. A pass
        at other parameters (forward(table, voltage)) leaves them in the
        tree, and a window made later takes the tree's as its defaults."""
        self._load(self.SLOT_START)
        self._set_params(self.table0, self.voltage0)

    def forward(self, table=None, voltage=None):
        """Runs the window from its start state with the transport table
        (n_cols, n_rows) and the applied voltage (the value V of the rule
        2V - c; defaults: the run's), keeping the checkpoints. The run's
        tree then holds the final state."""
        table = self.table0 if table is None else np.asarray(table, float)
        voltage = self.voltage0 if voltage is None else float(voltage)
        self._load(self.SLOT_START)
        self._set_params(table, voltage)
        record = self.dts is None
        dts = [] if record else self.dts
        amr = self.run.c.steps_amr
        for s in range(self.n):
            if record:
                dts.append(self.run.max_dt() if s % amr == 0 else dts[-1])
            if self.keep_all:
                self._save(self._slot_step(s, 0))
                self._step(dts[s], keep=(self._slot_step(s, 1), self._slot_step(s, 2)))
                continue
            if s % self.k == 0:
                self._save(self.SLOT_CKPT + s // self.k)
            self._step(dts[s])
        self._save(self.SLOT_FINAL)
        self.dts = dts
        self._ckpt_key = (table.tobytes(), voltage)
        return self

    # --- reverse ----------------------------------------------------------
    def _compute_fld_reverse(self):
        T, v = self.T, self.v
        T.gc_tree_adjoint(S.I_FLD, v.lE, S.RB_INTERP, S.BC_NEUMANN)
        T.fld_from_pot_adjoint(v.lE, v.lfld, v.lphi)
        vbar = 0.0
        for _ in range(S.N_FMG_CYCLES):
            vbar += T.fas_fmg_adjoint([v.lphi, v.lE, v.lrhs, v.lw])[self.top - 1]
        T.space_charge_adjoint(S.ELEC / S.EPS0, v.lrhs, v.le, v.li)
        return vbar

    def _transport_reverse(self, dt):
        """The reverse of fluxes_koren, consistent fluxes and update_solution
        on the state in the primal variables (the substep's start), whose
        fluxes are recomputed first."""
        T, adj = self.T, self.v.transport()
        T.fluxes_koren(dt)
        T.consistent_fluxes([S.F_ELEC])
        T.update_solution_adjoint(dt, adj)
        T.consistent_fluxes_adjoint(self.v.lF)
        T.fluxes_koren_adjoint(dt, adj)

    def _ghosts_restrict_reverse(self):
        T, v = self.T, self.v
        T.gc_tree_adjoint(S.I_ELEC, v.le, S.RB_INTERP_LIM, S.BC_NEUMANN)
        T.gc_tree_adjoint(S.I_PION, v.li, S.RB_INTERP_LIM, S.BC_NEUMANN)
        T.restrict_tree_adjoint(v.le)
        T.restrict_tree_adjoint(v.li)

    def _step_reverse(self, dt, s0, s1, s2):
        """The reverse of one step whose end state is in the primal
        variables and whose states S0 (start), S1 (after the first
        substep's field solve) and S2 (after the second substep) are in
        store slots s0, s1, s2: every operation's reverse, last first.
        Returns the voltage's share."""
        T, v = self.T, self.v
        vbar = self._compute_fld_reverse()                 # the last field solve
        T.average_dens_adjoint(v.le, v.leo, v.li, v.lio)
        self._load(s2)                                     # the second substep
        self._ghosts_restrict_reverse()
        self._load(s1)
        self._transport_reverse(dt)
        vbar += self._compute_fld_reverse()                # the first substep's field solve
        self._ghosts_restrict_reverse()
        self._load(s0)                                     # the first substep
        self._transport_reverse(dt)
        T.tree_copy_cc_adjoint(v.le, v.leo)                # the copies to the old densities
        T.tree_copy_cc_adjoint(v.li, v.lio)
        return vbar

    def backward(self, g_ne, g_E, table=None, voltage=None):
        """The gradient of a loss whose gradients with respect to the final
        n_e and |E| (arrays shaped as Tree.get_cc_all) are g_ne, g_E, at the
        given table and voltage (defaults: the run's). Returns (d table
        (n_cols, n_rows), d voltage, d n_e, d n_i of the start state)."""
        table = self.table0 if table is None else np.asarray(table, float)
        voltage = self.voltage0 if voltage is None else float(voltage)
        if self.dts is None or self._ckpt_key != (table.tobytes(), voltage):
            self.forward(table, voltage)
        T, v = self.T, self.v
        self._set_params(table, voltage)
        T.state_store(0, v.cells, v.faces, clear=True)
        T.set_cc_all(v.le, np.asarray(g_ne, float))
        T.set_cc_all(v.lE, np.asarray(g_E, float))
        T.table_adjoint(clear=True)
        vbar = 0.0
        for seg in reversed(range(self.n_seg)):
            lo, hi = seg * self.k, min(self.n, (seg + 1) * self.k)
            # recompute the segment from its checkpoint, keeping the three
            # states of every step that its reverse reads (kept by the
            # forward pass already with keep_all)
            if not self.keep_all:
                self._load(self.SLOT_CKPT + seg)
                for s in range(lo, hi):
                    j = s - lo
                    self._save(self._slot_step(j, 0))
                    self._step(self.dts[s], keep=(self._slot_step(j, 1), self._slot_step(j, 2)))
            for s in reversed(range(lo, hi)):
                j = s - lo
                # the step's end state: the next step's start, the next
                # segment's checkpoint or the window's final state
                if s + 1 < hi:
                    self._load(self._slot_step(j + 1, 0))
                elif seg + 1 < self.n_seg:
                    self._load(self.SLOT_CKPT + seg + 1)
                else:
                    self._load(self.SLOT_FINAL)
                vbar += self._step_reverse(self.dts[s], *(self._slot_step(j, w) for w in range(3)))
        g_table = T.table_adjoint(clear=True)
        return g_table, vbar, T.get_cc_all(v.le), T.get_cc_all(v.li)


def streamer_window_function():
    """The torch.autograd.Function of a FrozenWindow: (table, voltage,
    window) -> (n_e, |E|) of the final state, float64 tensors shaped as
    Tree.get_cc_all; its backward runs the window's reverse pass."""
    torch = _torch()

    class StreamerWindow(torch.autograd.Function):
        @staticmethod
        def forward(ctx, table, voltage, window):
            tb = table.detach().cpu().numpy().astype(float)
            vo = float(voltage.detach().cpu())
            window.forward(tb, vo)
            ne = torch.from_numpy(window.T.get_cc_all(S.I_ELEC))
            E = torch.from_numpy(window.T.get_cc_all(S.I_FLD))
            ctx.window, ctx.params = window, (tb, vo)
            return ne, E

        @staticmethod
        def backward(ctx, g_ne, g_E):
            tb, vo = ctx.params
            gt, gv, _, _ = ctx.window.backward(g_ne.detach().cpu().numpy(),
                                               g_E.detach().cpu().numpy(), tb, vo)
            return (torch.from_numpy(gt), torch.tensor(gv, dtype=torch.float64), None)

    return StreamerWindow


def streamer_window(window, table, voltage):
    """(n_e, |E|) after the window's steps, differentiable in the table
    (n_cols, n_rows) and the voltage (torch float64 tensors)."""
    return streamer_window_function().apply(table, voltage, window)


class LeafGeometry:
    """Per cell of the Tree.get_cc_all layout: the leaf-interior mask, the
    cell volume (dr^D; on a cylindrical (r, z) tree the ring 2 pi r dr^2,
    r the cell centre's radius) and the coordinate along the last axis (the
    propagation axis of the drivers' streamers)."""

    def __init__(self, T):
        b = T.boxes()
        n, nc, D = len(b["lvl"]), T.n_cell, T.dim
        m = nc + 2
        self.mask = np.zeros((n, m ** D), bool)
        self.vol = np.zeros((n, m ** D))
        self.z = np.zeros((n, m ** D))
        inner = np.zeros([m] * D, bool)
        inner[(slice(1, -1),) * D] = True
        inner = inner.ravel(order="F")
        # the first and the last axis' index of each cell of a box (x fastest)
        first = np.arange(m ** D) % m
        last = (np.arange(m ** D) // m ** (D - 1)) if D > 1 else np.arange(m)
        cyl = T.info()["coord"] == 1
        # the leaves from the box records (in use, no children): the level
        # lists of the top level are not rebuilt once refinement reaches
        # lvls_max (m_aX_core.f90:497-531), and its new leaves are missing
        # there until a tidy_up
        leaves = np.nonzero((b["in_use"] != 0) & (b["children"][:, 0] == 0) & (b["lvl"] >= 1))[0]
        for r in leaves:
            dr = b["dr"][r]
            self.mask[r] = inner
            if cyl:
                rad = b["r_min"][r, 0] + (first[inner] - 0.5) * dr
                self.vol[r, inner] = 2 * np.pi * rad * dr ** 2
            else:
                self.vol[r, inner] = dr ** D
            self.z[r] = b["r_min"][r, D - 1] + (last - 0.5) * dr


def window_loss(ne, E, geo, domain_len, e_ref, tau=1e-4):
    """The adjoint workload's loss: the electron count (sum of n_e dr^D over
    the leaf cells) plus the head position, the last-axis coordinate of the
    leaf cells weighted by softmax(L |E| / (tau e_ref)): a smooth stand-in
    for the location of the maximum field (src/streamer_cyl.f90:703-723),
    with temperature tau (1e-4 m) on the length scale where the domain
    length L corresponds to the reference field e_ref. Returns (loss,
    count, head) as torch scalars."""
    torch = _torch()
    vol = torch.from_numpy(geo.vol)
    count = (ne * vol).sum()
    m = torch.from_numpy(geo.mask)
    beta = domain_len / (tau * e_ref)
    p = torch.softmax(beta * E[m], dim=0)
    head = (p * torch.from_numpy(geo.z)[m]).sum()
    return count + head, count, head
