"""Differentiation of the streamer's transport substep on a frozen mesh,
as a torch.autograd.Function over the a5g C ABI.

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

No counterpart in the reference, which computes no derivatives.
"""
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
