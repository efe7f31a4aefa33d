"""ctypes binding of liba5g.so (include/a5g.h).

Every call checks the returned status and raises A5GError with the
library's message. Arrays are numpy; box ids and variable indices are
1-based as in Afivo. There is no fallback: if the library is not built the
import of Tree methods fails loudly.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liba5g.so")
_lib = None

ip = C.POINTER(C.c_int)
dp = C.POINTER(C.c_double)


class A5GError(RuntimeError):
    pass


def lib_path():
    return _LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise A5GError("liba5g.so is not built: run python -m a5_streamer_amd.build")
        L = C.CDLL(_LIB)
        L.a5g_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _call(name, *args):
    L = load()
    st = getattr(L, name)(*args)
    if st != 0:
        raise A5GError("%s: %s" % (name, L.a5g_last_error().decode()))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ip)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(dp)


class Tree:
    """One a5g tree: Afivo topology on the host, cell data in HBM."""

    def __init__(self, dim, n_cell, n_cc, n_fc, dr, r_min=None, coarsen_to=-1,
                 lvls_max=30, coord=0, n_boxes=100):
        self.dim = dim
        self.n_cell = n_cell
        self.n_cc = n_cc
        self.dr_base = dr
        self.h = C.c_void_p()
        r_min = np.zeros(dim) if r_min is None else np.asarray(r_min, float)
        rm, prm = _d(r_min)
        _call("a5g_init", C.byref(self.h), C.c_int(dim), C.c_int(n_cell),
              C.c_int(n_cc), C.c_int(n_fc), C.c_double(dr), prm,
              C.c_int(coarsen_to), C.c_int(lvls_max), C.c_int(coord),
              C.c_int(n_boxes))

    def close(self):
        if self.h:
            _call("a5g_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- tree -----------------------------------------------------------
    def set_base(self, ix_list, nb_list):
        ix, pix = _i(np.asarray(ix_list))
        nb, pnb = _i(np.asarray(nb_list))
        _call("a5g_set_base", self.h, C.c_int(ix.shape[0]), pix, pnb)

    def adjust(self, flags_by_id):
        fl, pfl = _i(flags_by_id)
        a, r = C.c_int(), C.c_int()
        _call("a5g_adjust_refinement", self.h, C.c_int(len(fl)), pfl,
              C.byref(a), C.byref(r))
        return a.value, r.value

    def ref_info_lvl(self, lvl):
        a, r = C.c_int(), C.c_int()
        _call("a5g_ref_info_lvl", self.h, C.c_int(lvl), None, None, C.byref(a),
              C.byref(r))
        add = np.zeros(max(a.value, 1), np.int32)
        rm = np.zeros(max(r.value, 1), np.int32)
        _call("a5g_ref_info_lvl", self.h, C.c_int(lvl), _i(add)[1], _i(rm)[1],
              C.byref(a), C.byref(r))
        return add[:a.value], rm[:r.value]

    def tidy_up(self, max_frac, goal_frac, n_clean_min, reorder):
        _call("a5g_tidy_up", self.h, C.c_double(max_frac), C.c_double(goal_frac),
              C.c_int(n_clean_min), C.c_int(int(reorder)))

    def info(self):
        out = np.zeros(8, np.int32)
        _call("a5g_info", self.h, _i(out)[1])
        keys = ("max_id", "max_lvl", "min_lvl", "lvls_max", "n_storage",
                "n_cell", "coord", "ready")
        return dict(zip(keys, (int(x) for x in out)))

    def boxes(self):
        n = self.info()["max_id"]
        D, nch, nnb = self.dim, 2 ** self.dim, 2 * self.dim
        out = dict(lvl=np.zeros(n, np.int32), ix=np.zeros((n, D), np.int32),
                   parent=np.zeros(n, np.int32),
                   children=np.zeros((n, nch), np.int32),
                   neighbors=np.zeros((n, nnb), np.int32),
                   in_use=np.zeros(n, np.int32), n_cell=np.zeros(n, np.int32),
                   dr=np.zeros(n), r_min=np.zeros((n, D)))
        ptrs = [out[k].ctypes.data_as(ip) for k in
                ("lvl", "ix", "parent", "children", "neighbors", "in_use",
                 "n_cell")]
        ptrs += [out[k].ctypes.data_as(dp) for k in ("dr", "r_min")]
        _call("a5g_get_boxes", self.h, C.c_int(n), *ptrs)
        return out

    def level(self, lvl):
        cnt = np.zeros(3, np.int32)
        _call("a5g_level_counts", self.h, C.c_int(lvl), _i(cnt)[1])
        ids = np.zeros(max(cnt[0], 1), np.int32)
        lv = np.zeros(max(cnt[1], 1), np.int32)
        pa = np.zeros(max(cnt[2], 1), np.int32)
        if cnt.any():
            _call("a5g_level_lists", self.h, C.c_int(lvl), _i(ids)[1],
                  _i(lv)[1], _i(pa)[1])
        return ids[:cnt[0]], lv[:cnt[1]], pa[:cnt[2]]

    # --- data -----------------------------------------------------------
    def box_len(self, nc):
        return (nc + 2) ** self.dim

    def set_cc(self, id_, iv, data):
        d, pd = _d(data)
        _call("a5g_set_cc", self.h, C.c_int(id_), C.c_int(iv), pd)

    def get_cc(self, id_, iv, nc=None):
        nc = self.n_cell if nc is None else nc
        out = np.zeros(self.box_len(nc))
        _call("a5g_get_cc", self.h, C.c_int(id_), C.c_int(iv), _d(out)[1])
        return out

    def set_cc_all(self, iv, data):
        d, pd = _d(data)
        _call("a5g_set_cc_all", self.h, C.c_int(iv), pd)

    def get_cc_all(self, iv):
        n = self.info()["max_id"]
        out = np.zeros((n, self.box_len(self.n_cell)))
        _call("a5g_get_cc_all", self.h, C.c_int(iv), _d(out)[1])
        return out

    def set_fc(self, id_, dim, ifc, data):
        d, pd = _d(data)
        _call("a5g_set_fc", self.h, C.c_int(id_), C.c_int(dim), C.c_int(ifc), pd)

    def get_fc(self, id_, dim, ifc):
        nc = self.n_cell
        out = np.zeros((nc + 1) * nc ** (self.dim - 1))
        _call("a5g_get_fc", self.h, C.c_int(id_), C.c_int(dim), C.c_int(ifc),
              _d(out)[1])
        return out

    def sync(self):
        _call("a5g_sync", self.h)

    # --- multigrid and grid ops -----------------------------------------
    def mg_setup(self, i_phi, i_tmp, i_rhs, i_eps=-1, n_down=2, n_up=2,
                 n_base=2, bc_kind=None, bc_val=None, auto_ops=False):
        nnb = 2 * self.dim
        k = np.zeros(nnb, np.int32) if bc_kind is None else bc_kind
        v = np.zeros(nnb) if bc_val is None else bc_val
        ka, pk = _i(k)
        va, pv = _d(v)
        _call("a5g_mg_setup", self.h, C.c_int(i_phi), C.c_int(i_tmp),
              C.c_int(i_rhs), C.c_int(i_eps), C.c_int(n_down), C.c_int(n_up),
              C.c_int(n_base), pk, pv, C.c_int(int(auto_ops)))

    def mg_set_bc_value(self, nb, value):
        _call("a5g_mg_set_bc_value", self.h, C.c_int(nb), C.c_double(value))

    def mg_set_helmholtz(self, kappa2):
        _call("a5g_mg_set_helmholtz", self.h, C.c_double(kappa2))

    def fas_fmg(self, set_residual, first_call):
        _call("a5g_fas_fmg", self.h, C.c_int(int(set_residual)),
              C.c_int(int(first_call)))

    def fas_vcycle(self, max_lvl, set_residual):
        _call("a5g_fas_vcycle", self.h, C.c_int(max_lvl), C.c_int(int(set_residual)))

    def fas_vcycle_adjoint(self, max_lvl, adj):
        """Reverse mode of fas_vcycle(max_lvl, False)
        (a5g_fas_vcycle_adjoint): adj and the returned boundary-value
        adjoints as for fas_fmg_adjoint."""
        a, pa = _i(np.asarray(adj))
        vbar = np.zeros(2 * self.dim)
        _call("a5g_fas_vcycle_adjoint", self.h, C.c_int(max_lvl), pa, _d(vbar)[1])
        return vbar

    def fas_fmg_adjoint(self, adj):
        """Reverse mode of fas_fmg(False, False) (a5g_fas_fmg_adjoint):
        adj = (lp, lt, lr, lw) cell variables; returns the adjoints of the
        boundary values V per face."""
        a, pa = _i(np.asarray(adj))
        vbar = np.zeros(2 * self.dim)
        _call("a5g_fas_fmg_adjoint", self.h, pa, _d(vbar)[1])
        return vbar

    SHAPE_FIELDS = ("kernel", "dim", "nc", "size_class", "shape", "boxes", "us", "chosen")

    def set_shape(self, kernel, nc, cls, shape):
        _call("a5g_set_shape", self.h, C.c_int(kernel), C.c_int(nc), C.c_int(cls),
              C.c_int(shape))

    def get_shape(self, kernel, nc, cls):
        s = C.c_int()
        _call("a5g_get_shape", self.h, C.c_int(kernel), C.c_int(nc), C.c_int(cls), C.byref(s))
        return s.value

    def autotune(self, reps=20):
        """Launch-shape autotuner (a5g_autotune): a list of dicts, one per
        (kernel, box size, size class, shape) timed, SHAPE_FIELDS as keys."""
        n = C.c_int()
        rec = np.zeros(8 * 5 * 6 * 3 * 4)
        _call("a5g_autotune", self.h, C.c_int(reps), C.c_int(len(rec) // 8), _d(rec)[1],
              C.byref(n))
        out = []
        for i in range(min(n.value, len(rec) // 8)):
            row = rec[8 * i:8 * i + 8]
            out.append({k: (float(v) if k == "us" else int(v))
                        for k, v in zip(self.SHAPE_FIELDS, row)})
        return out

    def solve_mixed(self, tol=1e-10, max_outer=30, n_inner=1, zero_start=True):
        it, rel = C.c_int(), C.c_double()
        _call("a5g_mg_solve_mixed", self.h, C.c_double(tol), C.c_int(max_outer),
              C.c_int(n_inner), C.c_int(int(zero_start)), C.byref(it), C.byref(rel))
        return it.value, rel.value

    def residual_norms(self):
        r, b = C.c_double(), C.c_double()
        _call("a5g_mg_residual_norms", self.h, C.byref(r), C.byref(b))
        return r.value, b.value

    def gc_tree(self, iv, rb, bc):
        _call("a5g_gc_tree", self.h, C.c_int(iv), C.c_int(rb), C.c_int(bc))

    def restrict_tree(self, iv):
        _call("a5g_restrict_tree", self.h, C.c_int(iv))

    def tree_copy_cc(self, a, b):
        _call("a5g_tree_copy_cc", self.h, C.c_int(a), C.c_int(b))

    # --- transport table and streamer physics -----------------------------
    def table_create(self, x_min, x_max, n_rows, n_cols):
        self._tbl = (n_rows, n_cols)
        _call("a5g_table_create", self.h, C.c_double(x_min), C.c_double(x_max),
              C.c_int(n_rows), C.c_int(n_cols))

    def table_set_col(self, col, xx, yy):
        xa, px = _d(xx)
        ya, py = _d(yy)
        _call("a5g_table_set_col", self.h, C.c_int(col), C.c_int(len(xa)), px, py)

    def table_get(self):
        n_rows, n_cols = self._tbl
        rc = np.zeros(n_rows * n_cols)
        xm, idx = C.c_double(), C.c_double()
        _call("a5g_table_get", self.h, _d(rc)[1], C.byref(xm), C.byref(idx))
        return rc.reshape(n_cols, n_rows), xm.value, idx.value

    def set_table(self, rows_cols, x_min, inv_dx):
        rc = np.asarray(rows_cols, float)
        n_cols, n_rows = rc.shape
        self._tbl = (n_rows, n_cols)
        _call("a5g_set_table", self.h, C.c_int(n_rows), C.c_int(n_cols),
              C.c_double(x_min), C.c_double(inv_dx), _d(rc)[1])

    def streamer_setup(self, elec=1, pion=2, elec_old=3, pion_old=4, phi=5, fld=6,
                       rhs=7, pho=-1, eps=9, f_elec=1, f_fld=2, t_mob=1, t_dif=2,
                       t_alpha=3, t_eta=4, dielectric_rule=True):
        v = np.array([elec, pion, elec_old, pion_old, phi, fld, rhs, pho, eps,
                      f_elec, f_fld, t_mob, t_dif, t_alpha, t_eta,
                      int(dielectric_rule)], np.int32)
        _call("a5g_streamer_setup", self.h, _i(v)[1])

    def space_charge(self, fac):
        _call("a5g_space_charge", self.h, C.c_double(fac))

    def fld_from_pot(self):
        _call("a5g_fld_from_pot", self.h)

    def fluxes_koren(self, dt):
        _call("a5g_fluxes_koren", self.h, C.c_double(dt))

    def consistent_fluxes(self, f_ixs):
        a, p = _i(f_ixs)
        _call("a5g_consistent_fluxes", self.h, C.c_int(len(a)), p)

    def update_solution(self, dt):
        _call("a5g_update_solution", self.h, C.c_double(dt))

    def average_dens(self):
        _call("a5g_average_dens", self.h)

    def photoi_helmholtz(self, coeff, amp, kappa, i_phi, i_rhs, i_tmp, tol=1e-8,
                         max_cycles=30):
        a, pa = _d(amp)
        k, pk = _d(kappa)
        if len(a) != len(k):
            raise A5GError("photoi_helmholtz: amp and kappa differ in length")
        cyc = C.c_int()
        _call("a5g_photoi_helmholtz", self.h, C.c_double(coeff), C.c_int(len(a)), pa, pk,
              C.c_int(i_phi), C.c_int(i_rhs), C.c_int(i_tmp), C.c_double(tol),
              C.c_int(max_cycles), C.byref(cyc))
        return cyc.value

    def photoi_helmholtz_terms(self, coeff, amp, kappa, i_terms, i_rhs, i_tmp, warm,
                               tol=1e-8, max_cycles=30):
        """The Helmholtz photoionization with term j in variables of its own
        (solution i_terms[j], rhs i_rhs[j], scratch i_tmp[j]), the terms'
        solves side by side; warm starts each term's V-cycles from what an
        earlier call left in i_terms[j] instead of an FMG from zero."""
        a, pa = _d(amp)
        k, pk = _d(kappa)
        iv = [np.ascontiguousarray(x, dtype=np.int32) for x in (i_terms, i_rhs, i_tmp)]
        if not all(len(x) == len(a) for x in [k] + iv):
            raise A5GError("photoi_helmholtz_terms: amp, kappa and the variables differ in "
                           "length")
        cyc = C.c_int()
        pi = [x.ctypes.data_as(C.POINTER(C.c_int)) for x in iv]
        _call("a5g_photoi_helmholtz_terms", self.h, C.c_double(coeff), C.c_int(len(a)), pa, pk,
              pi[0], pi[1], pi[2], C.c_int(int(bool(warm))), C.c_double(tol),
              C.c_int(max_cycles), C.byref(cyc))
        return cyc.value

    def box_summaries(self):
        n = self.info()["max_id"]
        out = np.zeros((n, 3))
        _call("a5g_box_summaries", self.h, _d(out)[1])
        return out

    def set_ref_rules(self, seeds, regions, init_window, crv_refine, n_base_last=1):
        """seeds (runs, n_seed, 2 dim + 1), regions (n_reg, 1 + 2 dim)."""
        sd = np.asarray(seeds, float)
        rg = np.asarray(regions, float).reshape(-1, 1 + 2 * self.dim)
        a, pa = _d(sd)
        b, pb = _d(rg if len(rg) else np.zeros(1))
        _call("a5g_set_ref_rules", self.h, C.c_int(sd.shape[1] if sd.ndim == 3 else 0), pa,
              C.c_int(len(rg)), pb, C.c_double(init_window), C.c_double(crv_refine),
              C.c_int(n_base_last))

    def ref_flags(self, times):
        n = self.info()["max_id"]
        out = np.zeros(n, np.int32)
        t, pt = _d(np.atleast_1d(np.asarray(times, float)))
        _call("a5g_ref_flags", self.h, pt, _i(out)[1])
        return out

    def instance_maxima(self, iv, runs=1):
        v = np.zeros(runs)
        lv = np.zeros(runs, np.int32)
        _call("a5g_instance_maxima", self.h, C.c_int(iv), _d(v)[1], _i(lv)[1])
        return v, lv

    # --- reverse mode of the transport substep (a5g_*_adjoint) ------------
    def update_solution_adjoint(self, dt, adj):
        a, pa = _i(np.asarray(adj))
        _call("a5g_update_solution_adjoint", self.h, C.c_double(dt), pa)

    def consistent_fluxes_adjoint(self, lF):
        _call("a5g_consistent_fluxes_adjoint", self.h, C.c_int(lF))

    def fluxes_koren_adjoint(self, dt, adj):
        a, pa = _i(np.asarray(adj))
        _call("a5g_fluxes_koren_adjoint", self.h, C.c_double(dt), pa)

    def table_adjoint(self, clear=False):
        """The accumulated table-row adjoints as (n_cols, n_rows)."""
        n_rows, n_cols = self._tbl
        out = np.zeros(n_rows * n_cols)
        _call("a5g_table_adjoint", self.h, C.c_int(int(bool(clear))), _d(out)[1])
        return out.reshape(n_cols, n_rows)

    # --- reverse mode of the rest of the step (a5g_*_adjoint) -------------
    def tree_copy_cc_adjoint(self, lfrom, lto):
        _call("a5g_tree_copy_cc_adjoint", self.h, C.c_int(lfrom), C.c_int(lto))

    def average_dens_adjoint(self, le, le_old, li, li_old):
        a, pa = _i(np.array([le, le_old, li, li_old]))
        _call("a5g_average_dens_adjoint", self.h, pa)

    def space_charge_adjoint(self, fac, lrhs, le, li):
        a, pa = _i(np.array([lrhs, le, li]))
        _call("a5g_space_charge_adjoint", self.h, C.c_double(fac), pa)

    def restrict_tree_adjoint(self, il):
        _call("a5g_restrict_tree_adjoint", self.h, C.c_int(il))

    def gc_tree_adjoint(self, iv, il, rb, bc):
        _call("a5g_gc_tree_adjoint", self.h, C.c_int(iv), C.c_int(il), C.c_int(rb), C.c_int(bc))

    def fld_from_pot_adjoint(self, lE, lfld, lphi):
        a, pa = _i(np.array([lE, lfld, lphi]))
        _call("a5g_fld_from_pot_adjoint", self.h, pa)

    def state_store(self, slot, ivs, fvs=(), load=False, clear=False):
        """Saves (default) or restores (load) whole cell variables ivs and
        face variables fvs in device store slot `slot`; clear sets them to
        zero instead (slot unused)."""
        a, pa = _i(np.asarray(list(ivs) or [0]))
        b, pb = _i(np.asarray(list(fvs) or [0]))
        op = 2 if clear else (1 if load else 0)
        _call("a5g_state_store", self.h, C.c_int(slot), C.c_int(len(ivs)), pa, C.c_int(len(fvs)),
              pb, C.c_int(op))

    def box_summaries_noalloc(self):
        """box_summaries through the library's fixed scratch: no device
        allocation (at most 131072 boxes; the tree must be uploaded)."""
        n = self.info()["max_id"]
        out = np.zeros((n, 3))
        _call("a5g_box_summaries_noalloc", self.h, _d(out)[1])
        return out

    def prolong1_to(self, ids, iv, add=False):
        if add:
            raise A5GError("prolong1_to: only add=False (prolong_to_new_boxes) is provided")
        a, p = _i(np.atleast_1d(ids))
        _call("a5g_prolong1_to", self.h, C.c_int(len(a)), p, C.c_int(iv))

    def fill_cc(self, ids, iv, value):
        a, p = _i(np.atleast_1d(ids))
        _call("a5g_fill_cc", self.h, C.c_int(len(a)), p, C.c_int(iv), C.c_double(value))

    def gc_boxes(self, ids, iv, rb, bc):
        a, p = _i(np.atleast_1d(ids))
        _call("a5g_gc_boxes", self.h, C.c_int(len(a)), p, C.c_int(iv), C.c_int(rb),
              C.c_int(bc))

    def set_gaussians(self, iv, blobs, base=0.0):
        b, pb = _d(np.asarray(blobs, float).reshape(-1, self.dim + 2))
        _call("a5g_set_gaussians", self.h, C.c_int(iv), C.c_int(b.shape[0]), pb,
              C.c_double(base))

    def set_gaussians_instances(self, iv, blobs, base):
        """blobs (n_inst, n_blob, dim + 2) in tree coordinates, base (n_inst,)."""
        base = np.asarray(base, float)
        b, pb = _d(np.asarray(blobs, float).reshape(len(base), -1, self.dim + 2))
        bb, pbb = _d(base)
        _call("a5g_set_gaussians_instances", self.h, C.c_int(iv), C.c_int(b.shape[1]), pb, pbb)

    def prof_enable(self, on):
        _call("a5g_prof_enable", self.h, C.c_int(int(on)))

    def prof_read(self, kind, lvl=-999):
        n, ms, by = C.c_int(), C.c_double(), C.c_double()
        _call("a5g_prof_read", self.h, C.c_int(kind), C.c_int(lvl), C.byref(n),
              C.byref(ms), C.byref(by))
        return n.value, ms.value, by.value

    def tree_maxdiff(self, iv_a, iv_b):
        d, b = C.c_double(), C.c_double()
        _call("a5g_tree_maxdiff_cc", self.h, C.c_int(iv_a), C.c_int(iv_b), C.byref(d),
              C.byref(b))
        return d.value, b.value

    def min_dr(self):
        """a$D_min_dr (afivo/src/m_aX_utils.f90:196-200): dr of the finest level."""
        return self.dr_base * 0.5 ** (self.info()["max_lvl"] - 1)

    def tree_minmax(self, iv):
        lo, hi = C.c_double(), C.c_double()
        _call("a5g_tree_minmax_cc", self.h, C.c_int(iv), C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    # --- independent runs in one tree ----------------------------------------
    def set_instances(self, n_inst, stride):
        """n_inst independent runs stacked along the last dimension, one every
        stride (a5g_set_instances)."""
        _call("a5g_set_instances", self.h, C.c_int(n_inst), C.c_double(stride))

    def set_instance_values(self, what, values):
        """Per-run time steps (what = 0) or values V of the rule 2V - c on
        face what (1..2 dim)."""
        v, p = _d(np.atleast_1d(values))
        _call("a5g_set_instance_values", self.h, C.c_int(what), C.c_int(len(v)), p)

    def set_active_instances(self, active=None):
        """Restrict the tree-wide calls to the runs with active[g] (None: all)."""
        if active is None:
            _call("a5g_set_active_instances", self.h, C.c_int(0), None)
            return
        a, p = _i(np.asarray(active, np.int32))
        _call("a5g_set_active_instances", self.h, C.c_int(len(a)), p)

    def set_instance_groups(self, group=None):
        """Groups of runs for fas_fmg_groups (group[g] >= 0, or -1 to leave
        run g out; None clears them)."""
        if group is None:
            _call("a5g_set_instance_groups", self.h, C.c_int(0), None)
            return
        a, p = _i(np.asarray(group, np.int32))
        _call("a5g_set_instance_groups", self.h, C.c_int(len(a)), p)

    def fas_fmg_groups(self, set_residual, first_call):
        """mg$D_fas_fmg of every group of runs up to its own finest level,
        the groups side by side."""
        _call("a5g_fas_fmg_groups", self.h, C.c_int(int(set_residual)),
              C.c_int(int(first_call)))

    # --- output (afivo/src/m_aX_io.f90:20-209, m_vtk.f90) -------------------
    def leaf_cells(self, ixs_cc, ixs_fc=()):
        """a$D_write_vtk's cell data: array (n_cells, columns), leaves level
        by level in list order, cells x fastest; the face columns (one per
        dimension) are 0.5 * the sum over the cell's two faces and over all
        face variables ixs_fc."""
        cc, pcc = _i(np.atleast_1d(np.asarray(ixs_cc, np.int32)))
        fc, pfc = _i(np.atleast_1d(np.asarray(ixs_fc, np.int32)))
        n = C.c_longlong()
        _call("a5g_leaf_cells", self.h, C.c_int(len(cc)), pcc, C.c_int(len(fc)), pfc,
              C.c_longlong(0), None, C.byref(n))
        cols = len(cc) + (self.dim if len(fc) else 0)
        out = np.zeros((cols, n.value))
        if out.size:
            _call("a5g_leaf_cells", self.h, C.c_int(len(cc)), pcc, C.c_int(len(fc)), pfc,
                  C.c_longlong(out.size), _d(out)[1], C.byref(n))
        return out.T

    def write_vtk(self, filename, cc_names, n_cycle, time, ixs_cc=None, fc_names=(),
                  ixs_fc=()):
        """a$D_write_vtk: <filename>.vtu with the leaves as pixels/voxels."""
        cn = (C.c_char_p * max(len(cc_names), 1))(*[n.encode() for n in cc_names])
        fn = (C.c_char_p * max(len(fc_names), 1))(*[n.encode() for n in fc_names])
        if len(fc_names) != len(ixs_fc) * self.dim:
            raise A5GError("write_vtk: size(fc_names) /= size(ixs_fc) * dim")
        icc = None if ixs_cc is None else _i(np.asarray(ixs_cc))
        ifc = _i(np.atleast_1d(np.asarray(ixs_fc, np.int32)))
        _call("a5g_write_vtk", self.h, str(filename).encode(), C.c_int(len(cc_names)), cn,
              None if icc is None else icc[1], C.c_int(len(ixs_fc)), fn, ifc[1],
              C.c_int(n_cycle), C.c_double(time))

    def dump_leaves(self, filename, ivs, cycle=0, time=0.0):
        """The raw leaf-box dump of cell variables ivs (read_leaf_dump)."""
        a, p = _i(np.atleast_1d(ivs))
        _call("a5g_dump_leaves", self.h, str(filename).encode(), C.c_int(len(a)), p,
              C.c_int(cycle), C.c_double(time))

    # --- streamer diagnostics (src/streamer_cyl.f90:682-956) ---------------
    def get_loc(self, r, max_lvl=0):
        """(id, ix) of the leaf cell holding point r (a$D_get_loc); (-1,
        [-1, ...]) outside the base boxes. max_lvl < 1: no level limit."""
        ix = np.zeros(self.dim, np.int32)
        idv = C.c_int()
        _call("a5g_get_loc", self.h, _d(np.asarray(r, float))[1], C.c_int(max_lvl),
              C.byref(idv), _i(ix)[1])
        return idv.value, ix

    def reduction_loc(self, kind, var, d, thr=0.0, init_val=-1e99):
        """(value, id, ix) of the located maximum over the leaves
        (a$D_reduction_loc with the box procedures of get_streamer_properties):
        kind 1 = 0.5 |f(ix) + f(ix + e_d)| of face variable var, kind 2 = the
        cell-centre coordinate d where cc(var) > thr; d is 1-based."""
        ix = np.zeros(self.dim, np.int32)
        val, idv = C.c_double(), C.c_int()
        _call("a5g_reduction_loc", self.h, C.c_int(kind), C.c_int(var), C.c_int(d),
              C.c_double(thr), C.c_double(init_val), C.byref(val), C.byref(idv), _i(ix)[1])
        return val.value, idv.value, ix

    def cc_axis(self, ixs_cc, ixs_fc):
        """get_cc_axis (2D): array (1 + len(ixs_cc) + len(ixs_fc), n) of the
        cell-centre y, the cell values and the face averages along x = 0."""
        cc, pcc = _i(np.atleast_1d(ixs_cc))
        fc, pfc = _i(np.atleast_1d(ixs_fc))
        n = C.c_int()
        _call("a5g_cc_axis", self.h, C.c_int(len(cc)), pcc, C.c_int(len(fc)), pfc,
              C.c_int(0), None, C.byref(n))
        out = np.zeros((n.value, len(cc) + len(fc) + 1))
        if n.value:
            _call("a5g_cc_axis", self.h, C.c_int(len(cc)), pcc, C.c_int(len(fc)), pfc,
                  C.c_int(n.value), _d(out)[1], C.byref(n))
        return out.T


def read_leaf_dump(path):
    """Read a raw leaf-box dump (a5g_dump_leaves): a dict with the header
    fields, per-leaf id, lvl, ix (n, 3), dr, r_min (n, 3), the variable
    indices and data (n_var, n_leaves) + (n_cell,) * dim in (x, y[, z])
    index order."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != b"A5GLEAF1":
        raise A5GError("%s is not an a5g leaf dump" % path)
    dim, nc, n, nv, coord, cycle = np.frombuffer(raw, "<i4", 6, 8)
    time = float(np.frombuffer(raw, "<f8", 1, 32)[0])
    off = 40
    ivs = np.frombuffer(raw, "<i4", nv, off)
    off += 4 * nv
    rec = np.dtype([("id", "<i4"), ("lvl", "<i4"), ("ix", "<i4", 3), ("dr", "<f8"),
                    ("r_min", "<f8", 3)])
    leaves = np.frombuffer(raw, rec, n, off)
    off += rec.itemsize * n
    cells = nc ** dim
    data = np.frombuffer(raw, "<f8", nv * n * cells, off)
    if off + data.nbytes != len(raw):
        raise A5GError("%s: size does not match its header" % path)
    # x fastest in the file: reverse the axes so that data[v, leaf][i, j(, k)]
    data = data.reshape((nv, n) + (nc,) * dim).transpose((0, 1) + tuple(range(dim + 1, 1, -1)))
    return dict(dim=int(dim), n_cell=int(nc), coord=int(coord), cycle=int(cycle), time=time,
                ivs=ivs.copy(), id=leaves["id"].copy(), lvl=leaves["lvl"].copy(),
                ix=leaves["ix"][:, :dim].copy(), dr=leaves["dr"].copy(),
                r_min=leaves["r_min"][:, :dim].copy(), data=np.ascontiguousarray(data))
