"""ctypes view of oracle/liba5oracle.so, the C restatement of the a5_streamer
hot path (tree core, ghost cells, multigrid, streamer physics).

Test infrastructure only: tests, the smoke check and the CPU-baseline leg of
bench.py use it as the checker. The API mirrors oracle/refharness.py so the
same scenario can be run on the reference, the restatement and the HIP path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liba5oracle.so")

_lib = None

ip = C.POINTER(C.c_int)
dp = C.POINTER(C.c_double)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liba5oracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liba5oracle.so is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.o_tree_new.restype = C.c_void_p
        L.o_tree_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                 dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.o_tree_free.argtypes = [C.c_void_p]
        L.o_min_dr.restype = C.c_double
        L.o_min_dr.argtypes = [C.c_void_p]
        L.o_max_residual.restype = C.c_double
        L.o_max_residual.argtypes = [C.c_void_p]
        L.o_koren_mlim.restype = C.c_double
        L.o_koren_mlim.argtypes = [C.c_double, C.c_double]
        L.o_lt_create.restype = C.c_void_p
        L.o_lt_create.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int]
        L.o_lt_free.argtypes = [C.c_void_p]
        L.o_lt_set_col.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp]
        L.o_lt_get_col.restype = C.c_double
        L.o_lt_get_col.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.o_lt_data.argtypes = [C.c_void_p, dp, dp, dp]
        L.o_norm2.restype = C.c_double
        L.o_norm2.argtypes = [C.c_int, dp]
        L.o_reduction_loc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, dp, ip, ip]
        L.o_cc_axis.restype = C.c_int
        L.o_leaf_cells.restype = C.c_longlong
        L.o_space_charge_adj.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]
        L.o_store_size.restype = C.c_size_t
        L.o_store_size.argtypes = [C.c_void_p, C.c_int, C.c_int]
        for name in ("o_space_charge", "o_fld_from_pot", "o_average_dens"):
            getattr(L, name).argtypes = None
        _lib = L
    return _lib


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ip)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(dp)


class OTree:
    """A tree of the C restatement; same methods as refharness.RefTree."""

    def __init__(self, dim, n_cell, n_cc, n_fc, dr, r_min=None, coarsen_to=-1,
                 lvls_max=30, coord=0, n_boxes=100):
        self.dim = dim
        self.n_cell = n_cell
        self.n_cc = n_cc
        self.L = lib()
        r_min = np.zeros(dim) if r_min is None else np.asarray(r_min, float)
        self._rm = _d(r_min)[0]
        self.h = C.c_void_p(self.L.o_tree_new(dim, n_cell, n_cc, n_fc, dr,
                                              self._rm.ctypes.data_as(dp),
                                              coarsen_to, lvls_max, coord,
                                              n_boxes))

    def __del__(self):
        try:
            self.L.o_tree_free(self.h)
        except Exception:
            pass

    def set_base(self, ix_list, nb_list):
        ix, pix = _i(np.asarray(ix_list))
        nb, pnb = _i(np.asarray(nb_list))
        self.L.o_set_base(self.h, C.c_int(ix.shape[0]), pix, pnb)

    def adjust(self, flags_by_id):
        fl, pfl = _i(flags_by_id)
        a, r = C.c_int(), C.c_int()
        self.L.o_adjust(self.h, C.c_int(len(fl)), pfl, C.byref(a), C.byref(r))
        return a.value, r.value

    def ref_info_lvl(self, lvl):
        a, r = C.c_int(), C.c_int()
        self.L.o_ref_info_lvl(self.h, C.c_int(lvl), None, None, C.byref(a),
                              C.byref(r))
        add = np.zeros(max(a.value, 1), np.int32)
        rm = np.zeros(max(r.value, 1), np.int32)
        self.L.o_ref_info_lvl(self.h, C.c_int(lvl), _i(add)[1], _i(rm)[1],
                              C.byref(a), C.byref(r))
        return add[:a.value], rm[:r.value]

    def tidy_up(self, max_frac, goal_frac, n_clean_min, reorder):
        self.L.o_tidy_up(self.h, C.c_double(max_frac), C.c_double(goal_frac),
                         C.c_int(n_clean_min), C.c_int(int(reorder)))

    def info(self):
        out = np.zeros(8, np.int32)
        self.L.o_info(self.h, _i(out)[1])
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
        self.L.o_get_boxes(self.h, C.c_int(n), *ptrs)
        return out

    def level(self, lvl):
        cnt = np.zeros(3, np.int32)
        self.L.o_level_counts(self.h, C.c_int(lvl), _i(cnt)[1])
        ids = np.zeros(max(cnt[0], 1), np.int32)
        lv = np.zeros(max(cnt[1], 1), np.int32)
        pa = np.zeros(max(cnt[2], 1), np.int32)
        if cnt[0] or cnt[1] or cnt[2]:
            self.L.o_level_lists(self.h, C.c_int(lvl), _i(ids)[1], _i(lv)[1],
                                 _i(pa)[1])
        return ids[:cnt[0]], lv[:cnt[1]], pa[:cnt[2]]

    def box_len(self, nc):
        return (nc + 2) ** self.dim

    def set_cc(self, id_, iv, data):
        d, pd = _d(data)
        self.L.o_set_cc(self.h, C.c_int(id_), C.c_int(iv), pd)

    def get_cc(self, id_, iv, nc=None):
        nc = self.n_cell if nc is None else nc
        out = np.zeros(self.box_len(nc))
        self.L.o_get_cc(self.h, C.c_int(id_), C.c_int(iv), _d(out)[1])
        return out

    def set_fc(self, id_, dim, ifc, data):
        d, pd = _d(data)
        self.L.o_set_fc(self.h, C.c_int(id_), C.c_int(dim), C.c_int(ifc), pd)

    def get_fc(self, id_, dim, ifc):
        nc = self.n_cell
        out = np.zeros((nc + 1) * nc ** (self.dim - 1))
        self.L.o_get_fc(self.h, C.c_int(id_), C.c_int(dim), C.c_int(ifc),
                        _d(out)[1])
        return out

    def mg_setup(self, i_phi, i_tmp, i_rhs, i_eps=-1, n_down=2, n_up=2,
                 n_base=2, bc_kind=None, bc_val=None, auto_ops=False):
        nnb = 2 * self.dim
        k = np.zeros(nnb, np.int32) if bc_kind is None else bc_kind
        v = np.zeros(nnb) if bc_val is None else bc_val
        ka, pk = _i(k)
        va, pv = _d(v)
        self.L.o_mg_setup(self.h, C.c_int(i_phi), C.c_int(i_tmp),
                          C.c_int(i_rhs), C.c_int(i_eps), C.c_int(n_down),
                          C.c_int(n_up), C.c_int(n_base), pk, pv,
                          C.c_int(int(auto_ops)))

    def mg_set_bc_value(self, nb, value):
        self.L.o_mg_set_bc_value(self.h, C.c_int(nb), C.c_double(value))

    def mg_set_helmholtz(self, kappa2):
        self.L.o_mg_set_helmholtz(self.h, C.c_double(kappa2))

    def fas_fmg(self, set_residual, first_call):
        self.L.o_fas_fmg(self.h, C.c_int(int(set_residual)),
                         C.c_int(int(first_call)))

    def fas_vcycle(self, max_lvl, set_residual):
        self.L.o_fas_vcycle(self.h, C.c_int(max_lvl), C.c_int(int(set_residual)))

    def box_gsrb(self, id_, cntr):
        self.L.o_box_gsrb(self.h, C.c_int(id_), C.c_int(cntr))

    def box_op(self, id_, i_out):
        self.L.o_box_op(self.h, C.c_int(id_), C.c_int(i_out))

    def box_rstr(self, id_, iv):
        self.L.o_box_rstr(self.h, C.c_int(id_), C.c_int(iv))

    def box_corr(self, id_):
        self.L.o_box_corr(self.h, C.c_int(id_))

    def gc_box(self, id_, iv, rb, bc):
        self.L.o_gc_box(self.h, C.c_int(id_), C.c_int(iv), C.c_int(rb),
                        C.c_int(bc))

    def gc_tree(self, iv, rb, bc):
        self.L.o_gc_tree(self.h, C.c_int(iv), C.c_int(rb), C.c_int(bc))

    def gc2_box(self, id_, iv):
        nc = self.n_cell
        out = np.zeros(nc ** (self.dim - 1) * 2 * self.dim)
        self.L.o_gc2_box(self.h, C.c_int(id_), C.c_int(iv), _d(out)[1])
        return out

    def restrict_tree(self, iv):
        self.L.o_restrict_tree(self.h, C.c_int(iv))

    def consistent_fluxes(self, f_ixs):
        a, p = _i(f_ixs)
        self.L.o_consistent_fluxes(self.h, C.c_int(len(a)), p)

    def prolong1_to(self, id_, iv, add=False):
        self.L.o_prolong1_to(self.h, C.c_int(id_), C.c_int(iv), C.c_int(int(add)))

    def prolong0_to(self, id_, iv):
        self.L.o_prolong0_to(self.h, C.c_int(id_), C.c_int(iv))

    def tree_copy_cc(self, a, b):
        self.L.o_tree_copy_cc(self.h, C.c_int(a), C.c_int(b))

    def tree_minmax(self, iv):
        lo, hi = C.c_double(), C.c_double()
        self.L.o_tree_minmax_cc(self.h, C.c_int(iv), C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    def min_dr(self):
        return self.L.o_min_dr(self.h)

    def max_residual(self):
        return self.L.o_max_residual(self.h)

    def get_loc(self, r, max_lvl=0):
        """(id, ix) of the leaf cell holding point r (a$D_get_loc);
        (-1, [-1, ...]) outside the base boxes."""
        ix, pix = _i(np.zeros(self.dim))
        idv = C.c_int()
        self.L.o_get_loc(self.h, _d(np.asarray(r, float))[1], C.c_int(max_lvl),
                         C.byref(idv), pix)
        return idv.value, ix

    def reduction_loc(self, kind, var, d, thr=0.0, init_val=-1e99):
        """(value, id, ix) of the located maximum (a$D_reduction_loc)."""
        ix, pix = _i(np.zeros(self.dim))
        val, idv = C.c_double(), C.c_int()
        self.L.o_reduction_loc(self.h, C.c_int(kind), C.c_int(var), C.c_int(d - 1),
                               C.c_double(thr), C.c_double(init_val), C.byref(val),
                               C.byref(idv), pix)
        return val.value, idv.value, ix

    def leaf_cells(self, ixs_cc, ixs_fc):
        """a$D_write_vtk's cell data: array (n_cells, columns)."""
        cc, pcc = _i(np.atleast_1d(ixs_cc))
        fc, pfc = _i(np.atleast_1d(ixs_fc))
        n = self.L.o_leaf_cells(self.h, len(cc), pcc, len(fc), pfc, None)
        cols = len(cc) + (self.dim if len(fc) else 0)
        out = np.zeros((cols, n))
        if n and cols:
            self.L.o_leaf_cells(self.h, len(cc), pcc, len(fc), pfc, _d(out)[1])
        return out.T

    def vtk_geometry(self):
        """(coords (n_nodes, dim), connects, offsets, types) of a$D_write_vtk."""
        D, nc = self.dim, self.n_cell
        n_grids = sum(len(self.level(l)[1]) for l in range(1, self.info()["max_lvl"] + 1))
        n_nodes, n_cells = n_grids * (nc + 1) ** D, n_grids * nc ** D
        coords = np.zeros((n_nodes, D))
        conn = np.zeros(n_cells * 2 ** D, np.int32)
        offs = np.zeros(n_cells, np.int32)
        types = np.zeros(n_cells, np.int32)
        if n_grids:
            self.L.o_vtk_geometry(self.h, coords.ctypes.data_as(dp),
                                  conn.ctypes.data_as(ip), offs.ctypes.data_as(ip),
                                  types.ctypes.data_as(ip))
        return coords, conn, offs, types

    def cc_axis(self, ixs_cc, ixs_fc):
        """get_cc_axis: array (len(ixs_cc) + len(ixs_fc) + 1, n)."""
        cc, pcc = _i(ixs_cc)
        fc, pfc = _i(ixs_fc)
        n = self.L.o_cc_axis(self.h, len(cc), pcc, len(fc), pfc, None, 0)
        out = np.zeros((n, len(cc) + len(fc) + 1))
        self.L.o_cc_axis(self.h, len(cc), pcc, len(fc), pfc, _d(out)[1] if n else None, n)
        return out.T


def koren_mlim(a, b):
    return lib().o_koren_mlim(a, b)


class OVars(C.Structure):
    """Variable indices of the streamer models (src/m_streamer.f90:16-43)."""
    _fields_ = [(n, C.c_int) for n in (
        "elec", "pion", "elec_old", "pion_old", "phi", "fld", "rhs", "pho", "eps",
        "f_elec", "f_fld", "t_mob", "t_dif", "t_alpha", "t_eta",
        "dielectric_rule", "photoi")]


def streamer_vars(dielectric_rule=True, photoi=False):
    return OVars(1, 2, 3, 4, 5, 6, 7, 8, 9, 1, 2, 1, 2, 3, 4,
                 int(dielectric_rule), int(photoi))


class OTable:
    """fosito's uniform lookup table (m_lookup_table.f90)."""

    def __init__(self, x_min, x_max, n_rows, n_cols):
        self.L = lib()
        self.n_rows, self.n_cols = n_rows, n_cols
        self.h = C.c_void_p(self.L.o_lt_create(x_min, x_max, n_rows, n_cols))

    def __del__(self):
        try:
            self.L.o_lt_free(self.h)
        except Exception:
            pass

    def set_col(self, col, xx, yy):
        xa, px = _d(xx)
        ya, py = _d(yy)
        self.L.o_lt_set_col(self.h, col, len(xa), px, py)

    def get_col(self, col, x):
        return self.L.o_lt_get_col(self.h, col, x)

    def set_rows(self, rows_cols):
        """Overwrite the table with rows_cols (n_cols, n_rows), as data() gives it."""
        a, p = _d(np.asarray(rows_cols, float).reshape(-1))
        self.L.o_lt_set_rows(self.h, p)

    def data(self):
        rc = np.zeros(self.n_rows * self.n_cols)
        xm, idx = C.c_double(), C.c_double()
        self.L.o_lt_data(self.h, _d(rc)[1], C.byref(xm), C.byref(idx))
        return rc.reshape(self.n_cols, self.n_rows), xm.value, idx.value


def norm2(x):
    a, p = _d(x)
    return lib().o_norm2(len(a), p)


def space_charge(T, v, fac):
    T.L.o_space_charge(T.h, C.byref(v), C.c_double(fac))


def fld_from_pot(T, v):
    T.L.o_fld_from_pot(T.h, C.byref(v))


def fluxes_koren(T, v, tbl, dt):
    T.L.o_fluxes_koren(T.h, C.byref(v), tbl.h, C.c_double(dt))


def update_solution(T, v, tbl, dt):
    T.L.o_update_solution(T.h, C.byref(v), tbl.h, C.c_double(dt))


def average_dens(T, v):
    T.L.o_average_dens(T.h, C.byref(v))


def photoi_rate(T, v, tbl, coeff):
    T.L.o_photoi_rate(T.h, C.byref(v), tbl.h, C.c_double(coeff))


class OAdj(C.Structure):
    """Adjoint variables of the reverse-mode transport substep: cell
    variables le, li, lE, lpho and face variables lF, lfld."""
    _fields_ = [(n, C.c_int) for n in ("le", "li", "lE", "lpho", "lF", "lfld")]


def _tg(tgrad):
    return None if tgrad is None else tgrad.ctypes.data_as(C.POINTER(C.c_double))


def update_solution_adj(T, v, tbl, dt, a, tgrad=None):
    T.L.o_update_solution_adj(T.h, C.byref(v), tbl.h, C.c_double(dt), C.byref(a), _tg(tgrad))


def consistent_fluxes_adj(T, lF):
    T.L.o_consistent_fluxes_adj(T.h, C.c_int(lF))


def fluxes_koren_adj(T, v, tbl, dt, a, tgrad=None):
    T.L.o_fluxes_koren_adj(T.h, C.byref(v), tbl.h, C.c_double(dt), C.byref(a), _tg(tgrad))


def box_summaries(T, v):
    n = T.info()["max_id"]
    out = np.zeros((n, 3))
    T.L.o_box_summaries(T.h, C.byref(v), _d(out)[1])
    return out


class OStreamTree(OTree):
    """OTree with the streamer calls under the names and argument forms of
    a5_streamer_amd.lib.Tree, so one host driver can step either."""

    def table_create(self, x_min, x_max, n_rows, n_cols):
        self.tbl = OTable(x_min, x_max, n_rows, n_cols)

    def table_set_col(self, col, xx, yy):
        self.tbl.set_col(col, xx, yy)

    def table_get(self):
        return self.tbl.data()

    def streamer_setup(self, dielectric_rule=True, pho=-1):
        self.v = streamer_vars(dielectric_rule, photoi=pho > 0)

    def space_charge(self, fac):
        space_charge(self, self.v, fac)

    def fld_from_pot(self):
        fld_from_pot(self, self.v)

    def fluxes_koren(self, dt):
        fluxes_koren(self, self.v, self.tbl, dt)

    def update_solution(self, dt):
        update_solution(self, self.v, self.tbl, dt)

    def average_dens(self):
        average_dens(self, self.v)

    def photoi_helmholtz(self, coeff, amp, kappa, i_phi, i_rhs, i_tmp, tol=1e-8,
                         max_cycles=30):
        a, pa = _d(amp)
        k, pk = _d(kappa)
        v = OVars(*[getattr(self.v, f) for f, _ in OVars._fields_])
        v.photoi = 1
        cyc = C.c_int()
        st = self.L.o_photoi_helmholtz(self.h, C.byref(v), self.tbl.h, C.c_double(coeff),
                                       C.c_int(len(a)), pa, pk, C.c_int(i_phi), C.c_int(i_rhs),
                                       C.c_int(i_tmp), C.c_double(tol), C.c_int(max_cycles),
                                       C.byref(cyc))
        if st != 0:
            raise RuntimeError("photoionization: Helmholtz solve did not converge")
        return cyc.value

    def photoi_helmholtz_terms(self, coeff, amp, kappa, i_terms, i_rhs, i_tmp, warm,
                               tol=1e-8, max_cycles=30):
        a, pa = _d(amp)
        k, pk = _d(kappa)
        iv = [np.ascontiguousarray(x, dtype=np.int32) for x in (i_terms, i_rhs, i_tmp)]
        if not all(len(x) == len(a) for x in [k] + iv):
            raise ValueError("photoi_helmholtz_terms: amp, kappa and the variables differ")
        v = OVars(*[getattr(self.v, f) for f, _ in OVars._fields_])
        v.photoi = 1
        cyc = C.c_int()
        pi = [x.ctypes.data_as(C.POINTER(C.c_int)) for x in iv]
        st = self.L.o_photoi_helmholtz_terms(
            self.h, C.byref(v), self.tbl.h, C.c_double(coeff), C.c_int(len(a)), pa, pk,
            pi[0], pi[1], pi[2], C.c_int(int(bool(warm))), C.c_double(tol),
            C.c_int(max_cycles), C.byref(cyc))
        if st != 0:
            raise RuntimeError("photoionization: Helmholtz solve did not converge")
        return cyc.value

    def box_summaries(self):
        return box_summaries(self, self.v)

    # --- the interface of a5_streamer_amd.adjoint.FrozenWindow -------------
    def set_cc_all(self, iv, data):
        """Tree.set_cc_all: row id-1 holds box id's block at its start."""
        b = self.boxes()
        for i in range(len(b["lvl"])):
            if b["in_use"][i]:
                n = self.box_len(int(b["n_cell"][i]))
                self.set_cc(i + 1, iv, np.asarray(data[i][:n], float))

    def get_cc_all(self, iv):
        b = self.boxes()
        out = np.zeros((len(b["lvl"]), self.box_len(self.n_cell)))
        for i in range(len(b["lvl"])):
            if b["in_use"][i]:
                nc = int(b["n_cell"][i])
                out[i, :self.box_len(nc)] = self.get_cc(i + 1, iv, nc)
        return out

    def set_table(self, rows_cols, x_min, inv_dx):
        self.tbl.set_rows(rows_cols)

    def table_adjoint(self, clear=False):
        """The accumulated table-row adjoints as (n_cols, n_rows)."""
        if getattr(self, "tgrad", None) is None:
            self.tgrad = np.zeros(self.tbl.n_rows * self.tbl.n_cols)
        out = self.tgrad.reshape(self.tbl.n_cols, self.tbl.n_rows).copy()
        if clear:
            self.tgrad[:] = 0
        return out

    def _adj(self, adj):
        a = OAdj(*[int(x) for x in adj])
        if getattr(self, "tgrad", None) is None:
            self.tgrad = np.zeros(self.tbl.n_rows * self.tbl.n_cols)
        return a

    def update_solution_adjoint(self, dt, adj):
        update_solution_adj(self, self.v, self.tbl, dt, self._adj(adj), self.tgrad)

    def consistent_fluxes_adjoint(self, lF):
        consistent_fluxes_adj(self, lF)

    def fluxes_koren_adjoint(self, dt, adj):
        fluxes_koren_adj(self, self.v, self.tbl, dt, self._adj(adj), self.tgrad)

    def tree_copy_cc_adjoint(self, lfrom, lto):
        self.L.o_tree_copy_cc_adj(self.h, C.c_int(lfrom), C.c_int(lto))

    def average_dens_adjoint(self, le, le_old, li, li_old):
        self.L.o_average_dens_adj(self.h, C.c_int(le), C.c_int(le_old), C.c_int(li),
                                  C.c_int(li_old))

    def space_charge_adjoint(self, fac, lrhs, le, li):
        self.L.o_space_charge_adj(self.h, fac, lrhs, le, li)

    def restrict_tree_adjoint(self, il):
        self.L.o_restrict_tree_adj(self.h, C.c_int(il))

    def gc_tree_adjoint(self, iv, il, rb, bc):
        if rb not in (1, 2) or bc not in (1, 2) or iv == il:
            raise ValueError("gc_tree_adjoint: rb, bc in (1, 2) and il != iv")
        self.L.o_gc_tree_adj(self.h, C.c_int(iv), C.c_int(il), C.c_int(rb), C.c_int(bc))

    def fld_from_pot_adjoint(self, lE, lfld, lphi):
        self.L.o_fld_from_pot_adj(self.h, C.byref(self.v), C.c_int(lE), C.c_int(lfld),
                                  C.c_int(lphi))

    def fas_vcycle_adjoint(self, max_lvl, adj):
        a, pa = _i(np.asarray(adj))
        vbar = np.zeros(6)
        self.L.o_fas_vcycle_adj(self.h, C.c_int(max_lvl), pa, _d(vbar)[1])
        return vbar[:2 * self.dim].copy()

    def fas_fmg_adjoint(self, adj):
        a, pa = _i(np.asarray(adj))
        vbar = np.zeros(6)
        if self.L.o_fas_fmg_adj(self.h, pa, _d(vbar)[1]) != 0:
            raise ValueError("fas_fmg_adjoint: no permittivity or Helmholtz term")
        return vbar[:2 * self.dim].copy()

    def state_store(self, slot, ivs, fvs=(), load=False, clear=False):
        """Tree.state_store: whole variables kept in host memory per slot."""
        ivs, fvs = [int(x) for x in ivs], [int(x) for x in fvs]
        a, pa = _i(np.asarray(ivs or [0]))
        f, pf = _i(np.asarray(fvs or [0]))
        n = int(self.L.o_store_size(self.h, len(ivs), len(fvs)))
        if clear:
            buf = np.zeros(n)
            self.L.o_store_vars(self.h, len(ivs), pa, len(fvs), pf, _d(buf)[1], 1)
            return
        if not hasattr(self, "_store"):
            self._store = {}
        key = (tuple(ivs), tuple(fvs), self.info()["max_id"])
        if load:
            if slot not in self._store or self._store[slot][0] != key:
                raise ValueError("state_store: slot %d holds no such state" % slot)
            buf = self._store[slot][1]
            self.L.o_store_vars(self.h, len(ivs), pa, len(fvs), pf, _d(buf)[1], 1)
        else:
            buf = np.zeros(n)
            self.L.o_store_vars(self.h, len(ivs), pa, len(fvs), pf, _d(buf)[1], 0)
            self._store[slot] = (key, buf)

    def prolong1_to(self, ids, iv, add=False):
        for i in np.atleast_1d(ids):
            OTree.prolong1_to(self, int(i), iv, add)

    def fill_cc(self, ids, iv, value):
        nc = self.boxes()["n_cell"]
        for i in np.atleast_1d(ids):
            self.set_cc(int(i), iv, np.full(self.box_len(int(nc[int(i) - 1])), value))

    def gc_boxes(self, ids, iv, rb, bc):
        for i in np.atleast_1d(ids):
            self.gc_box(int(i), iv, rb, bc)
