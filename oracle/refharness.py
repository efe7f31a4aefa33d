"""ctypes view of oracle/_ref/libafivo_ref.so: the reference's own Fortran
(Afivo m_a2_*/m_a3_*, fosito, m_streamer) driven through the harness in
oracle/ref_harness.f90.in and oracle/ref_streamer.f90.

Test infrastructure only: tests use it to pin the C restatement and the HIP
path against the reference on identical inputs. The harness holds one tree
per dimension in module state, so a RefTree object is a view of that state.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libafivo_ref.so")

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref/libafivo_ref.so is not built "
                               "(make -C oracle ref)")
        _lib = C.CDLL(LIB_PATH)
        for name in ("ref_koren_mlim", "ref_lt_get_col", "ref2_min_dr",
                     "ref3_min_dr", "ref_norm2"):
            getattr(_lib, name).restype = C.c_double
        _lib.ref_morton_from_ix2.restype = C.c_int64
        _lib.ref_morton_from_ix3.restype = C.c_int64
    return _lib


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


class RefTree:
    """The reference tree of dimension `dim` (2 or 3)."""

    def __init__(self, dim, n_cell, n_cc, n_fc, dr, r_min=None, coarsen_to=-1,
                 lvls_max=30, coord=0, n_boxes=100):
        self.dim = dim
        self.p = "ref%d_" % dim
        self.L = lib()
        r_min = np.zeros(dim) if r_min is None else np.asarray(r_min, float)
        rm, prm = _d(r_min)
        self.f("init")(C.c_int(n_cell), C.c_int(n_cc), C.c_int(n_fc),
                       C.c_double(dr), prm, C.c_int(coarsen_to),
                       C.c_int(lvls_max), C.c_int(coord), C.c_int(n_boxes))
        self.n_cell = n_cell

    @classmethod
    def view(cls, dim, n_cell):
        """A view of the harness tree of dimension dim as it is (no a$D_init)."""
        t = cls.__new__(cls)
        t.dim, t.p, t.L, t.n_cell = dim, "ref%d_" % dim, lib(), n_cell
        return t

    def f(self, name):
        return getattr(self.L, self.p + name)

    # --- tree -----------------------------------------------------------
    def set_base(self, ix_list, nb_list):
        # (n, D) row-major has the memory layout of Fortran's ix_list(D, n)
        ix, pix = _i(np.asarray(ix_list))
        nb, pnb = _i(np.asarray(nb_list))
        self.f("set_base")(C.c_int(ix.shape[0]), pix, pnb)

    def adjust(self, flags_by_id):
        fl, pfl = _i(flags_by_id)
        self.f("set_flags")(C.c_int(len(fl)), pfl)
        a, r = C.c_int(), C.c_int()
        self.f("adjust")(C.byref(a), C.byref(r))
        return a.value, r.value

    def ref_info_lvl(self, lvl):
        a, r = C.c_int(), C.c_int()
        self.f("ref_info_lvl")(C.c_int(lvl), C.byref(a), C.byref(r))
        add = np.zeros(max(a.value, 1), np.int32)
        rm = np.zeros(max(r.value, 1), np.int32)
        if a.value or r.value:
            self.f("ref_info_ids")(C.c_int(lvl), _i(add)[1], _i(rm)[1])
        return add[:a.value], rm[:r.value]

    def tidy_up(self, max_frac, goal_frac, n_clean_min, reorder):
        self.f("tidy_up")(C.c_double(max_frac), C.c_double(goal_frac),
                          C.c_int(n_clean_min), C.c_int(int(reorder)))

    def info(self):
        out = np.zeros(8, np.int32)
        self.f("info")(_i(out)[1])
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
        ptrs = []
        for k in ("lvl", "ix", "parent", "children", "neighbors", "in_use",
                  "n_cell"):
            ptrs.append(out[k].ctypes.data_as(C.POINTER(C.c_int)))
        for k in ("dr", "r_min"):
            ptrs.append(out[k].ctypes.data_as(C.POINTER(C.c_double)))
        self.f("get_boxes")(C.c_int(n), *ptrs)
        return out

    def level(self, lvl):
        cnt = np.zeros(3, np.int32)
        self.f("level_counts")(C.c_int(lvl), _i(cnt)[1])
        ids = np.zeros(max(cnt[0], 1), np.int32)
        lv = np.zeros(max(cnt[1], 1), np.int32)
        pa = np.zeros(max(cnt[2], 1), np.int32)
        if cnt[0]:
            self.f("level_lists")(C.c_int(lvl), _i(ids)[1], _i(lv)[1],
                                  _i(pa)[1])
        return ids[:cnt[0]], lv[:cnt[1]], pa[:cnt[2]]

    # --- data -----------------------------------------------------------
    def box_len(self, nc):
        return (nc + 2) ** self.dim

    def set_cc(self, id_, iv, data):
        d, pd = _d(data)
        self.f("set_cc")(C.c_int(id_), C.c_int(iv), pd)

    def get_cc(self, id_, iv, nc=None):
        nc = self.n_cell if nc is None else nc
        out = np.zeros(self.box_len(nc))
        self.f("get_cc")(C.c_int(id_), C.c_int(iv), _d(out)[1])
        return out

    def set_fc(self, id_, dim, ifc, data):
        d, pd = _d(data)
        self.f("set_fc")(C.c_int(id_), C.c_int(dim), C.c_int(ifc), pd)

    def get_fc(self, id_, dim, ifc):
        nc = self.n_cell
        out = np.zeros((nc + 1) * nc ** (self.dim - 1))
        self.f("get_fc")(C.c_int(id_), C.c_int(dim), C.c_int(ifc), _d(out)[1])
        return out

    # --- multigrid / grid ops --------------------------------------------
    def mg_setup(self, i_phi, i_tmp, i_rhs, i_eps=-1, n_down=2, n_up=2,
                 n_base=2, bc_kind=None, bc_val=None, auto_ops=False):
        nnb = 2 * self.dim
        k = np.zeros(nnb, np.int32) if bc_kind is None else bc_kind
        v = np.zeros(nnb) if bc_val is None else bc_val
        ka, pk = _i(k)
        va, pv = _d(v)
        self.f("mg_setup")(C.c_int(i_phi), C.c_int(i_tmp), C.c_int(i_rhs),
                           C.c_int(i_eps), C.c_int(n_down), C.c_int(n_up),
                           C.c_int(n_base), pk, pv, C.c_int(int(auto_ops)))

    def fas_fmg(self, set_residual, first_call):
        self.f("fas_fmg")(C.c_int(int(set_residual)), C.c_int(int(first_call)))

    def fas_vcycle(self, max_lvl, set_residual):
        self.f("fas_vcycle")(C.c_int(max_lvl), C.c_int(int(set_residual)))

    def box_gsrb(self, id_, cntr):
        self.f("box_gsrb")(C.c_int(id_), C.c_int(cntr))

    def box_op(self, id_, i_out):
        self.f("box_op")(C.c_int(id_), C.c_int(i_out))

    def box_rstr(self, id_, iv):
        self.f("box_rstr")(C.c_int(id_), C.c_int(iv))

    def box_corr(self, id_):
        self.f("box_corr")(C.c_int(id_))

    def gc_box(self, id_, iv, rb, bc):
        self.f("gc_box")(C.c_int(id_), C.c_int(iv), C.c_int(rb), C.c_int(bc))

    def gc_tree(self, iv, rb, bc):
        self.f("gc_tree")(C.c_int(iv), C.c_int(rb), C.c_int(bc))

    def gc2_box(self, id_, iv):
        nc = self.n_cell
        out = np.zeros(nc ** (self.dim - 1) * 2 * self.dim)
        self.f("gc2_box")(C.c_int(id_), C.c_int(iv), _d(out)[1])
        return out

    def restrict_tree(self, iv):
        self.f("restrict_tree")(C.c_int(iv))

    def consistent_fluxes(self, f_ixs):
        a, p = _i(f_ixs)
        self.f("consistent_fluxes")(C.c_int(len(a)), p)

    def prolong1_to(self, id_, iv, add=False):
        self.f("prolong1_to")(C.c_int(id_), C.c_int(iv), C.c_int(int(add)))

    def prolong0_to(self, id_, iv):
        self.f("prolong0_to")(C.c_int(id_), C.c_int(iv))

    def tree_copy_cc(self, a, b):
        self.f("tree_copy_cc")(C.c_int(a), C.c_int(b))

    def tree_minmax(self, iv):
        lo, hi = C.c_double(), C.c_double()
        self.f("tree_minmax_cc")(C.c_int(iv), C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    def get_loc(self, r, max_lvl=0):
        ix = np.zeros(self.dim, np.int32)
        idv = C.c_int()
        self.f("get_loc")(_d(np.asarray(r, float))[1], C.c_int(max_lvl), C.byref(idv),
                          _i(ix)[1])
        return idv.value, ix

    def reduction_loc(self, kind, var, d, thr=0.0, init_val=-1e99):
        ix = np.zeros(self.dim, np.int32)
        val, idv = C.c_double(), C.c_int()
        self.f("reduction_loc")(C.c_int(kind), C.c_int(var), C.c_int(d), C.c_double(thr),
                                C.c_double(init_val), C.byref(val), C.byref(idv), _i(ix)[1])
        return val.value, idv.value, ix

    def min_dr(self):
        return self.f("min_dr")()


def vtk_write(fname, coords, connects, offsets, types, names, variables, cycle, time):
    """A .vtu file written by the reference's m_vtk with the call sequence of
    a$D_write_vtk; coords (n_nodes, dim), variables (n_cells, n_vars)."""
    coords = np.ascontiguousarray(coords, np.float64)
    n_nodes, dim = coords.shape
    v = np.asfortranarray(variables, np.float64)
    n_cells, n_vars = v.shape
    nm = b"".join(n.encode().ljust(100) for n in names)
    lib().ref_vtk_write(str(fname).encode() + b"\0", C.c_int(n_nodes), C.c_int(n_cells),
                        C.c_int(dim), _d(coords.ravel())[1], _i(connects)[1], _i(offsets)[1],
                        _i(types)[1], C.c_int(n_vars), C.c_char_p(nm),
                        v.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(cycle),
                        C.c_double(time))


def set_threads(n):
    """OpenMP threads of the reference routines called afterwards."""
    lib().ref_set_threads(C.c_int(n))


def koren_mlim(a, b):
    return lib().ref_koren_mlim(C.c_double(a), C.c_double(b))


def norm2(x):
    a, p = _d(x)
    return lib().ref_norm2(C.c_int(len(a)), p)


def morton2(i, j):
    return lib().ref_morton_from_ix2(C.c_int(i), C.c_int(j))


def morton3(i, j, k):
    return lib().ref_morton_from_ix3(C.c_int(i), C.c_int(j), C.c_int(k))


def morton_rank(keys):
    k = np.ascontiguousarray(keys, dtype=np.int64)
    r = np.zeros(len(k), np.int32)
    lib().ref_morton_rank(C.c_int(len(k)), k.ctypes.data_as(C.POINTER(C.c_int64)),
                          _i(r)[1])
    return r


class RefTable:
    """The reference's uniform lookup table (fosito m_lookup_table)."""

    def __init__(self, x_min, x_max, n_rows, n_cols):
        self.n_rows, self.n_cols = n_rows, n_cols
        lib().ref_lt_create(C.c_double(x_min), C.c_double(x_max),
                            C.c_int(n_rows), C.c_int(n_cols))

    def set_col(self, col, xx, yy):
        xa, px = _d(xx)
        ya, py = _d(yy)
        lib().ref_lt_set_col(C.c_int(col), C.c_int(len(xa)), px, py)

    def get_col(self, col, x):
        return lib().ref_lt_get_col(C.c_int(col), C.c_double(x))

    def data(self):
        x = np.zeros(self.n_rows)
        cr = np.zeros((self.n_rows, self.n_cols))
        lib().ref_lt_get_data(C.c_int(self.n_rows), C.c_int(self.n_cols),
                              _d(x)[1], _d(cr)[1])
        return x, cr


def td_read(fname, gas, name, max_n=1000):
    xx = np.zeros(max_n)
    yy = np.zeros(max_n)
    n = C.c_int()
    lib().ref_td_read(fname.encode() + b"\0", gas.encode() + b"\0",
                      name.encode() + b"\0", C.c_int(max_n), C.byref(n),
                      _d(xx)[1], _d(yy)[1])
    return xx[:n.value], yy[:n.value]


# --- the reference's driver programs (oracle/ref_drivers.f90) ---------------

DRIVER_VARIANTS = {"2d": 1, "3d": 2, "cyl": 3}


class RefDriver:
    """One of the reference's driver programs (src/streamer_2d.f90,
    src/streamer_3d.f90, src/streamer_cyl.f90) built as a module: its box
    procedures act on the harness tree of its dimension (a RefTree), and
    run() executes its main program from init_tree to the end of its time
    loop on its own tree, which then becomes the harness tree."""

    def __init__(self, variant):
        self.variant = variant
        self.vid = DRIVER_VARIANTS[variant]
        self.dim = 3 if variant == "3d" else 2
        self.L = lib()
        self.L.refdrv_time.restype = C.c_double
        self.L.refdrv_max_dt.restype = C.c_double

    def config(self, domain_len, box_size, applied_fld, dt_max=1e-8, end_time=0.0,
               dt_out=1e99, epsilon_diel=1.0, gas_pressure=1.0, fld_mod_t0=1e99,
               fld_sin_amplitude=0.0, fld_sin_freq=0.2e9, fld_lin_deriv=0.0, time=0.0,
               steps_amr=2, photoi=False):
        d = np.array([domain_len, applied_fld, dt_max, end_time, dt_out, epsilon_diel,
                      gas_pressure, fld_mod_t0, fld_sin_amplitude, fld_sin_freq,
                      fld_lin_deriv, time], float)
        iv = np.array([box_size, steps_amr, int(bool(photoi))], np.int32)
        self.L.refdrv_config(_d(d)[1], _i(iv)[1])

    def init_cond(self, bg_dens, r0, r1, dens, width, falloff):
        r0 = np.asarray(r0, float).reshape(-1, self.dim)
        r1 = np.asarray(r1, float).reshape(-1, self.dim)
        n = r0.shape[0]
        # (n, dim) row-major has the memory layout of Fortran's r0(dim, n)
        self.L.refdrv_init_cond(C.c_int(self.dim), C.c_int(n), C.c_double(bg_dens),
                                _d(r0)[1], _d(r1)[1], _d(np.asarray(dens, float))[1],
                                _d(np.asarray(width, float))[1],
                                _i(np.asarray(falloff))[1])

    def table(self, n_rows, max_fld, x, cols):
        x = np.asarray(x, float)
        # cols(n, 4) column-major = the four columns one after the other
        c = np.ascontiguousarray(np.stack([np.asarray(y, float) for y in cols]))
        self.L.refdrv_table(C.c_int(n_rows), C.c_double(max_fld), C.c_int(len(x)),
                            _d(x)[1], _d(c)[1])

    def time(self):
        dt = C.c_double()
        t = self.L.refdrv_time(C.byref(dt))
        return t, dt.value

    def run(self, n_cell):
        self.L.refdrv_run(C.c_int(self.vid))
        return RefTree.view(self.dim, n_cell)

    def counters(self, reset=True):
        """(time steps, leaf cell-steps) the driver's main loop has run
        since the last reset."""
        n, c = C.c_int(), C.c_double()
        self.L.refdrv_counters(C.c_int(self.vid), C.byref(n), C.byref(c), C.c_int(int(reset)))
        return n.value, c.value

    def mg_init(self):
        self.L.refdrv_mg_init(C.c_int(self.vid))

    def compute_fld(self, n_cycles, no_guess):
        self.L.refdrv_compute_fld(C.c_int(self.vid), C.c_int(n_cycles),
                                  C.c_int(int(bool(no_guess))))

    def fld_from_pot(self):
        self.L.refdrv_fld_from_pot(C.c_int(self.vid))

    def fluxes_koren(self, dt):
        self.L.refdrv_fluxes_koren(C.c_int(self.vid), C.c_double(dt))

    def update_solution(self, dt):
        self.L.refdrv_update_solution(C.c_int(self.vid), C.c_double(dt))

    def average_dens(self):
        self.L.refdrv_average_dens(C.c_int(self.vid))

    def photoi_rate(self, coeff):
        self.L.refdrv_photoi_rate(C.c_int(self.vid), C.c_double(coeff))

    def set_init_cond(self):
        self.L.refdrv_set_init_cond(C.c_int(self.vid))

    def ref_flags(self, n):
        out = np.zeros(n, np.int32)
        self.L.refdrv_ref_flags(C.c_int(self.vid), C.c_int(n), _i(out)[1])
        return out

    def max_dt(self):
        return self.L.refdrv_max_dt(C.c_int(self.vid))
