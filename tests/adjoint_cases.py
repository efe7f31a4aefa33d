"""Seeded cases of the reverse-mode transport substep (fluxes_koren,
a$D_consistent_fluxes, update_solution on a frozen mesh), shared by the
oracle's check against finite differences and the GPU's checks.

The substep maps the inputs x = (n_e, n_i on every cell of every box, ghost
cells included; |E| on every cell; the face field; the photoionization
source; the transport table) to n_e', n_i' on the leaves. The loss is
L = sum over leaf cells of w_e n_e' + w_i n_i' with seeded weights, so that
the adjoints of n_e', n_i' are (w_e, w_i)."""
import numpy as np

import scenarios as S

EPS0, QE = 8.8541878176e-12, 1.6022e-19
N_CC, N_FC = 13, 4
LE, LI, LE_FLD, LPHO = 10, 11, 12, 13   # adjoint cell variables
LF, LFLD = 3, 4                          # adjoint face variables
CELL_IN = (1, 2, 6, 8, 9)                # n_e, n_i, |E|, photoionization, eps (not perturbed)


def rules(lvl, c):
    if lvl < 2:
        return 1
    return 1 if (lvl < 4 and np.linalg.norm(c - 0.4) < 0.22) else -1


def table_data():
    """Transport data shaped like the air model (mobility 0.03, diffusion 0.2,
    alpha = 4.3e5 exp(-2e7/E), eta 3e3 exp(-1e7/E) so that it is not zero),
    on 0..3.2e7 V/m."""
    x = np.linspace(0, 3.2e7, 433)
    alpha = 4.3e5 * np.exp(-2e7 / np.maximum(x, 1.0))
    eta = 3e3 * np.exp(-1e7 / np.maximum(x, 1.0))
    return x, [0.03 * (1 + 0.1 * np.sin(x / 4e6)), np.full_like(x, 0.2), alpha, eta]


CASES = {
    # dim, coord, n_base, dielectric (rule and eps 1.5 cells), photoionization
    "2d": (2, 0, (2, 1, 1), True, False),
    "2d_photoi": (2, 0, (2, 1, 1), True, True),
    "cyl": (2, 1, (1, 2, 1), False, False),
    "3d": (3, 0, (1, 1, 1), True, True),
}


def build(M, name, seed):
    """The tree of case `name` in the tree class M, filled with seeded data."""
    dim, coord, base, diel, photoi = CASES[name]
    nc, L = 8, 1e-3
    T = M(dim, nc, N_CC, N_FC, L / nc, coarsen_to=2, coord=coord)
    S.build_tree(T, dim, nc, lambda l, c: rules(l, c / L), n_base=base)
    rng = np.random.default_rng(seed)
    b = T.boxes()
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        n = (b["n_cell"][i] + 2) ** dim
        vals = {1: 1e18 * (1 + rng.random(n)), 2: 1e18 * (1 + rng.random(n)),
                6: 1e7 * (0.3 + 1.2 * rng.random(n)), 8: 1e25 * rng.random(n),
                9: np.where(rng.random(n) < 0.1, 1.5, 1.0)}
        for iv, v in vals.items():
            T.set_cc(i + 1, iv, v)
        if b["lvl"][i] >= 1:
            nf = (b["n_cell"][i] + 1) * b["n_cell"][i] ** (dim - 1)
            for d in range(1, dim + 1):
                T.set_fc(i + 1, d, 2, rng.standard_normal(nf) * 1e7)
    return T, diel, photoi


def leaf_cells(T):
    """(id, flat interior indices) of every leaf box."""
    info = T.info()
    b = T.boxes()
    out = []
    for lvl in range(1, info["max_lvl"] + 1):
        for i in T.level(lvl)[1]:
            nc = int(b["n_cell"][int(i) - 1])
            m = np.zeros([nc + 2] * T.dim, bool)
            m[(slice(1, -1),) * T.dim] = True
            out.append((int(i), np.nonzero(m.ravel(order="F"))[0]))
    return out


def weights(leaves, seed):
    rng = np.random.default_rng(seed)
    return {(i, iv): (ix, rng.standard_normal(len(ix))) for i, ix in leaves for iv in (1, 2)}


def ids_in_use(T):
    b = T.boxes()
    return [(i + 1, int(b["n_cell"][i]), int(b["lvl"][i]))
            for i in range(len(b["lvl"])) if b["in_use"][i]]


def get_state(T, table_rows):
    """The inputs x as a dict of arrays."""
    x = {}
    for i, nc, lvl in ids_in_use(T):
        for iv in CELL_IN:
            x[("c", i, iv)] = T.get_cc(i, iv, nc)
        if lvl >= 1:
            for d in range(1, T.dim + 1):
                x[("f", i, d)] = T.get_fc(i, d, 2)
    x[("t",)] = np.array(table_rows, float)
    return x


def set_state(T, x, set_table):
    for k, v in x.items():
        if k[0] == "c":
            T.set_cc(k[1], k[2], v)
        elif k[0] == "f":
            T.set_fc(k[1], k[2], 2, v)
        else:
            set_table(v)


def directions(x, seed, n):
    """n seeded relative directions: every input scaled by a normal deviate
    (the permittivity and the table's zero entries stay put)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dx = {}
        for k, v in x.items():
            if k[0] == "c" and k[2] == 9:
                dx[k] = np.zeros_like(v)
            else:
                dx[k] = v * rng.standard_normal(v.shape)
        out.append(dx)
    return out


def loss(T, w):
    s = 0.0
    nc = {i: n for i, n, _ in ids_in_use(T)}
    for (i, iv), (ix, wv) in w.items():
        s += float(np.dot(wv, T.get_cc(i, iv, nc[i])[ix]))
    return s


def seed_adjoint(T, w):
    """The adjoints of n_e', n_i' (the loss weights) into LE, LI; every other
    adjoint variable zero."""
    nc = {}
    for i, n, lvl in ids_in_use(T):
        nc[i] = n
        cells = (n + 2) ** T.dim
        for iv in (LE, LI, LE_FLD, LPHO):
            T.set_cc(i, iv, np.zeros(cells))
        if lvl >= 1:
            nf = (n + 1) * n ** (T.dim - 1)
            for d in range(1, T.dim + 1):
                for f in (LF, LFLD):
                    T.set_fc(i, d, f, np.zeros(nf))
    for (i, iv), (ix, wv) in w.items():
        a = np.zeros((nc[i] + 2) ** T.dim)
        a[ix] = wv
        T.set_cc(i, LE if iv == 1 else LI, a)


def gradient(T, table_grad):
    """The adjoints of the inputs, in the layout of get_state."""
    g = {}
    for i, nc, lvl in ids_in_use(T):
        g[("c", i, 1)] = T.get_cc(i, LE, nc)
        g[("c", i, 2)] = T.get_cc(i, LI, nc)
        g[("c", i, 6)] = T.get_cc(i, LE_FLD, nc)
        g[("c", i, 8)] = T.get_cc(i, LPHO, nc)
        g[("c", i, 9)] = np.zeros((nc + 2) ** T.dim)
        if lvl >= 1:
            for d in range(1, T.dim + 1):
                g[("f", i, d)] = T.get_fc(i, d, LFLD)
    g[("t",)] = np.asarray(table_grad, float)
    return g


def dot(a, b):
    return sum(float(np.dot(a[k].ravel(), b[k].ravel())) for k in a)
