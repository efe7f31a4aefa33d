"""Seeded tree and field scenarios, run identically on every backend that
offers the reference-shaped API (the reference harness, the C restatement
and the HIP library's Python binding)."""
import numpy as np


def box_centres(b, nc):
    return b["r_min"] + 0.5 * nc * b["dr"][:, None]


def build_tree(T, dim, nc, rules, n_base=(1, 1, 1), max_iter=12):
    """Set a base of prod(n_base) boxes and refine with `rules`, a function
    (lvl, centre) -> flag, until nothing is added. Neighbours on the domain
    faces are physical boundaries (-1), as in the reference drivers."""
    nb_shape = n_base[:dim]
    ix_list, nb_list = [], []
    idx = {}
    n = 0
    for k in range(nb_shape[2] if dim == 3 else 1):
        for j in range(nb_shape[1]):
            for i in range(nb_shape[0]):
                n += 1
                key = (i, j, k) if dim == 3 else (i, j)
                idx[key] = n
                ix_list.append([i + 1, j + 1, k + 1][:dim])
    for key, n in sorted(idx.items(), key=lambda kv: kv[1]):
        nbs = []
        for d in range(dim):
            for s in (-1, 1):
                q = list(key)
                q[d] += s
                nbs.append(idx.get(tuple(q), -1))
        nb_list.append(nbs)
    T.set_base(np.array(ix_list, np.int32), np.array(nb_list, np.int32))
    for _ in range(max_iter):
        b = T.boxes()
        c = box_centres(b, nc)
        flags = np.array([rules(int(b["lvl"][i]), c[i]) if b["in_use"][i] else 0
                          for i in range(len(b["lvl"]))], np.int32)
        a, r = T.adjust(flags)
        if a == 0 and r == 0:
            break
    return T


def fill_random(T, dim, ivs, seed, boxes=None):
    """Fill variables `ivs` of every box (ghosts included) with seeded
    values; returns the dict of arrays so other backends get the same."""
    rng = np.random.default_rng(seed)
    b = T.boxes() if boxes is None else boxes
    data = {}
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        nc = int(b["n_cell"][i])
        for iv in ivs:
            v = rng.standard_normal((nc + 2) ** dim)
            data[(i + 1, iv)] = v
            T.set_cc(i + 1, iv, v)
    return data


def gaussian_rhs(T, dim, iv, centres, widths, amps, boxes=None):
    b = T.boxes() if boxes is None else boxes
    for i in range(len(b["lvl"])):
        if not b["in_use"][i]:
            continue
        nc = int(b["n_cell"][i])
        dr = b["dr"][i]
        ax = [b["r_min"][i][d] + (np.arange(nc + 2) - 0.5) * dr for d in range(dim)]
        grids = np.meshgrid(*ax, indexing="ij")
        val = np.zeros(grids[0].shape)
        for c, w, a in zip(centres, widths, amps):
            r2 = sum((g - cc) ** 2 for g, cc in zip(grids, c))
            val += a * np.exp(-r2 / w ** 2)
        # Fortran order: i fastest
        T.set_cc(i + 1, iv, val.ravel(order="F"))


def all_cc(T, dim, iv):
    b = T.boxes()
    out = {}
    for i in range(len(b["lvl"])):
        if b["in_use"][i]:
            out[i + 1] = T.get_cc(i + 1, iv, int(b["n_cell"][i]))
    return out
